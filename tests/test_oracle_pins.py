"""Pins of the oracle (CPU): the reference's own closed-form tests, its
published numbers, cross-implementation agreement, and invariants that a
dropped term, a wrong sign or a transposed index would break.

The oracle has two restatements (C: oracle/fiberoracle.c, numpy:
oracle/np_ref.py); both are checked here.
"""
from __future__ import annotations

import itertools
import json
import math
import pathlib

import numpy as np
import pytest

import oracle
from oracle import np_ref as R

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def _smf(nt=256, window=3.0, length=0.05, power=0.05, fR=0.18, preset="SMF28", beta_order=3):
    sim = R.sim_params(1550e-9, 1, nt, window, beta_order)
    p = R.SINGLE_MODE_PRESETS[preset]
    rf = R.raman_fields(sim, fR)
    D = R.dispersion_user(sim, p["betas"])
    S = R.shock(sim)
    _, u0 = R.initial_state([1.0], power, 185e-15, 80.5e6, "sech_sq", sim)
    g = np.full((1, 1, 1, 1), p["gamma"])
    return sim, oracle.Fiber(D, g, rf["hR"], S, rf["one_m_fR"]), u0, (D, g, rf["hR"], S, rf["one_m_fR"])


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# --------------------------------------------------------------- setup


@pytest.mark.parametrize("nt", [64, 128, 256, 512])
def test_raman_response_closed_form(nt):
    """test_numerical_kernels.jl:65-95: the analytic single-oscillator
    transform, written in the test's amplitude*resonance form, rtol 5e-14;
    DC = fR, Nyquist real, Hermitian."""
    sim = R.sim_params(1550e-9, 1, nt, 3.0, 2)
    hR = R.raman_fields(sim, 0.18, 12.2, 32.0)["hR"]
    tau1, tau2 = 12.2e-3, 32.0e-3
    omega = 2 * np.pi * R.fftfreq(nt, 1 / sim["dt"])
    amp = (tau1 ** 2 + tau2 ** 2) / (tau1 * tau2 ** 2)
    expected = 0.18 * amp * (1 / tau1) / ((1 / tau2 + 1j * omega) ** 2 + (1 / tau1) ** 2)
    expected = np.exp(1j * (-omega * sim["time_window"] / 2)) * expected
    expected[nt // 2] = expected[nt // 2].real
    assert np.allclose(hR, expected, rtol=5e-14, atol=5e-14)
    assert abs(hR[0].real - 0.18) <= 5e-14 * 0.18
    assert hR[nt // 2].imag == 0
    assert np.allclose(hR[1:nt // 2], np.conj(hR[nt // 2 + 1:][::-1]))


def test_grid_validator_rejects_10ps_window_at_8192():
    """test_numerical_kernels.jl:45-53: Nt=8192 at 12 ps is rejected."""
    with pytest.raises(ValueError, match="use at least 21.178 ps"):
        R.sim_params(1550e-9, 1, 8192, 12.0, 2)


def test_fft_conventions():
    """test_numerical_kernels.jl:259-275 and the C FFT against numpy."""
    nt = 64
    temporal = np.exp(2j * np.pi * np.arange(nt) / nt)
    spectrum = oracle.fft(temporal, inverse=True)
    assert np.allclose(oracle.fft(spectrum), temporal, rtol=1e-14, atol=1e-14)
    assert math.isclose(nt * np.sum(np.abs(spectrum) ** 2), np.sum(np.abs(temporal) ** 2), rel_tol=1e-14)
    rng = np.random.default_rng(1)
    for n in (8, 256, 4096):
        x = rng.normal(size=n) + 1j * rng.normal(size=n)
        assert _rel(oracle.fft(x), np.fft.fft(x)) < 1e-14
        assert _rel(oracle.fft(x, inverse=True), np.fft.ifft(x)) < 1e-14
    sim = R.sim_params(1550e-9, 1, nt, 10.0, 2)
    const = R.field_metrics(np.fft.ifft(np.ones((nt, 1)), axis=0), sim)
    imp = np.zeros((nt, 1), complex)
    imp[nt // 2] = 1
    impm = R.field_metrics(np.fft.ifft(imp, axis=0), sim)
    assert const["temporal_edge_fraction"] == 6 / nt and const["spectral_edge_fraction"] == 0
    assert impm["temporal_edge_fraction"] == 0 and impm["spectral_edge_fraction"] == 6 / nt


def test_raman_convolution_grid_independence():
    """test_numerical_kernels.jl:297-327: 1024 vs 2048 samples, < 1e-12."""
    def launch_and_raman(nt):
        sim = R.sim_params(1550e-9, 1, nt, 10.0, 2)
        hR = R.raman_fields(sim, 0.18)["hR"]
        _, uw = R.initial_state([1.0], 0.05, 185e-15, 80.5e6, "sech_sq", sim)
        inten = np.abs(np.fft.fft(uw, axis=0)[:, 0]) ** 2
        raman = np.fft.fftshift(np.fft.ifft(hR * np.fft.fft(inten.astype(complex)))).real
        return inten, raman

    i1, r1 = launch_and_raman(1024)
    i2, r2 = launch_and_raman(2048)
    assert np.max(np.abs(i1 - i2[::2])) / np.max(i2) < 1e-13
    assert np.max(np.abs(r1 - r2[::2])) / np.max(np.abs(r2)) < 1e-12


def test_taylor_basis_values():
    """test_physical_controls.jl:232-279."""
    coarse = R.sim_params(1550e-9, 1, 16, 4.0, 2, validate=False)
    fine = R.sim_params(1550e-9, 1, 32, 8.0, 2, validate=False)
    cb = R.taylor_phase_basis(coarse, (2, 3, 4))
    fb = R.taylor_phase_basis(fine, (2, 3, 4))
    cf = R.fftfreq(16, 1 / coarse["dt"])
    ff = R.fftfreq(32, 1 / fine["dt"])
    for i, f in enumerate(cf):
        j = int(np.nonzero(ff == f)[0][0])
        assert np.allclose(cb[i], fb[j])
    expected = (125.0 * 1e-6 / 2) * (2 * np.pi * cf) ** 2
    assert np.allclose(R.taylor_phase_basis(coarse, (2,))[:, 0] * 125.0, expected)
    assert np.allclose(R.taylor_phase_basis(coarse, (2, 3, 4), (125.0, 2000.0, 5e4)), cb * [125.0, 2000.0, 5e4])
    assert np.all(cb[0] == 0)
    assert math.isclose(cb[1, 1], -cb[-1, 1])


# ------------------------------------------------------- propagation


def test_c_and_numpy_restatements_agree():
    """Two independent restatements (own radix-2 FFT vs pocketfft)."""
    sim, fib, u0, args = _smf(nt=256, power=0.2)
    z = 0.013
    assert _rel(oracle.rhs(fib, z, u0), R.rhs(z, u0, *args)) < 1e-14
    assert _rel(oracle.rk4_forward(fib, 0.05, 30, u0), R.rk4(u0, 0.05, 30, *args)) < 1e-13


def test_linear_dispersion_closed_form():
    """gamma = 0: RK4 output equals cis(D L) u0 to 2e-15 (test_scalar_crosscheck.jl:27-37)."""
    sim = R.sim_params(1550e-9, 1, 128, 4.0, 3)
    rf = R.raman_fields(sim, 0.18)
    D = R.dispersion_user(sim, [-2.17e-26, 1.2e-40])
    _, u0 = R.initial_state([1.0], 0.02, 185e-15, 80.5e6, "sech_sq", sim)
    fib = oracle.Fiber(D, np.zeros((1, 1, 1, 1)), rf["hR"], R.shock(sim), rf["one_m_fR"])
    out = oracle.rk4_forward(fib, 0.02, 4, u0)
    expected = np.exp(1j * D * 0.02) * u0
    assert np.max(np.abs(out - expected)) <= 2e-15 + 2e-15 * np.max(np.abs(expected))


def test_cw_kerr_rk4_order():
    """test_scalar_crosscheck.jl:39-56: gamma=1, P=5, L=2, 32/64/128 steps."""
    sim = R.sim_params(1550e-9, 1, 64, 2.0, 2)
    rf = R.raman_fields(sim, 0.0)
    launch = np.fft.ifft(np.full((64, 1), np.sqrt(5.0), dtype=complex), axis=0)
    fib = oracle.Fiber(np.zeros((64, 1)), np.ones((1, 1, 1, 1)), rf["hR"], R.shock(sim), 1.0)
    exact = launch * np.exp(1j * 10.0)
    errs = [np.linalg.norm(oracle.rk4_forward(fib, 2.0, s, launch) - exact) for s in (32, 64, 128)]
    assert np.log2(errs[0] / errs[1]) > 3.5 and np.log2(errs[1] / errs[2]) > 3.5
    assert errs[-1] / np.linalg.norm(exact) < 1e-4


def test_cw_kerr_tight():
    """test_numerical_kernels.jl:329-353: gamma=0.73, P=2, L=5, no dispersion,
    no Raman, shock off; tight error < 1e-8."""
    nt = 64
    sim = R.sim_params(1550e-9, 1, nt, 10.0, 2)
    launch = np.fft.ifft(np.full((nt, 1), np.sqrt(2.0), dtype=complex), axis=0)
    fib = oracle.Fiber(np.zeros((nt, 1)), np.full((1, 1, 1, 1), 0.73), np.zeros(nt), np.ones(nt), 1.0)
    out = oracle.rk4_forward(fib, 5.0, 2000, launch)
    expected = launch * np.exp(1j * 0.73 * 2.0 * 5.0)
    assert _rel(out, expected) < 1e-8


def test_stokes_sign_and_photon_number():
    """test_numerical_kernels.jl:204-221: HNLF_zero_disp, 0.02 m, 0.05 W,
    Nt=256, 3 ps: centroid shift < -1e-3 THz, photon drift < 1e-6."""
    sim, fib, u0, _ = _smf(nt=256, window=3.0, length=0.02, power=0.05, preset="HNLF_zero_disp")
    out = oracle.rk4_forward(fib, 0.02, 400, u0)
    shift = R.centroid_thz(out, sim) - R.centroid_thz(u0, sim)
    assert shift < -1e-3
    p0 = R.field_metrics(u0, sim)["photon_number"]
    p1 = R.field_metrics(out, sim)["photon_number"]
    assert abs(p1 / p0 - 1) < 1e-6


def test_small_problem_photon_number():
    """test_propagation.jl:19-73: SMF28_beta2_only, 1e-4 m, 1e-5 W, Gaussian
    250 fs at 50 MHz, Nt=128, 3 ps: photon drift < 1e-8; input energy equals
    P/rep within 1e-5."""
    sim = R.sim_params(1550e-9, 1, 128, 3.0, 2)
    rf = R.raman_fields(sim, 0.18)
    D = R.dispersion_user(sim, [-2.17e-26])
    _, u0 = R.initial_state([1.0], 1e-5, 250e-15, 50e6, "gauss", sim)
    fib = oracle.Fiber(D, np.full((1, 1, 1, 1), 1.1e-3), rf["hR"], R.shock(sim), rf["one_m_fR"])
    out = oracle.rk4_forward(fib, 1e-4, 20, u0)
    m0, m1 = R.field_metrics(u0, sim), R.field_metrics(out, sim)
    assert math.isclose(m0["energy_pj"], 1e-5 / 50e6 * 1e12, rel_tol=1e-5)
    assert abs(m1["photon_number"] / m0["photon_number"] - 1) < 1e-8


def test_published_counterfactual_centroid_shift():
    """Published Delta C (counterfactual-raman-benchmark.md:75-79; problem
    construction examples/05_counterfactual_raman.jl:114-160): SMF28 0.5 m,
    0.1 W, 185 fs sech2, Nt=1024, 10 ps; 2000 RK4 steps.  Tolerance 2.5e-6
    THz, inside the reference's 5.95e-6 numerical envelope (:92-95)."""
    pins = json.loads((GOLDEN / "published_pins.json").read_text())["counterfactual_shift_thz"]
    sim = R.sim_params(1550e-9, 1, 1024, 10.0, 3)
    D = R.dispersion_user(sim, [-2.17e-26, 1.2e-40])
    _, u0 = R.initial_state([1.0], 0.1, 185e-15, 80.5e6, "sech_sq", sim)
    basis = R.taylor_phase_basis(sim, (2, 3))
    fibs = {fr: oracle.Fiber(D, np.full((1, 1, 1, 1), 1.1e-3), R.raman_fields(sim, fr)["hR"], R.shock(sim),
                             1 - fr) for fr in (0.18, 0.0)}
    for pin in pins:
        launch = u0 * np.exp(1j * (basis @ np.array([pin["phi2_fs2"], pin["phi3_fs3"]])))[:, None]
        c = [R.centroid_thz(oracle.rk4_forward(fibs[fr], 0.5, 2000, launch), sim) for fr in (0.18, 0.0)]
        assert abs((c[0] - c[1]) - pin["delta_c_thz"]) <= 2.5e-6, (pin, c[0] - c[1])


def test_mechanism_controls():
    """counterfactual-raman-benchmark.md:105-107: gamma = 0 gives Delta C ~ 0
    (|Delta C| < 1e-7 THz)."""
    sim = R.sim_params(1550e-9, 1, 256, 10.0, 3)
    D = R.dispersion_user(sim, [-2.17e-26, 1.2e-40])
    _, u0 = R.initial_state([1.0], 0.1, 185e-15, 80.5e6, "sech_sq", sim)
    outs = []
    for fr in (0.18, 0.0):
        rf = R.raman_fields(sim, fr)
        fib = oracle.Fiber(D, np.zeros((1, 1, 1, 1)), rf["hR"], R.shock(sim), rf["one_m_fR"])
        outs.append(R.centroid_thz(oracle.rk4_forward(fib, 0.5, 50, u0), sim))
    assert abs(outs[0] - outs[1]) < 1e-7


# ---------------------------------------------------------- multimode


def _sym_gamma(m, rng, scale=1e-2):
    g = rng.uniform(0.5, 1.5, size=(m,) * 4) * scale
    s = np.zeros_like(g)
    for p in itertools.permutations(range(4)):
        s += np.transpose(g, p)
    return s / 24


def _mm_setup(m, nt=128, fR=0.18, shock=True, seed=0):
    rng = np.random.default_rng(seed)
    sim = R.sim_params(1550e-9, m, nt, 3.0, 3)
    rf = R.raman_fields(sim, fR)
    D = np.hstack([R.dispersion_user(sim, [-2.17e-26 * (1 + 0.1 * k), 1.2e-40]) + 2.0 * k for k in range(m)])
    modes = rng.normal(size=m) + 1j * rng.normal(size=m)
    _, u0 = R.initial_state(modes / np.linalg.norm(modes), 0.05, 185e-15, 80.5e6, "sech_sq", sim)
    S = R.shock(sim) if shock else np.ones(nt)
    return sim, D, rf, S, u0, rng


def test_multimode_reduces_to_scalar_with_diagonal_gamma():
    """Diagonal-only gamma decouples into independent scalar NLSEs."""
    sim, D, rf, S, u0, rng = _mm_setup(3)
    g = np.zeros((3,) * 4)
    gd = [1e-2, 2e-2, 3e-2]
    for k in range(3):
        g[k, k, k, k] = gd[k]
    out = oracle.rk4_forward(oracle.Fiber(D, g, rf["hR"], S, rf["one_m_fR"]), 0.05, 20, u0)
    for k in range(3):
        fk = oracle.Fiber(D[:, [k]], np.full((1, 1, 1, 1), gd[k]), rf["hR"], S, rf["one_m_fR"])
        assert _rel(out[:, [k]], oracle.rk4_forward(fk, 0.05, 20, u0[:, [k]])) < 1e-13


def test_multimode_energy_and_photon_number():
    """fR = 0 and shock off: total energy conserved; with Raman and shock the
    photon number is conserved (to the RK4 truncation)."""
    sim, D, rf, S, u0, rng = _mm_setup(3, fR=0.0, shock=False)
    g = _sym_gamma(3, rng, 5e-3)
    out = oracle.rk4_forward(oracle.Fiber(D, g, rf["hR"], S, 1.0), 0.05, 200, u0)
    assert abs(np.sum(np.abs(out) ** 2) / np.sum(np.abs(u0) ** 2) - 1) < 1e-9
    sim, D, rf, S, u0, rng = _mm_setup(3, fR=0.18, shock=True)
    out = oracle.rk4_forward(oracle.Fiber(D, _sym_gamma(3, rng, 5e-3), rf["hR"], S, rf["one_m_fR"]), 0.05, 200, u0)
    p0, p1 = R.field_metrics(u0, sim)["photon_number"], R.field_metrics(out, sim)["photon_number"]
    assert abs(p1 / p0 - 1) < 1e-8


def test_multimode_c_vs_numpy():
    sim, D, rf, S, u0, rng = _mm_setup(3)
    g = _sym_gamma(3, rng)
    fib = oracle.Fiber(D, g, rf["hR"], S, rf["one_m_fR"])
    assert _rel(oracle.rhs(fib, 0.02, u0), R.rhs(0.02, u0, D, g, rf["hR"], S, rf["one_m_fR"])) < 1e-14


# -------------------------------------------------------------- adjoint


def test_rhs_vjp_matches_finite_differences():
    """The pullback satisfies <gy, J dx> = <VJP(gy), dx> (real inner product)."""
    for m in (1, 2):
        sim, D, rf, S, u0, rng = _mm_setup(m, seed=m)
        g = _sym_gamma(m, rng, 2e-1)
        fib = oracle.Fiber(D, g, rf["hR"], S, rf["one_m_fR"])
        z, eps = 0.013, 1e-6
        gy = rng.normal(size=u0.shape) + 1j * rng.normal(size=u0.shape)
        dx = (rng.normal(size=u0.shape) + 1j * rng.normal(size=u0.shape)) * np.abs(u0).max()
        fd = np.sum(np.real(np.conj(gy) * (oracle.rhs(fib, z, u0 + eps * dx) - oracle.rhs(fib, z, u0 - eps * dx))
                            / (2 * eps)))
        ad = np.sum(np.real(np.conj(oracle.rhs_vjp(fib, z, u0, gy)) * dx))
        assert abs(fd - ad) <= 1e-8 * abs(fd)
        assert _rel(oracle.rhs_vjp(fib, z, u0, gy), R.rhs_vjp(z, u0, gy, D, g, rf["hR"], S, rf["one_m_fR"])) < 1e-13


def test_discrete_adjoint_phase_gradient_vs_central_fd():
    """Band objective J = E_band/E (physics_helpers.jl:222-232) with the
    chain rule g = 2 Re(conj(lam0) i u0) (physics_models.jl:1042)."""
    sim, fib, u0, _ = _smf(nt=256, window=3.0, power=0.05)
    fib = oracle.Fiber(fib.D, np.full((1, 1, 1, 1), 5.5e-3), fib.hR, fib.shock, fib.one_m_fR)
    L, steps = 0.3, 100
    mask = R.frequency_band_mask(sim, -16.0, -4.0)
    J, lamL = R.band_cost(oracle.rk4_forward(fib, L, steps, u0), mask)
    g = R.phase_gradient(oracle.rk4_adjoint(fib, L, steps, u0, lamL, every=16), u0)
    for k in (2, 20, 250):
        e = 1e-5
        ph = np.zeros(256)
        ph[k] = e
        jp = R.band_cost(oracle.rk4_forward(fib, L, steps, u0 * np.exp(1j * ph)[:, None]), mask)[0]
        jm = R.band_cost(oracle.rk4_forward(fib, L, steps, u0 * np.exp(-1j * ph)[:, None]), mask)[0]
        fd = (jp - jm) / (2 * e)
        assert abs(fd - g[k]) <= 1e-6 * abs(fd) + 1e-12


def test_discrete_adjoint_checkpointing_invariant():
    """Checkpoint spacing does not change the discrete adjoint; C equals numpy."""
    sim, fib, u0, args = _smf(nt=128, window=3.0, power=0.1)
    lam = np.random.default_rng(3).normal(size=u0.shape) * (1 + 0j)
    a = oracle.rk4_adjoint(fib, 0.05, 12, u0, lam, every=1)
    b = oracle.rk4_adjoint(fib, 0.05, 12, u0, lam, every=5)
    c = R.rk4_adjoint(u0, lam, 0.05, 12, *args)
    assert np.array_equal(a, b)
    assert _rel(a, c) < 1e-13


def test_discrete_adjoint_converges_to_reference_continuous_adjoint():
    """The reference's continuous adjoint (adjoint_disp_mmf!, restated) and the
    discrete adjoint of RK4 converge to each other at 4th order."""
    sim, fib, u0, _ = _smf(nt=128, window=3.0, power=0.05)
    fib = oracle.Fiber(fib.D, np.full((1, 1, 1, 1), 5.5e-3), fib.hR, fib.shock, fib.one_m_fR)
    mask = R.frequency_band_mask(sim, -16.0, -4.0)
    errs = []
    for steps in (100, 200):
        _, lamL = R.band_cost(oracle.rk4_forward(fib, 0.3, steps, u0), mask)
        d = oracle.rk4_adjoint(fib, 0.3, steps, u0, lamL, every=10)
        c = oracle.continuous_adjoint(fib, 0.3, steps, u0, lamL)
        errs.append(_rel(d, c))
    assert errs[1] < 1e-3
    assert np.log2(errs[0] / errs[1]) > 3.5


def test_adjoint_linear_in_seed():
    """test_numerical_kernels.jl:166-202: the adjoint is linear in the seed."""
    sim, fib, u0, _ = _smf(nt=128, window=3.0, power=0.1)
    lam = np.random.default_rng(4).normal(size=u0.shape) + 0j
    a = oracle.rk4_adjoint(fib, 0.05, 10, u0, lam, every=5)
    b = oracle.rk4_adjoint(fib, 0.05, 10, u0, -1e-4 * lam, every=5)
    assert np.allclose(b, -1e-4 * a, rtol=1e-12, atol=0)


# ----------------------------------------------------- adaptive Tsit5

def test_tsit5_tableau_order_conditions():
    """Tsitouras 5(4) pair: row sums equal c, the 5th-order weights satisfy
    the order conditions through order 5 (all 17 trees), the embedded pair
    through order 4 (8 trees).  A mistyped coefficient fails one of them."""
    c, a, bt = oracle.tsit5_tableau()
    A = np.zeros((7, 7))
    A[:, :6] = a
    b = A[6].copy()
    bhat = b - bt
    assert np.allclose(A.sum(axis=1), c, atol=1e-14)
    C = np.diag(c)
    e = np.ones(7)
    trees = {  # (elementary weight vector, 1/gamma)
        1: [(e, 1)],
        2: [(c, 2)],
        3: [(c ** 2, 3), (A @ c, 6)],
        4: [(c ** 3, 4), (C @ A @ c, 8), (A @ c ** 2, 12), (A @ A @ c, 24)],
        5: [(c ** 4, 5), (C @ C @ A @ c, 10), (C @ A @ c ** 2, 15), (C @ A @ A @ c, 30), ((A @ c) ** 2, 20),
            (A @ c ** 3, 20), (A @ (C @ A @ c), 40), (A @ A @ c ** 2, 60), (A @ A @ A @ c, 120)],
    }
    for order, ts in trees.items():
        for phi, g in ts:
            assert abs(b @ phi - 1 / g) < 1e-13, (order, g)
            if order <= 4:
                assert abs(bhat @ phi - 1 / g) < 1e-13, ("embedded", order, g)
    # the embedded pair is not 5th order (otherwise the error estimate vanishes)
    assert abs(bhat @ c ** 4 - 1 / 5) > 1e-4


def test_tsit5_fixed_step_order_five():
    """The accepted-step map converges at 5th order on the SMF-28 problem."""
    sim, fib, u0, _ = _smf(nt=256, window=3.0, length=0.5, power=0.1)
    ref = oracle.tsit5_replay(fib, 0.5, np.full(2560, 0.5 / 2560), u0)
    errs = [_rel(oracle.tsit5_replay(fib, 0.5, np.full(n, 0.5 / n), u0), ref) for n in (40, 80)]
    assert 4.7 < np.log2(errs[0] / errs[1]) < 5.5


def test_tsit5_cw_kerr_tight_and_loose():
    """test_numerical_kernels.jl:329-353 with the adaptive solver: tight
    (1e-10, 1e-12) error < 1e-8 and < 1e-3 x the loose (1e-4, 1e-6) error."""
    nt = 64
    launch = np.fft.ifft(np.full((nt, 1), np.sqrt(2.0), dtype=complex), axis=0)
    fib = oracle.Fiber(np.zeros((nt, 1)), np.full((1, 1, 1, 1), 0.73), np.zeros(nt), np.ones(nt), 1.0)
    expected = launch * np.exp(1j * 0.73 * 2.0 * 5.0)
    tight = _rel(oracle.tsit5_forward(fib, 5.0, launch, rtol=1e-10, atol=1e-12).out, expected)
    loose = _rel(oracle.tsit5_forward(fib, 5.0, launch, rtol=1e-4, atol=1e-6).out, expected)
    assert tight < 1e-8 and tight < 1e-3 * loose


def test_tsit5_dense_output_order_conditions():
    """The continuous extension b_i(th) satisfies the eight order conditions
    through order 4 at every th, vanishes at th = 0 and equals the 5th-order
    weights at th = 1.  The coefficient table is overdetermined by these
    (22 free coefficients, > 27 conditions), so a mistyped digit fails."""
    c, a, _ = oracle.tsit5_tableau()
    A = np.zeros((7, 7))
    A[:, :6] = a
    Ac = A @ c
    assert np.max(np.abs(oracle.tsit5_dense_weights(0.0))) == 0.0
    assert np.max(np.abs(oracle.tsit5_dense_weights(1.0) - A[6])) < 1e-14
    for th in (0.05, 0.21, 0.5, 0.77, 0.93, 1.0):
        w = oracle.tsit5_dense_weights(th)
        conds = [w.sum() - th, w @ c - th ** 2 / 2, w @ c ** 2 - th ** 3 / 3, w @ Ac - th ** 3 / 6,
                 w @ c ** 3 - th ** 4 / 4, w @ (c * Ac) - th ** 4 / 8, w @ (A @ c ** 2) - th ** 4 / 12,
                 w @ (A @ Ac) - th ** 4 / 24]
        assert np.max(np.abs(conds)) < 1e-13, (th, conds)


def test_tsit5_saveat_closed_forms():
    """propagate(problem; saveat) samples (propagation.jl:99-110) from the
    adaptive solver's continuous extension.  CW Kerr (test_numerical_kernels.jl:
    329-353 setup): positions inside accepted steps match u0 cis(gamma P z)
    to < 1e-8 at (1e-10, 1e-12), far below the 1e-2 a sample-and-hold of the
    previous accepted state would leave.  gamma = 0 with dispersion: the
    samples are cis(D z) u0.  Endpoints are the stored states exactly."""
    nt = 64
    launch = np.fft.ifft(np.full((nt, 1), np.sqrt(2.0), dtype=complex), axis=0)
    fib = oracle.Fiber(np.zeros((nt, 1)), np.full((1, 1, 1, 1), 0.73), np.zeros(nt), np.ones(nt), 1.0)
    zs = np.array([0.0, 0.37, 1.234, 2.5, 3.9, 5.0])
    r = oracle.tsit5_forward(fib, 5.0, launch, rtol=1e-10, atol=1e-12, saveat=zs)
    bounds = np.concatenate([[0.0], np.cumsum(r.hs)])
    inside = [z for z in zs[1:-1] if np.min(np.abs(bounds - z)) > 1e-6]
    assert len(inside) >= 3
    for j, z in enumerate(zs):
        assert _rel(r.saved[j], launch * np.exp(1j * 0.73 * 2.0 * z)) < 1e-8, z
    assert np.array_equal(r.saved[0], launch) and np.array_equal(r.saved[-1], r.out)

    sim, fib_lin, u0, (D, _, hR, S, omf) = _smf(nt=128, window=4.0, length=0.3, power=0.05)
    lin = oracle.Fiber(D, np.zeros((1, 1, 1, 1)), hR, S, omf)
    zl = np.array([0.0, 0.0123, 0.1, 0.2999, 0.3])
    rl = oracle.tsit5_forward(lin, 0.3, u0, rtol=1e-8, atol=1e-6, saveat=zl)
    for j, z in enumerate(zl):
        assert _rel(rl.saved[j], np.exp(1j * D * z) * u0) < 1e-13, z


def test_tsit5_multimode_copies_reduce_to_scalar():
    """M = 2 identical copies of the SMF-28 problem with gamma_iiii only: the
    adaptive multimode solve takes the scalar solve's steps (the RMS error
    norm over N*M equal entries is the scalar norm) and each mode equals the
    scalar output.  A norm over N instead of N*M, or a cross-mode term, fails."""
    sim, fib, u0, (D, g, hR, S, omf) = _smf(nt=256, window=3.0, length=0.1, power=0.1)
    g2 = np.zeros((2, 2, 2, 2))
    g2[0, 0, 0, 0] = g2[1, 1, 1, 1] = g[0, 0, 0, 0]
    multi = oracle.Fiber(np.repeat(D.reshape(-1, 1), 2, axis=1), g2, hR, S, omf)
    rs = oracle.tsit5_forward(fib, 0.1, u0, rtol=1e-8, atol=1e-6)
    rm = oracle.tsit5_forward(multi, 0.1, np.repeat(u0, 2, axis=1), rtol=1e-8, atol=1e-6)
    assert rm.accepted == rs.accepted and rm.rejected == rs.rejected
    assert np.max(np.abs(rm.hs / rs.hs - 1)) < 1e-12
    for i in range(2):
        assert _rel(rm.out[:, i:i + 1], rs.out) < 1e-12


def test_tsit5_strict_vs_rk4_crosscheck():
    """test_scalar_crosscheck.jl:58-75: SMF28 0.02 m, 0.02 W, Nt=128, 4 ps;
    128 fixed RK4 steps vs the adaptive solver at (1e-11, 1e-12): < 1e-7,
    with and without Raman."""
    for fr in (0.18, 0.0):
        sim, fib, u0, _ = _smf(nt=128, window=4.0, length=0.02, power=0.02, fR=fr)
        strict = oracle.tsit5_forward(fib, 0.02, u0, rtol=1e-11, atol=1e-12).out
        assert _rel(oracle.rk4_forward(fib, 0.02, 128, u0), strict) < 1e-7


def test_tsit5_published_counterfactual_shift():
    """The published Delta C (counterfactual-raman-benchmark.md:75-79) was
    computed with the reference's adaptive solver at reltol 1e-8, abstol 1e-6
    (simulate_disp_mmf.jl:186-190).  The same solver settings reproduce the
    three published 6-decimal values within their rounding (5e-7 THz)."""
    pins = json.loads((GOLDEN / "published_pins.json").read_text())["counterfactual_shift_thz"]
    sim = R.sim_params(1550e-9, 1, 1024, 10.0, 3)
    D = R.dispersion_user(sim, [-2.17e-26, 1.2e-40])
    _, u0 = R.initial_state([1.0], 0.1, 185e-15, 80.5e6, "sech_sq", sim)
    basis = R.taylor_phase_basis(sim, (2, 3))
    fibs = {fr: oracle.Fiber(D, np.full((1, 1, 1, 1), 1.1e-3), R.raman_fields(sim, fr)["hR"], R.shock(sim),
                             1 - fr) for fr in (0.18, 0.0)}
    for pin in pins:
        launch = u0 * np.exp(1j * (basis @ np.array([pin["phi2_fs2"], pin["phi3_fs3"]])))[:, None]
        c = [R.centroid_thz(oracle.tsit5_forward(fibs[fr], 0.5, launch, rtol=1e-8, atol=1e-6).out, sim)
             for fr in (0.18, 0.0)]
        assert abs((c[0] - c[1]) - pin["delta_c_thz"]) <= 5e-7, (pin, c[0] - c[1])


def test_tsit5_discrete_adjoint_matches_fd_and_is_checkpoint_invariant():
    """Discrete adjoint of the accepted-step map against central differences
    of the band objective through the same step sequence, and identical
    results for checkpoint spacings 1, 3 and the whole run."""
    sim, fib, u0, _ = _smf(nt=128, window=3.0, length=0.2, power=0.2)
    mask = R.frequency_band_mask(sim, -16.0, -10.0)
    res = oracle.tsit5_forward(fib, 0.2, u0, rtol=1e-6, atol=1e-6)
    hs = res.hs
    assert res.accepted == hs.size > 3
    J0, seed = R.band_cost(oracle.tsit5_replay(fib, 0.2, hs, u0), mask)
    lams = [oracle.tsit5_adjoint(fib, 0.2, hs, u0, seed, every=e) for e in (1, 3, hs.size)]
    assert _rel(lams[1], lams[0]) < 1e-13 and _rel(lams[2], lams[0]) < 1e-13
    lam = lams[0]
    rng = np.random.default_rng(7)
    for _ in range(4):
        d = rng.standard_normal((128, 1)) + 1j * rng.standard_normal((128, 1))
        eps = 1e-6 * np.linalg.norm(u0) / np.linalg.norm(d)
        jp = R.band_cost(oracle.tsit5_replay(fib, 0.2, hs, u0 + eps * d), mask)[0]
        jm = R.band_cost(oracle.tsit5_replay(fib, 0.2, hs, u0 - eps * d), mask)[0]
        fd = (jp - jm) / (2 * eps)
        ad = 2 * np.real(np.vdot(lam, d))
        assert abs(fd - ad) <= 1e-6 * abs(ad) + 1e-12, (fd, ad)


def test_multimode_discrete_adjoint_shared_phase_vs_fd():
    """M = 3: band energy summed over modes (mmf_cost_sum, mmf_cost.jl:49-62),
    shared spectral phase on all modes (mmf_raman_optimization.jl:268-271):
    the discrete adjoint's gradient against central differences."""
    sim, D, rf, S, u0, rng = _mm_setup(3, nt=128, seed=5)
    g = _sym_gamma(3, rng, 5e-2)
    fib = oracle.Fiber(D, g, rf["hR"], S, rf["one_m_fR"])
    L, steps = 0.05, 40
    mask = R.frequency_band_mask(sim, -16.0, -4.0)
    J, lamL = R.band_cost(oracle.rk4_forward(fib, L, steps, u0), mask)
    lam0 = oracle.rk4_adjoint(fib, L, steps, u0, lamL, every=7)
    grad = np.sum(2 * np.real(np.conj(lam0) * 1j * u0), axis=1)
    for k in (3, 40, 100):
        e = 1e-5
        ph = np.zeros(128)
        ph[k] = e
        jp = R.band_cost(oracle.rk4_forward(fib, L, steps, u0 * np.exp(1j * ph)[:, None]), mask)[0]
        jm = R.band_cost(oracle.rk4_forward(fib, L, steps, u0 * np.exp(-1j * ph)[:, None]), mask)[0]
        fd = (jp - jm) / (2 * e)
        # absolute floor: FD roundoff on components far below the largest one
        assert abs(fd - grad[k]) <= 1e-6 * abs(fd) + 1e-8 * np.abs(grad).max(), (k, fd, grad[k])


# ------------------------------------------- multimode closed forms (CW)


@pytest.mark.parametrize("m,pattern", [(2, "quad"), (3, "phase"), (4, "quad"), (6, "phase"), (6, "quad")])
def test_multimode_cw_closed_form(m, pattern):
    """Cross-mode Kerr/Raman coupling of disp_mmf! (simulate_disp_mmf.jl:38-53)
    against a closed form: a CW launch at one bin with gamma = a on (i,i,i,i)
    and b on every arrangement of (i,i,j,j) turns every mode at the rate
    (a + 3(M-1)b) P (all in phase) or (a + (2M-3)b) P (half the modes a
    quarter period behind), times the shock factor of the bin, on top of the
    dispersion phase (tests/conftest.py: cw_multimode_case).  Dropping the
    exchange terms (i,j,i,j), (i,j,j,i) would give (a + (M-1)b) P; dropping
    the cross-phase terms (i,i,j,j) would give (a + 2(M-1)b) P."""
    from conftest import cw_multimode_case

    D, g, hR, S, omf, u0, expected, L = cw_multimode_case(m, pattern)
    fib = oracle.Fiber(D, g, hR, S, omf)
    assert _rel(oracle.rk4_forward(fib, L, 1000, u0), expected) < 1e-10
    if m <= 3:
        assert _rel(R.rk4(u0, L, 1000, D, g, hR, S, omf), expected) < 1e-10
    res = oracle.tsit5_forward(fib, L, u0, rtol=1e-12, atol=1e-12)
    assert _rel(res.out, expected) < 1e-9


def test_multimode_cw_closed_form_catches_dropped_exchange_terms():
    """The closed form is sensitive to the cross terms: the oracle run with the
    exchange entries gamma[i,j,i,j], gamma[i,j,j,i] (i != j) zeroed -- the
    mistake of contracting only the cross-phase part -- misses it by far more
    than the tolerance of the test above."""
    from conftest import cw_multimode_case

    D, g, hR, S, omf, u0, expected, L = cw_multimode_case(3, "phase")
    bad = g.copy()
    for i in range(3):
        for j in range(3):
            if i != j:
                bad[i, j, i, j] = bad[i, j, j, i] = 0.0
    out = oracle.rk4_forward(oracle.Fiber(D, bad, hR, S, omf), L, 400, u0)
    assert _rel(out, expected) > 0.1


# ------------------------------------------------------ multimode costs


@pytest.mark.parametrize("kind", ["sum", "fundamental", "worst_mode"])
def test_mmf_cost_seed_is_wirtinger_derivative(kind):
    """The seeds of mmf_cost.jl's three costs (restated in np_ref.mmf_costs)
    are dJ/dconj(u): for any perturbation du, dJ = 2 Re sum conj(seed) du
    (central differences, relative 1e-7)."""
    rng = np.random.default_rng(11)
    nt, m = 64, 3
    u = (rng.normal(size=(nt, m)) + 1j * rng.normal(size=(nt, m))) * np.array([1.0, 0.7, 1.3])
    mask = np.zeros(nt, dtype=bool)
    mask[5:20] = True
    J, seed = R.mmf_costs(u, mask, kind, tau=7.0)
    for trial in range(3):
        du = rng.normal(size=(nt, m)) + 1j * rng.normal(size=(nt, m))
        e = 1e-6
        fd = (R.mmf_costs(u + e * du, mask, kind, 7.0)[0] - R.mmf_costs(u - e * du, mask, kind, 7.0)[0]) / (2 * e)
        ad = 2 * np.real(np.sum(np.conj(seed) * du))
        assert abs(fd - ad) <= 1e-7 * max(abs(fd), 1e-12)


def test_mmf_costs_bounds_and_single_mode():
    """mmf_cost.jl:112-114: max(r) - log(K)/tau <= J_worst <= max(r) over the K
    modes with energy (a zero mode is skipped); fundamental = r_1; for one
    mode the three costs coincide with the band fraction."""
    rng = np.random.default_rng(12)
    nt = 64
    mask = np.zeros(nt, dtype=bool)
    mask[40:60] = True
    u = rng.normal(size=(nt, 4)) + 1j * rng.normal(size=(nt, 4))
    u[:, 2] = 0
    r = [np.sum(np.abs(u[mask, k]) ** 2) / np.sum(np.abs(u[:, k]) ** 2) for k in (0, 1, 3)]
    J, seed = R.mmf_costs(u, mask, "worst_mode", tau=50.0)
    assert max(r) - np.log(3) / 50 <= J <= max(r)
    assert np.all(seed[:, 2] == 0)
    assert math.isclose(R.mmf_costs(u, mask, "fundamental")[0], r[0], rel_tol=1e-14)
    u1 = u[:, :1]
    costs = [R.mmf_costs(u1, mask, k)[0] for k in ("sum", "fundamental", "worst_mode")]
    assert max(costs) - min(costs) <= 1e-15
    assert math.isclose(costs[0], R.band_cost(u1, mask)[0], rel_tol=1e-14)
