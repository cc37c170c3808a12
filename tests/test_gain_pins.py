"""Pump + signal propagation with gain (the YDFA amplifier model): pins of
the oracle's restatement (oracle/fiberoracle.c: orc_gain_rhs,
orc_gain_forward) and of the host's folding of the gain model into the
kernel's coefficients (fiberhip/gain.py), on the CPU.

Reference: disp_gain_smf! and solve_disp_gain_smf
(src/simulation/simulate_disp_gain_mmf.jl:21-63,200-239), YDFAParams,
get_ydfa_cross_sections, psd_from_uω and calculate_gain_YDFA
(src/gain_simulation/gain.jl:30-182).  The reference's own test
(test/core/test_numerical_kernels.jl:5-24) checks that the cross sections
are finite and non-negative on the grid and that a short solve succeeds with
finite output; the pins below add what the mathematics fixes:

* constant gain with the nonlinearity off has the closed form
  x(L) = x(0) exp(g L / 2), Pp(L) = Pp(0) exp(-L);
* with the spontaneous decay switched off (1/tau21 = 0) the steady-state rate
  equations conserve photons: the pump photon flux Pp / (h nu_p) plus the
  signal's sum Ps(nu) df / (h nu) is constant along the fiber, with Kerr,
  Raman and self-steepening on (they conserve the signal's photon number);
* with no pump and a vanishing signal the inversion is zero and every bin is
  absorbed as exp(-Gamma_s rho sigma_as(nu) L) -- which fixes the
  cross-section-to-bin order;
* the folded coefficients reproduce a literal restatement of gain.jl
  (shifted grid, PSD, W12/W21, n2, g, gP) bin by bin.
"""
from __future__ import annotations

import json
import math
import pathlib

import numpy as np
import pytest

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def yb_tables() -> dict:
    return json.loads((GOLDEN / "yb_cross_sections.json").read_text())


def gain_case(nt=512, window_ps=5.0, length_m=0.5, p_avg=0.05, pump=0.5, gamma=1e-3, fR=0.18, tau21=None,
              betas=(20e-27, 0.0), beta_order=3, pulse="gauss", fwhm_s=300e-15):
    """A YDFA amplifier problem at 1030 nm: (sim, YDFAParams, launch uω0,
    fiber dict, pump power)."""
    from fiberhip import gain as G
    from fiberhip import problem as P

    xs = yb_tables()
    sim = P.get_disp_sim_params(1030e-9, 1, nt, window_ps, beta_order)
    cs = G.get_ydfa_cross_sections(sim["fs"], xs["absorption"]["wavelength_nm"], xs["absorption"]["intensity"],
                                   xs["emission"]["wavelength_nm"], xs["emission"]["intensity"], xs["scale_m2"])
    p = G.get_YDFAParams(sim, cs)
    if tau21 is not None:
        p.tau21 = tau21
    _, u0 = P.get_initial_state([1.0], p_avg, fwhm_s, 80e6, pulse, sim)
    fib = P.get_disp_fiber_params_user_defined(length_m, sim, fR=fR, gamma_user=gamma, betas_user=list(betas))
    fib["gain_parameters"] = p
    return sim, p, u0, fib, pump


def oracle_gain(sim, fib, coef):
    import oracle
    from fiberhip import problem as P

    return (oracle.Fiber(fib["Dω"], fib["γ"], fib["hRω"], P.self_steepening(sim), fib["one_m_fR"]),
            oracle.Gain(coef))


def photon_flux(p, u, pump) -> float:
    """Pump photons Pp / (h nu_p) plus signal photons sum Ps df / (h nu) [1/s]."""
    from fiberhip import gain as G

    ps = G.psd_from_uω(np.asarray(u).reshape(-1), p)
    return pump / (G.H_PLANCK * p.νp) + float(np.sum(ps * p.df / (G.H_PLANCK * p.νs)))


def test_reference_gain_case_on_32_samples():
    """test_numerical_kernels.jl:5-24: 1030 nm, Nt = 32, 5 ps; cross sections
    of length Nt, finite and non-negative; a 10 um solve (gamma = 0, fR = 0,
    beta = 0, Gaussian 300 fs at 1e-4 W, pump 0.05 W) succeeds with finite
    output."""
    import oracle
    from fiberhip import gain as G

    sim, p, u0, fib, _ = gain_case(nt=32, window_ps=5.0, length_m=1e-5, p_avg=1e-4, gamma=0.0, fR=0.0, betas=(0.0,),
                                   beta_order=2)
    for s in (p.σas, p.σes):
        assert s.shape == (32,) and np.all(np.isfinite(s)) and np.all(s >= 0)
    of, og = oracle_gain(sim, fib, G.coefficients(p, 32))
    out, pump_l, _, stats = oracle.gain_forward(of, og, fib["L"], u0, 0.05)
    assert np.all(np.isfinite(out)) and math.isfinite(pump_l) and stats[0] > 0


def test_cross_section_interpolation_is_flat_outside_the_tables():
    from fiberhip import gain as G

    xs = yb_tables()
    lam_a = np.array(xs["absorption"]["wavelength_nm"])
    fs = 2.99792458e8 / (np.array([lam_a[0] - 20, lam_a[0], lam_a[-1], lam_a[-1] + 20]) * 1e-9) / 1e12
    cs = G.get_ydfa_cross_sections(fs, lam_a, xs["absorption"]["intensity"], xs["emission"]["wavelength_nm"],
                                   xs["emission"]["intensity"])
    a = np.array(xs["absorption"]["intensity"]) * 1e-27
    np.testing.assert_allclose(cs["sigma_as"], [a[0], a[0], a[-1], a[-1]], rtol=1e-12)
    # Marcuse overlap factors of the reference's defaults lie in (0, 1)
    p = G.YDFAParams(fs=np.array([291.0]), Δt=0.01, Nt=1)
    assert 0 < p.Gamma_s < 1 and 0 < p.Gamma_p < 1 and p.Gamma_p > p.Gamma_s


def test_folded_coefficients_reproduce_gain_jl():
    """The per-bin coefficients the kernel evaluates give the same W12/W21,
    n2, g(nu) (after ifftshift) and gP as a literal restatement of
    calculate_gain_YDFA on psd_from_uω's shifted PSD (gain.jl:115-166)."""
    from fiberhip import gain as G

    sim, p, u0, fib, pump = gain_case(nt=256)
    rng = np.random.default_rng(11)
    u = u0[:, 0] * (1 + 0.3 * rng.standard_normal(256)) * np.exp(1j * rng.uniform(0, 6.28, 256))
    g_ref, gp_ref = G.calculate_gain_YDFA(pump, G.psd_from_uω(u[:, None], p), p)
    c = G.ydfa_coefficients(p)
    a2 = np.abs(u) ** 2
    w12, w21 = np.sum(c.w12 * a2), np.sum(c.w21 * a2)
    n2 = (c.r12 * pump + w12) / (c.r12 * pump + c.r21 * pump + w12 + w21 + c.inv_tau)
    g = c.ge * n2 - c.ga * (1 - n2)
    np.testing.assert_allclose(g, np.fft.ifftshift(g_ref), rtol=1e-12, atol=1e-12 * np.abs(g_ref).max())
    assert abs((c.gpe * n2 - c.gpa * (1 - n2)) - gp_ref) <= 1e-12 * abs(gp_ref)


def test_oracle_gain_rhs_adds_gain_to_disp_mmf():
    """orc_gain_rhs = disp_mmf! + g/2 .* x, dPp = gP Pp (simulate_disp_gain_mmf.jl:61-62)."""
    import oracle
    from fiberhip import gain as G

    sim, p, u0, fib, pump = gain_case(nt=256)
    of, og = oracle_gain(sim, fib, G.ydfa_coefficients(p))
    z = 0.013
    dpp, dx = oracle.gain_rhs(of, og, z, pump, u0)
    # the gain is evaluated on the lab spectrum cis(Dz) x, whose |.|^2 is that of x
    lab = np.exp(1j * fib["Dω"][:, 0] * z) * u0[:, 0]
    g_ref, gp_ref = G.calculate_gain_YDFA(pump, G.psd_from_uω(lab[:, None], p), p)
    expect = oracle.rhs(of, z, u0)[:, 0] + 0.5 * np.fft.ifftshift(g_ref) * u0[:, 0]
    np.testing.assert_allclose(dx[:, 0], expect, rtol=0, atol=1e-12 * np.abs(expect).max())
    assert abs(dpp - gp_ref * pump) <= 1e-12 * abs(gp_ref * pump)


def test_constant_gain_closed_form():
    """compute_gain(uω, pGain::Number, Pp) (simulate_disp_gain_mmf.jl:76-80)
    with the nonlinearity off: cis(D L) x0 exp(g L / 2) and Pp0 exp(-L)."""
    import oracle
    from fiberhip import gain as G

    sim, p, u0, fib, _ = gain_case(nt=256, gamma=0.0, fR=0.0, length_m=0.7)
    g = 1.3
    of, og = oracle_gain(sim, fib, G.constant_coefficients(g, 256))
    out, pump_l, _, _ = oracle.gain_forward(of, og, fib["L"], u0, 0.2, rtol=1e-11, atol=1e-14)
    L = fib["L"]
    expect = np.exp(1j * fib["Dω"][:, 0] * L) * u0[:, 0] * math.exp(g * L / 2)
    assert np.linalg.norm(out[:, 0] - expect) / np.linalg.norm(expect) < 1e-9
    assert abs(pump_l - 0.2 * math.exp(-L)) < 1e-9 * 0.2


def test_photon_balance_without_spontaneous_decay():
    """1/tau21 = 0: pump + signal photon flux is conserved (Kerr, Raman and
    self-steepening on), while the pump visibly amplifies the signal; with the
    reference's tau21 the flux falls (spontaneous decay)."""
    import oracle
    from fiberhip import gain as G

    sim, p, u0, fib, pump = gain_case(nt=512, tau21=math.inf)
    c = G.ydfa_coefficients(p)
    assert c.inv_tau == 0.0
    of, og = oracle_gain(sim, fib, c)
    out, pump_l, _, _ = oracle.gain_forward(of, og, fib["L"], u0, pump, rtol=1e-10, atol=1e-12)
    n0, n1 = photon_flux(p, u0, pump), photon_flux(p, out, pump_l)
    assert abs(n1 / n0 - 1) < 1e-8
    e0, e1 = (float(np.sum(G.psd_from_uω(u, p)) * p.df) for u in (u0, out))
    assert e1 > 2 * e0 and pump_l < 0.9 * pump  # the signal gained, the pump depleted

    sim, p, u0, fib, pump = gain_case(nt=512)
    of, og = oracle_gain(sim, fib, G.ydfa_coefficients(p))
    out, pump_l, _, _ = oracle.gain_forward(of, og, fib["L"], u0, pump, rtol=1e-10, atol=1e-12)
    assert photon_flux(p, out, pump_l) < photon_flux(p, u0, pump) * (1 - 1e-6)


def test_unpumped_absorption_per_bin():
    """No pump, a 1e-12 W signal: n2 ~ 0 and |x_k(L)|^2 = |x_k(0)|^2
    exp(-Gamma_s rho sigma_as(nu_k) L) bin by bin (FFT order)."""
    import oracle
    from fiberhip import gain as G

    sim, p, u0, fib, _ = gain_case(nt=256, p_avg=1e-12, gamma=0.0, fR=0.0, length_m=0.3)
    of, og = oracle_gain(sim, fib, G.ydfa_coefficients(p))
    out, pump_l, _, _ = oracle.gain_forward(of, og, fib["L"], u0, 0.0, rtol=1e-11, atol=1e-30)
    assert pump_l == 0.0
    sig = np.abs(u0[:, 0]) ** 2 > 1e-6 * np.max(np.abs(u0) ** 2)
    ratio = np.abs(out[sig, 0]) ** 2 / np.abs(u0[sig, 0]) ** 2
    expect = np.exp(-p.Gamma_s * p.rho * np.fft.ifftshift(p.σas)[sig] * fib["L"])
    np.testing.assert_allclose(ratio, expect, rtol=1e-7)
    assert np.ptp(expect) > 0.05  # absorption differs across the bins: the order matters
