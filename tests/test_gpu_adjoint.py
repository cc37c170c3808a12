"""GPU parity of the fused discrete-adjoint gradient (single mode).

The oracle is the C restatement's discrete adjoint of the same fixed-step
RK4IP map (oracle/fiberoracle.c:orc_rk4_adjoint), itself pinned to central
finite differences and to the reference's continuous adjoint equations
(tests/test_oracle_pins.py).  FP64 path: lam0 within 1e-10 relative of the
oracle.  FP32-bulk path: gradient within 1e-4 relative (normwise) of the
FP64 path, cost within 1e-5.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import oracle_fiber, rel, smf_problem

pytestmark = pytest.mark.gpu


def _band_problem(nt, length, power, window, steps):
    import fiberhip

    problem = smf_problem(nt=nt, window_ps=window, length_m=length, power_w=power, policy="exact")
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
    return problem, mask


def _phases(problem, n, seed):
    import fiberhip

    rng = np.random.default_rng(seed)
    basis = fiberhip.taylor_phase_basis(problem, (2, 3))
    return [basis @ np.array([rng.uniform(-1e4, 1e4), rng.uniform(-5e5, 5e5)]) for _ in range(n)]


@pytest.mark.parametrize("every", [7, 50])
def test_fp64_gradient_matches_oracle(gpu_context, every):
    import fiberhip
    import oracle
    from oracle import np_ref as R

    problem, mask = _band_problem(1024, 0.5, 0.2, 10.0, 60)
    steps = 60
    plan = fiberhip.Plan.from_problem(problem, steps=steps, precision="fp64", band_mask=mask,
                                      checkpoint_every=every, max_batch=3)
    phases = _phases(problem, 3, 11)
    launches = np.stack([problem.uω0 * np.exp(1j * p)[:, None] for p in phases])
    print(f"case: fp64 gradient nt=1024 steps={steps} every={every}", flush=True)
    res = plan.gradient(launches)
    fib = oracle_fiber(problem)
    for b in range(3):
        uL = oracle.rk4_forward(fib, problem.fiber["L"], steps, launches[b])
        J, seed = R.band_cost(uL, mask)
        lam0 = oracle.rk4_adjoint(fib, problem.fiber["L"], steps, launches[b], seed, every=every)
        assert rel(res["uL"][b], uL) <= 1e-11
        assert abs(res["cost"][b] - J) <= 1e-12 * J
        assert rel(res["lam0"][b], lam0) <= 1e-10
        g = R.phase_gradient(lam0, launches[b])
        assert rel(res["grad"][b], g) <= 1e-10
    # caller seed gives the same answer as the on-device band objective
    seeded = plan.gradient(launches[1], seed=R.band_cost(res["uL"][1], mask)[1])
    assert rel(seeded["lam0"], res["lam0"][1]) <= 1e-12
    assert np.isnan(seeded["cost"])


def test_smf28_16384_reduced_fp64_and_fd(gpu_context):
    """SMF-28 gradient geometry (SMF-28, 2 m, 0.2 W, Nt=16384, 47.061 ps, band
    -16..-10 THz, checkpoint every 50) with 200 steps for the oracle run;
    central differences on random bins with h = 1e-4 rad."""
    import fiberhip
    import oracle
    from oracle import np_ref as R

    problem = smf_problem(nt=16384, window_ps=47.061, length_m=2.0, power_w=0.2, policy="exact")
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
    steps = 200
    phase = _phases(problem, 1, 2)[0]
    launch = problem.uω0 * np.exp(1j * phase)[:, None]
    plan = fiberhip.Plan.from_problem(problem, steps=steps, precision="fp64", band_mask=mask, checkpoint_every=50,
                                      max_batch=8)
    print("case: smf28-16384 reduced fp64 nt=16384 steps=200", flush=True)
    res = plan.gradient(launch)
    # eight identical solves in one launch return identical bits
    many = plan.gradient(np.stack([launch] * 8))
    for b in range(8):
        assert np.array_equal(many["lam0"][b], res["lam0"]), b
        assert np.array_equal(many["grad"][b], res["grad"]), b
    fib = oracle_fiber(problem)
    uL = oracle.rk4_forward(fib, 2.0, steps, launch)
    J, seed = R.band_cost(uL, mask)
    lam0 = oracle.rk4_adjoint(fib, 2.0, steps, launch, seed, every=50)
    assert rel(res["lam0"], lam0) <= 1e-10
    assert abs(res["cost"] - J) <= 1e-12 * J
    # FD on bins where the gradient is significant
    rng = np.random.default_rng(3)
    g = res["grad"]
    cand = np.nonzero(np.abs(g) > 1e-3 * np.abs(g).max())[0]
    bins = rng.choice(cand, size=8, replace=False)
    h = 1e-4
    fwd = fiberhip.Plan.from_problem(problem, steps=steps, precision="fp64", max_batch=16)
    pl = []
    for k in bins:
        for sgn in (1, -1):
            ph = phase.copy()
            ph[k] += sgn * h
            pl.append(problem.uω0 * np.exp(1j * ph)[:, None])
    outs, _ = fwd.forward(np.stack(pl))
    for j, k in enumerate(bins):
        jp = R.band_cost(outs[2 * j], mask)[0]
        jm = R.band_cost(outs[2 * j + 1], mask)[0]
        fd = (jp - jm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-5 * abs(fd) + 1e-9 * np.abs(g).max(), (k, fd, g[k])


def test_fp32_gradient_close_to_fp64(gpu_context):
    import fiberhip

    problem, mask = _band_problem(8192, 1.0, 0.1, 23.531, 400)
    steps = 400
    phases = _phases(problem, 4, 5)
    launches = np.stack([problem.uω0 * np.exp(1j * p)[:, None] for p in phases])
    res = {}
    for prec in ("fp64", "fp32"):
        plan = fiberhip.Plan.from_problem(problem, steps=steps, precision=prec, band_mask=mask, max_batch=4)
        print(f"case: gradient {prec} nt=8192 steps={steps}", flush=True)
        res[prec] = plan.gradient(launches)
    for b in range(4):
        gerr = rel(res["fp32"]["grad"][b], res["fp64"]["grad"][b])
        cerr = abs(res["fp32"]["cost"][b] - res["fp64"]["cost"][b]) / res["fp64"]["cost"][b]
        print(f"fp32 vs fp64 gradient rel {gerr:.3e} cost rel {cerr:.3e}", flush=True)
        assert gerr <= 1e-4
        assert cerr <= 1e-5


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_gradient_batch_members_equal_alone(gpu_context, precision):
    import fiberhip

    problem, mask = _band_problem(2048, 0.3, 0.1, 8.0, 30)
    plan = fiberhip.Plan.from_problem(problem, steps=30, precision=precision, band_mask=mask, max_batch=4,
                                      checkpoint_every=8)
    launches = np.stack([problem.uω0 * np.exp(1j * p)[:, None] for p in _phases(problem, 4, 9)])
    batch = plan.gradient(launches)
    for b in (0, 3):
        alone = plan.gradient(launches[b])
        assert np.array_equal(alone["grad"], batch["grad"][b])
        assert alone["cost"] == batch["cost"][b]


def test_adjoint_contract_api(gpu_context):
    """run_adjoint_step / check_adjoint_gradient through the GPU model, with
    the Nyquist-seam probe of test_numerical_kernels.jl:133-164 (weights
    objective, coordinates 2, nt/2, nt/2+1, nt in 1-based indexing)."""
    import fiberhip

    nt = 64
    sim = fiberhip.get_disp_sim_params(1550e-9, 1, nt, 3.0, 2)
    t_ps = sim["ts"] * 1e12
    temporal = (np.exp(-0.5 * (t_ps / 0.07) ** 2) * np.exp(1j * 8 * (t_ps / 0.07) ** 2))[:, None]
    fib = fiberhip.get_disp_fiber_params_user_defined(0.05, sim, fR=0.18, gamma_user=2.0, betas_user=[0.0])
    problem = fiberhip.fiber_field_problem(np.fft.ifft(temporal, axis=0), fib, sim)
    w = np.linspace(0.3, 1.7, nt)[:, None]
    objective = fiberhip.ObjectiveMap("nyquist_probe", cost=lambda u: float(np.sum(w * np.abs(u) ** 2)),
                                      terminal_adjoint=lambda u: w * u)
    model = fiberhip.fiber_model(problem, steps=200)
    check = fiberhip.check_adjoint_gradient(model, fiberhip.FullGridPhase(problem), objective, np.zeros(nt),
                                            coordinate_indices=[1, nt // 2 - 1, nt // 2, nt - 1], step=1e-6,
                                            atol=1e-12, rtol=1e-4)
    assert check.pass_
    assert np.max(check.relative_error) < 2e-5
    # Taylor-basis control: pullback is B^T g
    basis = fiberhip.taylor_phase_basis(problem, (2, 3), coefficient_scales_fs=(1000.0, 1e4))
    step = fiberhip.run_adjoint_step(model, fiberhip.PhaseBasis(basis), objective, np.array([0.3, -0.2]))
    assert np.allclose(step.control_gradient, basis.T @ step.physical_gradient, rtol=1e-14)
