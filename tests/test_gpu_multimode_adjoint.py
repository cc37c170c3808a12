"""GPU parity of the multimode discrete adjoint (k_grad_multi).

Oracle: the C restatement's discrete adjoint of the fixed-step RK4IP map for
M modes (oracle/fiberoracle.c: orc_rk4_adjoint), pinned on the CPU against
central differences of the band objective summed over modes
(tests/test_oracle_pins.py).  FP64: lam0 within 1e-10 relative of the oracle;
FP32-bulk: gradient within 1e-3 normwise of FP64 and cost within 1e-5.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def _grin(modes, nt, length=0.01, energy=2e-9, seed=1):
    from fiberhip import grin

    window = max(3.0, 1.2 * nt / (2 * 193.41449))
    return grin.grin_problem(modes, nt=nt, time_window_ps=window, length_m=length, energy_j=energy, nx=96,
                             window_um=60.0, seed=seed)


def _oracle_fiber(p):
    import fiberhip
    import oracle

    f = p.fiber
    return oracle.Fiber(f["Dω"], f["γ"], f["hRω"], fiberhip.self_steepening(p.sim), f["one_m_fR"])


@pytest.mark.parametrize("modes,nt", [(3, 512), (6, 1024), (10, 256)])
def test_fp64_multimode_gradient_matches_oracle(gpu_context, modes, nt):
    import fiberhip
    import oracle
    from oracle import np_ref as R

    p = _grin(modes, nt)
    mask = fiberhip.frequency_band_mask(p, (-16.0, -4.0))
    steps, L = 40, p.fiber["L"]
    print(f"case: multimode gradient fp64 M={modes} nt={nt}", flush=True)
    plan = fiberhip.Plan.from_problem(p, steps=steps, band_mask=mask, checkpoint_every=7)
    res = plan.gradient(p.uω0)
    fib = _oracle_fiber(p)
    uL = oracle.rk4_forward(fib, L, steps, p.uω0)
    J, seed = R.band_cost(uL, mask)
    lam0 = oracle.rk4_adjoint(fib, L, steps, p.uω0, seed, every=50)
    grad = np.sum(2 * np.real(np.conj(lam0) * 1j * p.uω0), axis=1)
    print("lam0 rel", rel(res["lam0"], lam0), "cost rel", abs(res["cost"] / J - 1), flush=True)
    assert rel(res["uL"], uL) <= 1e-10
    assert abs(res["cost"] - J) <= 1e-12 * J
    assert rel(res["lam0"], lam0) <= 1e-10
    assert rel(res["grad"], grad) <= 1e-10


def test_caller_seed_fundamental_mode_cost(gpu_context):
    """A caller seed (mmf_cost_fundamental, mmf_cost.jl:84-100: band fraction of
    mode 1 only) goes through the same pullback."""
    import fiberhip
    import oracle

    p = _grin(3, 512)
    fib = _oracle_fiber(p)
    mask = fiberhip.frequency_band_mask(p, (-16.0, -4.0))
    uL = oracle.rk4_forward(fib, p.fiber["L"], 40, p.uω0)
    u1 = uL[:, 0]
    J = np.sum(np.abs(u1[mask]) ** 2) / np.sum(np.abs(u1) ** 2)
    seed = np.zeros_like(uL)
    seed[:, 0] = u1 * (mask - J) / np.sum(np.abs(u1) ** 2)
    print("case: multimode caller seed", flush=True)
    res = fiberhip.Plan.from_problem(p, steps=40, band_mask=mask).gradient(p.uω0, seed=seed)
    lam0 = oracle.rk4_adjoint(fib, p.fiber["L"], 40, p.uω0, seed, every=50)
    assert rel(res["lam0"], lam0) <= 1e-10
    assert np.isnan(res["cost"])


def test_fp32_close_to_fp64_and_deterministic(gpu_context):
    import fiberhip

    p = _grin(6, 1024, length=0.02, energy=5e-9)
    mask = fiberhip.frequency_band_mask(p, (-16.0, -4.0))
    launches = np.stack([_grin(6, 1024, length=0.02, energy=5e-9, seed=s).uω0 for s in (1, 2, 3)])
    out = {}
    for precision in ("fp64", "fp32"):
        print(f"case: multimode gradient {precision} batch", flush=True)
        plan = fiberhip.Plan.from_problem(p, steps=80, band_mask=mask, precision=precision, max_batch=4)
        out[precision] = plan.gradient(launches)
        again = plan.gradient(np.stack([launches[1]] * 4))
        for b in range(4):
            assert np.array_equal(again["lam0"][b], out[precision]["lam0"][1])
    g64, g32 = out["fp64"]["grad"], out["fp32"]["grad"]
    err = np.max(np.linalg.norm(g32 - g64, axis=1) / np.linalg.norm(g64, axis=1))
    cerr = np.max(np.abs(out["fp32"]["cost"] / out["fp64"]["cost"] - 1))
    print("fp32 vs fp64 gradient", err, "cost", cerr, flush=True)
    assert err <= 1e-3 and cerr <= 1e-5


def test_nonsymmetric_gamma_rejected_like_the_reference(gpu_context):
    """sensitivity_disp_mmf.jl:239-240: the fiber adjoint requires a fully
    permutation-symmetric gamma tensor."""
    import fiberhip

    base = _grin(3, 256)
    g = base.fiber["γ"] * (1 + 0.3 * np.random.default_rng(5).uniform(size=(3, 3, 3, 3)))
    fiber = dict(base.fiber)
    fiber["γ"] = g
    prob = fiberhip.fiber_field_problem(base.uω0, fiber, base.sim)
    mask = fiberhip.frequency_band_mask(prob, (-16.0, -4.0))
    plan = fiberhip.Plan.from_problem(prob, steps=10, band_mask=mask)
    with pytest.raises(fiberhip.FiberHIPError, match="FH_ENOTSYM"):
        plan.gradient(prob.uω0)
