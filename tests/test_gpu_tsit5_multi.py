"""GPU parity of the adaptive multimode path (FH_SCHEME_TSIT5, M > 1).

Oracle: oracle/fiberoracle.c orc_tsit5_forward(_save) on the multimode
right-hand side (the adaptive solver and the multimode RHS are each pinned in
tests/test_oracle_pins.py; the copies-of-a-scalar reduction below pins the
N*M error norm).  FP64: same accepted and rejected step counts as the
oracle, per-mode output energy within 1e-10, saved spectra within 1e-10.
FP32-bulk: per-mode energy within 1e-5 of FP64.  The adaptive multimode
gradient: lam0, cost and gradient within 1e-10 of the oracle's discrete
adjoint of the same accepted steps (FP64); FP32 cost within 1e-5.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import mode_energy_rel, oracle_fiber, rel, smf_problem

pytestmark = pytest.mark.gpu


def _grin(m=3, nt=1024, length=0.05, energy=2e-9, seed=1, window_ps=10.0):
    from fiberhip import grin

    return grin.grin_problem(m, nt=nt, time_window_ps=window_ps, length_m=length, energy_j=energy, seed=seed)


def test_fp64_matches_oracle_steps_energy_and_saveat(gpu_context):
    import fiberhip
    import oracle

    gp = _grin()
    zs = [0.0, 0.0123, 0.025, 0.05]
    print("case: tsit5 multimode fp64 M=3 nt=1024", flush=True)
    plan = fiberhip.Plan.from_problem(gp, scheme="tsit5", rtol=1e-8, atol=1e-6)
    out, stats, saved = plan.forward_saveat(gp.uω0, zs)
    ref = oracle.tsit5_forward(oracle_fiber(gp), 0.05, gp.uω0, rtol=1e-8, atol=1e-6, saveat=zs)
    print("accepted", stats["accepted"], ref.accepted, "energy rel", mode_energy_rel(out, ref.out).max(), flush=True)
    assert stats["accepted"] == ref.accepted and stats["rejected"] == ref.rejected
    assert stats["rhs_evaluations"] == ref.rhs_evals
    assert mode_energy_rel(out, ref.out).max() <= 1e-10
    assert rel(out, ref.out) <= 1e-9
    for j in range(len(zs)):
        assert rel(saved[j], ref.saved[j]) <= 1e-10, zs[j]
    assert np.array_equal(saved[-1], out)
    hs = plan.step_sizes(1)[0]
    assert np.max(np.abs(hs[:-1] / ref.hs[:-1] - 1)) <= 1e-10


def test_identical_mode_copies_reduce_to_scalar(gpu_context):
    """M copies of one scalar problem with gamma_iiii only: every mode follows
    the scalar solve and the N*M error norm equals the scalar one, so the
    multimode driver takes the scalar driver's steps."""
    import fiberhip

    problem = smf_problem(nt=1024, window_ps=10.0, length_m=0.3, power_w=0.1)
    f = problem.fiber
    m = 3
    g = np.zeros((m, m, m, m))
    for i in range(m):
        g[i, i, i, i] = float(np.asarray(f["γ"]).reshape(-1)[0])
    D = np.repeat(np.asarray(f["Dω"]).reshape(-1, 1), m, axis=1)
    shock = fiberhip.self_steepening(problem.sim)
    launch = np.repeat(problem.uω0, m, axis=1)
    print("case: tsit5 multimode copies-of-scalar fp64", flush=True)
    multi = fiberhip.Plan(D, g, f["hRω"], shock, f["one_m_fR"], f["L"], None, scheme="tsit5", rtol=1e-8, atol=1e-6)
    single = fiberhip.Plan.from_problem(problem, scheme="tsit5", rtol=1e-8, atol=1e-6)
    om, sm = multi.forward(launch)
    os_, ss = single.forward(problem.uω0)
    assert sm["accepted"] == ss["accepted"] and sm["rejected"] == ss["rejected"]
    for i in range(m):
        assert rel(om[:, i], os_[:, 0]) <= 1e-12


def test_fp32_within_1e5_of_fp64(gpu_context):
    gp = _grin(m=6, nt=4096, length=0.05, energy=5e-9, window_ps=12.0)
    import fiberhip

    out = {}
    for precision in ("fp64", "fp32"):
        print(f"case: tsit5 multimode {precision} M=6 nt=4096", flush=True)
        plan = fiberhip.Plan.from_problem(gp, scheme="tsit5", rtol=1e-7, atol=1e-6, precision=precision)
        out[precision] = plan.forward(gp.uω0)[0]
    err = mode_energy_rel(out["fp32"], out["fp64"]).max()
    print("fp32 vs fp64 per-mode energy rel", err, flush=True)
    assert err <= 1e-5


def test_batch_members_run_their_own_steps(gpu_context):
    """Four launches of different mode weights (FP32, M=4): each keeps its own
    step sequence and returns the same bits alone as in the batch."""
    import fiberhip
    from fiberhip import grin

    base = _grin(m=4, nt=2048, length=0.1, energy=5e-9)
    launches = np.stack([grin.grin_problem(4, nt=2048, time_window_ps=10.0, length_m=0.1, energy_j=5e-9,
                                           seed=s).uω0 for s in (1, 2, 3, 4)])
    print("case: tsit5 multimode batch M=4 fp32", flush=True)
    plan = fiberhip.Plan.from_problem(base, scheme="tsit5", rtol=1e-6, atol=1e-6, precision="fp32", max_batch=4)
    batch, _ = plan.forward(launches)
    counts = plan.solve_stats(4)[:, 0]
    assert len(set(counts.tolist())) > 1, counts
    for b in range(4):
        alone, st = plan.forward(launches[b])
        assert st["accepted"] == counts[b]
        assert np.array_equal(alone, batch[b])


def test_propagate_multimode_default_is_adaptive(gpu_context):
    import fiberhip

    gp = _grin()
    print("case: propagate multimode adaptive + verify", flush=True)
    res = fiberhip.propagate(gp, saveat=[0.0, 0.025, 0.05])
    assert res.solver == "Tsit5" and res.spectra.shape == (1024, 3, 3)
    ev = fiberhip.metrics(res)
    assert ev["max_photon_number_drift"] < 1e-6
    assert fiberhip.verify(res)["pass_"]


def _small_grin(modes, nt, length=0.01, energy=2e-9, seed=1):
    from fiberhip import grin

    window = max(3.0, 1.2 * nt / (2 * 193.41449))
    return grin.grin_problem(modes, nt=nt, time_window_ps=window, length_m=length, energy_j=energy, nx=96,
                             window_um=60.0, seed=seed)


@pytest.mark.parametrize("modes,nt", [(3, 512), (6, 256)])
def test_fp64_adaptive_gradient_matches_oracle_adjoint_of_same_steps(gpu_context, modes, nt):
    """The adaptive multimode gradient is the discrete adjoint of the GPU's own
    accepted-step map: lam0, cost and the shared-phase gradient within 1e-10
    of the oracle's adjoint of that step sequence (orc_tsit5_adjoint, pinned
    against central differences for the scalar case and by the multimode
    RHS pullback pins)."""
    import fiberhip
    import oracle
    from oracle import np_ref as R

    p = _small_grin(modes, nt)
    mask = fiberhip.frequency_band_mask(p, (-16.0, -4.0))
    L = p.fiber["L"]
    print(f"case: adaptive multimode gradient fp64 M={modes} nt={nt}", flush=True)
    plan = fiberhip.Plan.from_problem(p, scheme="tsit5", rtol=1e-8, atol=1e-6, band_mask=mask, checkpoint_every=7)
    res = plan.gradient(p.uω0)
    hs = plan.step_sizes(1)[0]
    assert hs.size > 7
    fib = oracle_fiber(p)
    uL = oracle.tsit5_replay(fib, L, hs, p.uω0)
    J, seed = R.band_cost(uL, mask)
    lam0 = oracle.tsit5_adjoint(fib, L, hs, p.uω0, seed, every=50)
    grad = np.sum(2 * np.real(np.conj(lam0) * 1j * p.uω0), axis=1)
    print("steps", hs.size, "lam0 rel", rel(res["lam0"], lam0), "cost rel", abs(res["cost"] / J - 1), flush=True)
    assert rel(res["uL"], uL) <= 1e-10
    assert abs(res["cost"] - J) <= 1e-11 * J
    assert rel(res["lam0"], lam0) <= 1e-10
    assert rel(res["grad"], grad) <= 1e-10
    # the forward solve takes the same steps
    out, st = fiberhip.Plan.from_problem(p, scheme="tsit5", rtol=1e-8, atol=1e-6).forward(p.uω0)
    assert st["accepted"] == hs.size and rel(out, res["uL"]) <= 1e-13


def test_adaptive_gradient_fp32_and_identical_members(gpu_context):
    import fiberhip

    p = _small_grin(4, 1024, length=0.02, energy=3e-9)
    mask = fiberhip.frequency_band_mask(p, (-16.0, -4.0))
    res = {}
    for precision in ("fp64", "fp32"):
        print(f"case: adaptive multimode gradient identical members {precision} M=4", flush=True)
        plan = fiberhip.Plan.from_problem(p, scheme="tsit5", rtol=1e-7, atol=1e-6, band_mask=mask,
                                          precision=precision, max_batch=4)
        r = plan.gradient(np.stack([p.uω0] * 4))
        for key in ("lam0", "grad", "uL", "cost"):
            assert all(np.array_equal(r[key][b], r[key][0]) for b in range(4)), (precision, key)
        res[precision] = r
    cost_err = abs(res["fp32"]["cost"][0] / res["fp64"]["cost"][0] - 1)
    grad_err = rel(res["fp32"]["grad"][0], res["fp64"]["grad"][0])
    print("fp32 vs fp64 cost rel", cost_err, "grad normwise", grad_err, flush=True)
    assert cost_err <= 1e-5 and grad_err <= 1e-2
