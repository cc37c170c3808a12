"""GPU parity of the adaptive Tsit5 path (FH_SCHEME_TSIT5, single mode).

Oracle: the C restatement's adaptive Tsit5 (oracle/fiberoracle.c:
orc_tsit5_forward), which shares the tableau, error norm and step controller
and is pinned in tests/test_oracle_pins.py (order conditions, closed forms,
the published counterfactual shifts).  FP64: output spectral energy within
1e-10 of the oracle with the same accepted-step count; the discrete adjoint of
the GPU's own accepted-step sequence within 1e-10 of the oracle's adjoint of
that sequence.  FP32-bulk: within 1e-5 relative spectral energy of FP64.
"""
from __future__ import annotations

import json
import pathlib

import numpy as np
import pytest

from conftest import oracle_fiber, rel, smf_problem

pytestmark = pytest.mark.gpu
GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def _phases(problem, n, seed, orders=(2, 3)):
    import fiberhip

    rng = np.random.default_rng(seed)
    basis = fiberhip.taylor_phase_basis(problem, orders)
    scales = {2: 1e4, 3: 5e5, 4: 1e7}
    return [basis @ np.array([rng.uniform(-scales[o], scales[o]) for o in orders]) for _ in range(n)]


def test_fp64_forward_matches_oracle_steps_and_energy(gpu_context):
    import fiberhip
    import oracle

    problem = smf_problem(nt=1024, window_ps=10.0, length_m=0.5, power_w=0.1)
    print("case: tsit5 fp64 nt=1024 rtol=1e-8", flush=True)
    plan = fiberhip.Plan.from_problem(problem, scheme="tsit5", rtol=1e-8, atol=1e-6)
    out, stats = plan.forward(problem.uω0)
    ref = oracle.tsit5_forward(oracle_fiber(problem), 0.5, problem.uω0, rtol=1e-8, atol=1e-6)
    e_gpu = np.sum(np.abs(out) ** 2)
    e_ref = np.sum(np.abs(ref.out) ** 2)
    print("accepted", stats["accepted"], ref.accepted, "energy rel", abs(e_gpu / e_ref - 1), flush=True)
    assert stats["accepted"] == ref.accepted and stats["rejected"] == ref.rejected
    assert stats["rhs_evaluations"] == ref.rhs_evals
    assert abs(e_gpu / e_ref - 1) <= 1e-10
    assert rel(out, ref.out) <= 1e-9
    hs = plan.step_sizes(1)[0]
    # the last step is L - z: compare it against L (z carries rounding of ~1e-13)
    assert np.max(np.abs(hs[:-1] / ref.hs[:-1] - 1)) <= 1e-10
    assert abs(hs[-1] - ref.hs[-1]) <= 1e-12 * 0.5


def test_fp64_reproduces_published_counterfactual_shift(gpu_context):
    """The published Delta C (counterfactual-raman-benchmark.md:75-79) with the
    reference's solver settings (reltol 1e-8, abstol 1e-6), within the
    6-decimal rounding of the published values."""
    import fiberhip
    from oracle import np_ref as R

    pins = json.loads((GOLDEN / "published_pins.json").read_text())["counterfactual_shift_thz"]
    problems = {fr: smf_problem(nt=1024, window_ps=10.0, length_m=0.5, power_w=0.1, fR=fr) for fr in (0.18, 0.0)}
    plans = {fr: fiberhip.Plan.from_problem(p, scheme="tsit5", rtol=1e-8, atol=1e-6, max_batch=len(pins))
             for fr, p in problems.items()}
    sim = R.sim_params(1550e-9, 1, 1024, 10.0, 3)
    basis = R.taylor_phase_basis(sim, (2, 3))
    u0 = problems[0.18].uω0
    launches = np.stack([u0 * np.exp(1j * (basis @ np.array([p["phi2_fs2"], p["phi3_fs3"]])))[:, None] for p in pins])
    print("case: tsit5 fp64 published delta C", flush=True)
    outs = {fr: plans[fr].forward(launches)[0] for fr in (0.18, 0.0)}
    for i, pin in enumerate(pins):
        dc = R.centroid_thz(outs[0.18][i], sim) - R.centroid_thz(outs[0.0][i], sim)
        assert abs(dc - pin["delta_c_thz"]) <= 5e-7, (pin, dc)


def test_fp64_gradient_matches_oracle_adjoint_of_same_steps(gpu_context):
    import fiberhip
    import oracle
    from oracle import np_ref as R

    problem = smf_problem(nt=1024, window_ps=10.0, length_m=0.5, power_w=0.2)
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
    launch = problem.uω0 * np.exp(1j * _phases(problem, 1, 2)[0])[:, None]
    print("case: tsit5 fp64 gradient nt=1024", flush=True)
    plan = fiberhip.Plan.from_problem(problem, scheme="tsit5", rtol=1e-6, atol=1e-6, band_mask=mask,
                                      checkpoint_every=7)
    res = plan.gradient(launch)
    hs = plan.step_sizes(1)[0]
    assert hs.size > 7
    fib = oracle_fiber(problem)
    uL = oracle.tsit5_replay(fib, 0.5, hs, launch)
    J, seed = R.band_cost(uL, mask)
    lam0 = oracle.tsit5_adjoint(fib, 0.5, hs, launch, seed, every=50)
    print("lam0 rel", rel(res["lam0"], lam0), "cost rel", abs(res["cost"] - J) / J, flush=True)
    assert rel(res["uL"], uL) <= 1e-10
    assert abs(res["cost"] - J) <= 1e-11 * J
    assert rel(res["lam0"], lam0) <= 1e-10
    # the gradient and the forward agree on the step sequence
    out, stats = fiberhip.Plan.from_problem(problem, scheme="tsit5", rtol=1e-6, atol=1e-6).forward(launch)
    assert stats["accepted"] == hs.size and np.array_equal(out, res["uL"])


def test_candidates_run_their_own_step_sequences(gpu_context):
    """Candidate sweep shape (SMF-28 1 m, 0.1 W, Nt=8192, GDD/TOD/FOD phases):
    each candidate of one launch keeps its own step sequence and returns the
    same bits as when solved alone."""
    import fiberhip

    problem = smf_problem(nt=8192, window_ps=23.531, length_m=1.0, power_w=0.1)
    phases = _phases(problem, 6, 4, orders=(2, 3, 4))
    launches = np.stack([problem.uω0 * np.exp(1j * ph)[:, None] for ph in phases])
    for precision in ("fp64", "fp32"):
        print(f"case: tsit5 candidates {precision} nt=8192", flush=True)
        plan = fiberhip.Plan.from_problem(problem, scheme="tsit5", rtol=1e-6, atol=1e-6, precision=precision,
                                          max_batch=6)
        batch, _ = plan.forward(launches)
        counts = plan.solve_stats(6)[:, 0]
        assert len(set(counts.tolist())) > 1, counts
        for b in (0, 3, 5):
            alone, st = plan.forward(launches[b])
            assert st["accepted"] == counts[b]
            assert np.array_equal(alone, batch[b])


def test_fp32_bulk_within_1e5_of_fp64(gpu_context):
    import fiberhip

    problem = smf_problem(nt=8192, window_ps=23.531, length_m=1.0, power_w=0.1)
    launches = np.stack([problem.uω0 * np.exp(1j * ph)[:, None] for ph in _phases(problem, 4, 4, orders=(2, 3, 4))])
    out = {}
    for precision in ("fp64", "fp32"):
        print(f"case: tsit5 {precision} nt=8192 rtol=1e-6", flush=True)
        plan = fiberhip.Plan.from_problem(problem, scheme="tsit5", rtol=1e-6, atol=1e-6, precision=precision,
                                          max_batch=4)
        out[precision] = plan.forward(launches)[0]
    e64 = np.sum(np.abs(out["fp64"]) ** 2, axis=1)
    e32 = np.sum(np.abs(out["fp32"]) ** 2, axis=1)
    err = np.max(np.abs(e32 / e64 - 1))
    print("fp32 vs fp64 energy rel", err, flush=True)
    assert err <= 1e-5


def test_gradient_identical_members_and_step_limit(gpu_context):
    import fiberhip

    problem = smf_problem(nt=2048, window_ps=10.0, length_m=0.5, power_w=0.2)
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
    for precision in ("fp64", "fp32"):
        print(f"case: tsit5 gradient identical members {precision}", flush=True)
        plan = fiberhip.Plan.from_problem(problem, scheme="tsit5", rtol=1e-6, atol=1e-6, precision=precision,
                                          band_mask=mask, max_batch=8)
        res = plan.gradient(np.stack([problem.uω0] * 8))
        for key in ("lam0", "grad", "uL"):
            assert all(np.array_equal(res[key][b], res[key][0]) for b in range(8)), key
    print("case: tsit5 step limit", flush=True)
    small = fiberhip.Plan.from_problem(problem, scheme="tsit5", rtol=1e-10, atol=1e-10, max_steps=5)
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EMAXSTEPS"):
        small.forward(problem.uω0)
    from fiberhip import grin

    with pytest.raises(fiberhip.FiberHIPError, match="FH_EUNSUPPORTED"):  # no kernel for M = 5
        fiberhip.Plan.from_problem(grin.grin_problem(5, nt=1024, time_window_ps=10.0), scheme="tsit5")
