"""Kept tape segments of the gradients (tape_segments).

The forward sweep stores the stage tape of K checkpoint segments (the last K
of a fixed-step solve, the first K of an adaptive one; single and multimode)
and the backward sweep recomputes only the others.  The recomputation repeats
the forward sweep's arithmetic from the checkpoint, so every output (uL,
cost, lam0, grad) must be bit-identical for every K, in both precisions,
including a partial last segment and K at or beyond the segment count.  The
K = 0 path is the one pinned against the oracle's discrete adjoint in
tests/test_gpu_adjoint.py; one kept-tape case is also checked against the
oracle directly here.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import oracle_fiber, rel, smf_problem

pytestmark = pytest.mark.gpu


def _launches(problem, n, seed):
    import fiberhip

    rng = np.random.default_rng(seed)
    basis = fiberhip.taylor_phase_basis(problem, (2, 3))
    phases = [basis @ np.array([rng.uniform(-1e4, 1e4), rng.uniform(-5e5, 5e5)]) for _ in range(n)]
    return np.stack([problem.uω0 * np.exp(1j * p)[:, None] for p in phases])


def _gradients(problem, mask, launches, steps, every, precision, tapes):
    import fiberhip

    out = {}
    for k in tapes:
        print(f"case: tape_segments={k} nt={problem.sim['Nt']} steps={steps} every={every} {precision}", flush=True)
        plan = fiberhip.Plan.from_problem(problem, steps=steps, precision=precision, band_mask=mask,
                                          checkpoint_every=every, max_batch=len(launches), tape_segments=k)
        out[k] = plan.gradient(launches)
        del plan
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_kept_tape_is_bit_identical(gpu_context, precision):
    import fiberhip

    problem = smf_problem(nt=1024, window_ps=10.0, length_m=0.5, power_w=0.2)
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
    launches = _launches(problem, 3, 5)
    # 60 steps in segments of 7: 9 segments, the last one of 4 steps
    res = _gradients(problem, mask, launches, 60, 7, precision, (0, 1, 4, 9, 20))
    for k in (1, 4, 9, 20):
        for key in ("uL", "cost", "lam0", "grad"):
            assert np.array_equal(res[k][key], res[0][key]), (precision, k, key)


def test_kept_tape_fp64_matches_oracle(gpu_context):
    import fiberhip
    import oracle
    from oracle import np_ref as R

    problem = smf_problem(nt=1024, window_ps=10.0, length_m=0.5, power_w=0.2)
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
    launches = _launches(problem, 1, 9)
    steps, every = 60, 7
    res = _gradients(problem, mask, launches, steps, every, "fp64", (3,))[3]
    fib = oracle_fiber(problem)
    uL = oracle.rk4_forward(fib, problem.fiber["L"], steps, launches[0])
    J, seed = R.band_cost(uL, mask)
    lam0 = oracle.rk4_adjoint(fib, problem.fiber["L"], steps, launches[0], seed, every=every)
    assert rel(res["uL"][0], uL) <= 1e-11
    assert abs(res["cost"][0] - J) <= 1e-12 * J
    assert rel(res["lam0"][0], lam0) <= 1e-10
    assert rel(res["grad"][0], R.phase_gradient(lam0, launches[0])) <= 1e-10


def test_kept_tape_headline_kernel_bit_identical(gpu_context):
    """The 16384-sample FP32 kernel of the SMF-28 2 m gradient (1024 threads,
    state in global scratch), 100 steps in 4 segments, 2 of them kept."""
    import fiberhip

    problem = smf_problem(nt=16384, window_ps=47.061, length_m=0.1, power_w=0.1)
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
    launches = _launches(problem, 2, 3)
    res = _gradients(problem, mask, launches, 100, 25, "fp32", (0, 2))
    for key in ("uL", "cost", "lam0", "grad"):
        assert np.array_equal(res[2][key], res[0][key]), key


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adaptive_kept_tape_is_bit_identical(gpu_context, precision):
    """Adaptive single-mode gradient: the first K accepted-step segments keep
    the forward sweep's stage tape (rejected attempts overwrite their slots;
    the seventh stage is stage 0 of the next step).  Launches of different
    chirps take different step counts; segments of 7 accepted steps."""
    import fiberhip

    problem = smf_problem(nt=1024, window_ps=10.0, length_m=1.0, power_w=0.3)
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
    launches = _launches(problem, 3, 21)
    res = {}
    for k in (0, 1, 3, 400):
        print(f"case: adaptive tape_segments={k} nt=1024 every=7 {precision}", flush=True)
        plan = fiberhip.Plan.from_problem(problem, scheme="tsit5", rtol=1e-7, atol=1e-6, precision=precision,
                                          band_mask=mask, checkpoint_every=7, max_steps=2000, max_batch=3,
                                          tape_segments=k)
        res[k] = plan.gradient(launches)
        steps = plan.solve_stats(3)[:, 0]
        del plan
    print("accepted steps", steps, flush=True)
    assert steps.min() > 14  # more than two segments of 7 steps in every solve
    for k in (1, 3, 400):
        for key in ("uL", "cost", "lam0", "grad"):
            assert np.array_equal(res[k][key], res[0][key]), (precision, k, key)


def test_tape_segments_validation(gpu_context):
    import fiberhip
    from fiberhip import grin

    problem = smf_problem(nt=1024, window_ps=10.0, length_m=0.5, power_w=0.2)
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EINVAL"):
        fiberhip.Plan.from_problem(problem, steps=60, tape_segments=-1)
    gp = grin.grin_problem(3, nt=512, time_window_ps=4.0, length_m=0.002, energy_j=1e-9, nx=96, window_um=60.0)
    fiberhip.Plan.from_problem(gp, steps=10, tape_segments=1)  # every gradient accepts it
    fiberhip.Plan.from_problem(gp, scheme="tsit5", tape_segments=1)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_adaptive_multimode_kept_tape_is_bit_identical(gpu_context, precision):
    """Adaptive multimode gradient (GRIN M=3): the first K segments of 7
    accepted steps keep the forward sweep's stage tape."""
    import fiberhip
    from fiberhip import grin

    p = grin.grin_problem(3, nt=512, time_window_ps=4.0, length_m=0.02, energy_j=3e-9, nx=96, window_um=60.0)
    mask = fiberhip.frequency_band_mask(p, (-16.0, -4.0))
    launches = np.stack([grin.grin_problem(3, nt=512, time_window_ps=4.0, length_m=0.02, energy_j=3e-9, nx=96,
                                           window_um=60.0, seed=s).uω0 for s in (1, 2)])
    res = {}
    for k in (0, 1, 400):
        print(f"case: adaptive multimode tape_segments={k} M=3 nt=512 every=7 {precision}", flush=True)
        plan = fiberhip.Plan.from_problem(p, scheme="tsit5", rtol=1e-7, atol=1e-6, precision=precision,
                                          band_mask=mask, checkpoint_every=7, max_steps=2000, max_batch=2,
                                          tape_segments=k)
        res[k] = plan.gradient(launches)
        steps = plan.solve_stats(2)[:, 0]
        del plan
    print("accepted steps", steps, flush=True)
    assert steps.min() > 7  # at least two segments
    for k in (1, 400):
        for key in ("uL", "cost", "lam0", "grad"):
            assert np.array_equal(res[k][key], res[0][key]), (precision, k, key)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_multimode_kept_tape_is_bit_identical(gpu_context, precision):
    """Fixed-step multimode gradient (GRIN M=3): 40 steps in segments of 7
    (the last of 5), the last K segments kept."""
    import fiberhip
    from fiberhip import grin

    p = grin.grin_problem(3, nt=512, time_window_ps=4.0, length_m=0.01, energy_j=2e-9, nx=96, window_um=60.0)
    mask = fiberhip.frequency_band_mask(p, (-16.0, -4.0))
    launches = np.stack([grin.grin_problem(3, nt=512, time_window_ps=4.0, length_m=0.01, energy_j=2e-9, nx=96,
                                           window_um=60.0, seed=s).uω0 for s in (1, 2)])
    res = {}
    for k in (0, 2, 6, 10):
        print(f"case: multimode tape_segments={k} M=3 nt=512 steps=40 every=7 {precision}", flush=True)
        plan = fiberhip.Plan.from_problem(p, steps=40, precision=precision, band_mask=mask, checkpoint_every=7,
                                          max_batch=2, tape_segments=k)
        res[k] = plan.gradient(launches)
        del plan
    for k in (2, 6, 10):
        for key in ("uL", "cost", "lam0", "grad"):
            assert np.array_equal(res[k][key], res[0][key]), (precision, k, key)


@pytest.mark.parametrize("scheme,nt,window,length,power", [("rk4ip", 32768, 94.0, 0.05, 0.1),
                                                          ("tsit5", 16384, 47.061, 0.5, 0.5)])
def test_kept_tape_large_grids_bit_identical(gpu_context, scheme, nt, window, length, power):
    """The 1024-thread FP32 kernels: fixed step at 32768 samples (32 per
    thread, planar exchange; 40 steps in 7 segments) and adaptive at 16384
    samples (about 320 accepted steps with rejections, segments of 6), with
    and without kept segments."""
    import fiberhip

    problem = smf_problem(nt=nt, window_ps=window, length_m=length, power_w=power)
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
    launches = _launches(problem, 2, 13)
    res = {}
    for k in (0, 2):
        print(f"case: {scheme} tape_segments={k} nt={nt} fp32", flush=True)
        kw = {"steps": 40} if scheme == "rk4ip" else {"scheme": "tsit5", "rtol": 1e-8, "atol": 1e-6,
                                                      "max_steps": 2000}
        plan = fiberhip.Plan.from_problem(problem, precision="fp32", band_mask=mask, checkpoint_every=6, max_batch=2,
                                          tape_segments=k, **kw)
        res[k] = plan.gradient(launches)
        del plan
    for key in ("uL", "cost", "lam0", "grad"):
        assert np.array_equal(res[2][key], res[0][key]), (scheme, key)
