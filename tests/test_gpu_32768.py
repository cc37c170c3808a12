"""The 32768-point kernels.

FP32 (32 samples per thread, planar LDS exchange): single-mode fixed-step and
adaptive forwards, the gradient and a 3-mode forward against the oracle, and
bitwise reproducibility of identical solves.

FP64 (split transforms: one radix-2 step inside each thread and two
16384-point block transforms, frequency tables in split order; single mode
runs on the multimode kernels with M = 1): fixed-step and adaptive forwards
and gradients, a 3-mode forward and a 6-mode gradient against the oracle
(field, lam0 and phase gradient within 1e-10 normwise), the FP32 kernel
within 1e-5 of it, and bitwise reproducibility of identical solves."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import mode_energy_rel, oracle_fiber, rel, smf_problem

pytestmark = pytest.mark.gpu
NT, WINDOW = 32768, 94.122


def test_fp32_fixed_step_forward_and_gradient_vs_oracle(gpu_context):
    import fiberhip
    import oracle
    from oracle import np_ref as R

    p = smf_problem(nt=NT, window_ps=WINDOW, length_m=0.3, power_w=0.2)
    fib = oracle_fiber(p)
    print("case: fp32 nt=32768 fixed-step forward", flush=True)
    out, _ = fiberhip.Plan.from_problem(p, steps=60, precision="fp32").forward(p.uω0)
    ref = oracle.rk4_forward(fib, 0.3, 60, p.uω0)
    assert np.all(mode_energy_rel(out, ref) <= 1e-5) and rel(out, ref) <= 1e-4
    mask = fiberhip.frequency_band_mask(p, (-16.0, -10.0))
    print("case: fp32 nt=32768 gradient", flush=True)
    res = fiberhip.Plan.from_problem(p, steps=60, precision="fp32", band_mask=mask, checkpoint_every=20).gradient(p.uω0)
    J, seed = R.band_cost(ref, mask)
    lam0 = oracle.rk4_adjoint(fib, 0.3, 60, p.uω0, seed, every=20)
    print("lam0 rel", rel(res["lam0"], lam0), "cost rel", abs(res["cost"] / J - 1), flush=True)
    assert rel(res["lam0"], lam0) <= 1e-3 and abs(res["cost"] / J - 1) <= 1e-5


def test_fp32_adaptive_forward_vs_oracle(gpu_context):
    import fiberhip
    import oracle

    p = smf_problem(nt=NT, window_ps=WINDOW, length_m=0.3, power_w=0.2)
    print("case: fp32 nt=32768 adaptive forward", flush=True)
    out, stats = fiberhip.Plan.from_problem(p, scheme="tsit5", rtol=1e-6, atol=1e-6, precision="fp32").forward(p.uω0)
    ref = oracle.tsit5_forward(oracle_fiber(p), 0.3, p.uω0, rtol=1e-6, atol=1e-6)
    print("accepted", stats["accepted"], ref.accepted, flush=True)
    assert abs(stats["accepted"] - ref.accepted) <= 2
    assert np.all(mode_energy_rel(out, ref.out) <= 1e-5)


def test_fp32_three_mode_forward_vs_oracle(gpu_context):
    import fiberhip
    import oracle
    from fiberhip import grin

    p = grin.grin_problem(3, nt=NT, time_window_ps=WINDOW, length_m=0.005, energy_j=2e-9, nx=96, window_um=60.0)
    print("case: fp32 M=3 nt=32768 forward", flush=True)
    out, _ = fiberhip.Plan.from_problem(p, steps=20, precision="fp32").forward(p.uω0)
    f = p.fiber
    fib = oracle.Fiber(f["Dω"], f["γ"], f["hRω"], fiberhip.self_steepening(p.sim), f["one_m_fR"])
    ref = oracle.rk4_forward(fib, 0.005, 20, p.uω0)
    assert np.all(mode_energy_rel(out, ref) <= 1e-5)


def test_fp32_identical_members(gpu_context):
    import fiberhip

    p = smf_problem(nt=NT, window_ps=WINDOW, length_m=1.0, power_w=0.2)
    mask = fiberhip.frequency_band_mask(p, (-16.0, -10.0))
    print("case: fp32 nt=32768 identical members", flush=True)
    plan = fiberhip.Plan.from_problem(p, steps=100, precision="fp32", max_batch=8, band_mask=mask)
    res = plan.gradient(np.stack([p.uω0] * 8))
    for key in ("uL", "lam0", "grad"):
        assert all(np.array_equal(res[key][b], res[key][0]) for b in range(8)), key


def test_fp64_split_single_mode_fixed_step_vs_oracle(gpu_context):
    import fiberhip
    import oracle
    from oracle import np_ref as R

    p = smf_problem(nt=NT, window_ps=WINDOW, length_m=0.3, power_w=0.2)
    fib = oracle_fiber(p)
    mask = fiberhip.frequency_band_mask(p, (-16.0, -10.0))
    print("case: fp64 split nt=32768 fixed-step forward", flush=True)
    plan = fiberhip.Plan.from_problem(p, steps=60, precision="fp64", band_mask=mask, checkpoint_every=20,
                                      max_batch=4)
    out, _ = plan.forward(p.uω0)
    ref = oracle.rk4_forward(fib, 0.3, 60, p.uω0)
    print("fp64 split forward rel", rel(out, ref), flush=True)
    assert rel(out, ref) <= 1e-10
    out32, _ = fiberhip.Plan.from_problem(p, steps=60, precision="fp32").forward(p.uω0)
    assert np.all(mode_energy_rel(out32, out) <= 1e-5)
    print("case: fp64 split nt=32768 gradient", flush=True)
    res = plan.gradient(p.uω0)
    J, seed = R.band_cost(ref, mask)
    lam0 = oracle.rk4_adjoint(fib, 0.3, 60, p.uω0, seed, every=20)
    grad = R.phase_gradient(lam0, p.uω0)
    print("lam0 rel", rel(res["lam0"], lam0), "grad rel", rel(res["grad"], grad), "cost rel",
          abs(res["cost"] / J - 1), flush=True)
    assert rel(res["lam0"], lam0) <= 1e-10 and rel(res["grad"], grad) <= 1e-10
    assert abs(res["cost"] - J) <= 1e-12 * J
    print("case: fp64 split nt=32768 identical members", flush=True)
    r4 = plan.gradient(np.stack([p.uω0] * 4))
    for key in ("uL", "lam0", "grad"):
        assert all(np.array_equal(r4[key][b], res[key]) for b in range(4)), key


def test_fp64_split_single_mode_adaptive_vs_oracle(gpu_context):
    import fiberhip
    import oracle
    from oracle import np_ref as R

    p = smf_problem(nt=NT, window_ps=WINDOW, length_m=0.3, power_w=0.2)
    fib = oracle_fiber(p)
    mask = fiberhip.frequency_band_mask(p, (-16.0, -10.0))
    print("case: fp64 split nt=32768 adaptive forward", flush=True)
    plan = fiberhip.Plan.from_problem(p, scheme="tsit5", rtol=1e-8, atol=1e-6, precision="fp64", band_mask=mask,
                                      checkpoint_every=16)
    out, stats = plan.forward(p.uω0)
    ref = oracle.tsit5_forward(fib, 0.3, p.uω0, rtol=1e-8, atol=1e-6)
    print("accepted", stats["accepted"], ref.accepted, "rel", rel(out, ref.out), flush=True)
    assert stats["accepted"] == ref.accepted and rel(out, ref.out) <= 1e-10
    print("case: fp64 split nt=32768 adaptive gradient", flush=True)
    res = plan.gradient(p.uω0)
    hs = plan.step_sizes(1)[0]
    uL = oracle.tsit5_replay(fib, 0.3, hs, p.uω0)
    J, seed = R.band_cost(uL, mask)
    lam0 = oracle.tsit5_adjoint(fib, 0.3, hs, p.uω0, seed, every=50)
    print("steps", hs.size, "lam0 rel", rel(res["lam0"], lam0), flush=True)
    assert rel(res["uL"], uL) <= 1e-10 and rel(res["lam0"], lam0) <= 1e-10
    assert abs(res["cost"] - J) <= 1e-11 * J


def test_fp64_split_multimode_vs_oracle(gpu_context):
    import fiberhip
    import oracle
    from fiberhip import grin
    from oracle import np_ref as R

    p3 = grin.grin_problem(3, nt=NT, time_window_ps=WINDOW, length_m=0.005, energy_j=2e-9, nx=96, window_um=60.0)
    print("case: fp64 split M=3 nt=32768 forward", flush=True)
    out, _ = fiberhip.Plan.from_problem(p3, steps=20, precision="fp64").forward(p3.uω0)
    ref = oracle.rk4_forward(oracle_fiber(p3), 0.005, 20, p3.uω0)
    print("fp64 split M=3 rel", rel(out, ref), flush=True)
    assert rel(out, ref) <= 1e-10 and np.all(mode_energy_rel(out, ref) <= 1e-10)
    p6 = grin.grin_problem(6, nt=NT, time_window_ps=WINDOW, length_m=0.002, energy_j=2e-9, nx=96, window_um=60.0)
    mask = fiberhip.frequency_band_mask(p6, (-16.0, -4.0))
    print("case: fp64 split M=6 nt=32768 gradient", flush=True)
    res = fiberhip.Plan.from_problem(p6, steps=8, precision="fp64", band_mask=mask, checkpoint_every=4).gradient(p6.uω0)
    fib = oracle_fiber(p6)
    uL = oracle.rk4_forward(fib, 0.002, 8, p6.uω0)
    J, seed = R.band_cost(uL, mask)
    lam0 = oracle.rk4_adjoint(fib, 0.002, 8, p6.uω0, seed, every=4)
    print("M=6 lam0 rel", rel(res["lam0"], lam0), "uL rel", rel(res["uL"], uL), flush=True)
    assert rel(res["uL"], uL) <= 1e-10 and rel(res["lam0"], lam0) <= 1e-10
    assert abs(res["cost"] - J) <= 1e-11 * J
    assert rel(res["grad"], R.phase_gradient(lam0, p6.uω0)) <= 1e-10
