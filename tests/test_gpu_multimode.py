"""GPU parity of the multimode RK4IP kernel against the FP64 oracle.

The oracle (oracle/fiberoracle.c) restates disp_mmf! literally: M^2 complex
Raman transforms and the full gamma[i,j,k,l] contraction, so it also checks
the kernel's pair-symmetric reductions.  Tolerances: FP64 1e-10 relative per
mode (energy) and normwise; FP32 1e-5 relative per-mode spectral energy.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from conftest import mode_energy_rel, oracle_fiber, rel

pytestmark = pytest.mark.gpu


def _oracle(problem, steps, launch=None):
    import oracle

    u0 = problem.uω0 if launch is None else launch
    return oracle.rk4_forward(oracle_fiber(problem), problem.fiber["L"], steps, u0)


def test_grin6_geometry_fp64_matches_oracle(gpu_context):
    """GRIN 50 um / NA 0.2, M=6 HG modes on 256^2, Nt=4096, 12 ps,
    10 nJ sech2 100 fs, Dirichlet launch) at the benchmark step h = 0.075 mm
    over 7.5 mm (100 steps)."""
    import fiberhip
    from fiberhip import grin

    problem = grin.grin_problem(6, length_m=0.0075)
    print("case: grin6 fp64 M=6 nt=4096 steps=100", flush=True)
    gpu, _ = fiberhip.Plan.from_problem(problem, steps=100, precision="fp64").forward(problem.uω0)
    ref = _oracle(problem, 100)
    print("grin6 fp64 energy rel", mode_energy_rel(gpu, ref), "field rel", rel(gpu, ref), flush=True)
    assert np.all(mode_energy_rel(gpu, ref) <= 1e-10)
    assert rel(gpu, ref) <= 1e-10
    gpu32, _ = fiberhip.Plan.from_problem(problem, steps=100, precision="fp32").forward(problem.uω0)
    print("grin6 fp32 energy rel", mode_energy_rel(gpu32, ref), flush=True)
    assert np.all(mode_energy_rel(gpu32, ref) <= 1e-5)


@pytest.mark.parametrize("modes,nt", [(2, 256), (3, 512), (4, 256), (10, 512)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_mode_counts_match_oracle(gpu_context, modes, nt, precision):
    import fiberhip
    from fiberhip import grin

    problem = grin.grin_problem(modes, nt=nt, time_window_ps=max(3.0, 1.2 * nt / (2 * 193.41449)),
                                length_m=0.002, energy_j=2e-9, nx=96, window_um=60.0)
    steps = 20
    print(f"case: modes={modes} nt={nt} {precision}", flush=True)
    gpu, _ = fiberhip.Plan.from_problem(problem, steps=steps, precision=precision).forward(problem.uω0)
    ref = _oracle(problem, steps)
    if precision == "fp64":
        assert rel(gpu, ref) <= 1e-11
        assert np.all(mode_energy_rel(gpu, ref) <= 1e-11)
    else:
        # Kerr coupling mostly moves phase, so the field itself is checked as
        # well: a dropped or misplaced contraction pair changes it by ~1e-2
        # at this nonlinear phase (~0.04 rad), FP32 rounding by ~1e-7.
        print("fp32 field rel", rel(gpu, ref), flush=True)
        assert rel(gpu, ref) <= 1e-5
        assert np.all(mode_energy_rel(gpu, ref) <= 1e-5)


def test_nonsymmetric_gamma_forward(gpu_context):
    """The forward path accepts gamma without (i,j) symmetry (the reference's
    forward RHS does not require it): all M^2 output pairs are formed."""
    import fiberhip
    from fiberhip import grin

    base = grin.grin_problem(3, nt=256, time_window_ps=3.0, length_m=0.002, energy_j=2e-9, nx=96,
                             window_um=60.0)
    rng = np.random.default_rng(5)
    g = base.fiber["γ"] * (1 + 0.3 * rng.uniform(size=(3, 3, 3, 3)))
    fiber = dict(base.fiber)
    fiber["γ"] = g
    problem = fiberhip.fiber_field_problem(base.uω0, fiber, base.sim)
    gpu, _ = fiberhip.Plan.from_problem(problem, steps=20).forward(problem.uω0)
    assert rel(gpu, _oracle(problem, 20)) <= 1e-11


def test_diagonal_gamma_matches_single_mode_kernel(gpu_context):
    """Diagonal gamma decouples the modes: each mode equals the single-mode
    kernel's solve with its own dispersion and gamma."""
    import fiberhip
    from fiberhip import grin

    base = grin.grin_problem(3, nt=1024, time_window_ps=10.0, length_m=0.01, energy_j=1e-9, nx=96,
                             window_um=60.0)
    g = np.zeros((3, 3, 3, 3))
    for k, val in enumerate((1e-3, 2e-3, 3e-3)):
        g[k, k, k, k] = val
    fiber = dict(base.fiber)
    fiber["γ"] = g
    problem = fiberhip.fiber_field_problem(base.uω0, fiber, base.sim)
    multi, _ = fiberhip.Plan.from_problem(problem, steps=50).forward(problem.uω0)
    for k in range(3):
        single = fiberhip.Plan(fiber["Dω"][:, [k]], np.full((1, 1, 1, 1), g[k, k, k, k]), fiber["hRω"],
                               fiberhip.self_steepening(base.sim), fiber["one_m_fR"], fiber["L"], 50)
        out, _ = single.forward(problem.uω0[:, [k]])
        assert rel(multi[:, [k]], out) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_multimode_batch_members_equal_alone(gpu_context, precision):
    import fiberhip
    from fiberhip import grin

    problems = [grin.grin_problem(3, nt=512, time_window_ps=4.0, length_m=0.002, energy_j=1e-9, nx=96,
                                  window_um=60.0, seed=s) for s in (1, 2, 3)]
    plan = fiberhip.Plan.from_problem(problems[0], steps=20, precision=precision, max_batch=3)
    launches = np.stack([p.uω0 for p in problems])
    batch, _ = plan.forward(launches)
    for b in range(3):
        alone, _ = plan.forward(launches[b])
        assert np.array_equal(alone, batch[b])


def test_grin6_full_fp32_vs_fp64(gpu_context):
    """GRIN M=6 at full length (0.3 m, 4000 steps): FP32-bulk vs FP64 per-mode
    spectral energy within 1e-5 (stated per mode)."""
    import fiberhip
    from fiberhip import grin

    problem = grin.grin_problem(6)
    out = {}
    for precision in ("fp64", "fp32"):
        print(f"case: grin6 full {precision} 4000 steps", flush=True)
        out[precision], _ = fiberhip.Plan.from_problem(problem, steps=4000, precision=precision).forward(problem.uω0)
    err = mode_energy_rel(out["fp32"], out["fp64"])
    print("grin6 full fp32 vs fp64 per-mode energy rel err:", err, flush=True)
    assert np.all(err <= 1e-5)


@pytest.mark.parametrize("modes", [3, 6])
def test_fp64_16384_multimode_matches_oracle(gpu_context, modes):
    """FP64 multimode at 16384 samples (the serving trace's largest FP64
    class): field within 1e-11 of the oracle, and four identical launches of
    one batch return identical bits."""
    import fiberhip
    from fiberhip import grin

    problem = grin.grin_problem(modes, nt=16384, time_window_ps=47.061, length_m=0.004)
    steps = 8
    print(f"case: fp64 multimode M={modes} nt=16384 steps={steps}", flush=True)
    plan = fiberhip.Plan.from_problem(problem, steps=steps, precision="fp64", max_batch=4)
    out, _ = plan.forward(np.stack([problem.uω0] * 4))
    ref = _oracle(problem, steps)
    print("fp64 16384 field rel", rel(out[0], ref), flush=True)
    assert rel(out[0], ref) <= 1e-11
    assert np.all(mode_energy_rel(out[0], ref) <= 1e-11)
    for b in range(1, 4):
        assert np.array_equal(out[b], out[0])
