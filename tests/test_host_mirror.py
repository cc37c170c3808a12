"""CPU tests of the host side: the FiberLab-surface mirror against the oracle's
restatement of the same helpers, error behaviour, and the C ABI library
(loads, exports every symbol include/fiberhip.h declares).  No GPU calls."""
from __future__ import annotations

import ctypes
import itertools
import math
import re
import subprocess

import numpy as np
import pytest

import fiberhip
from fiberhip import _lib
from oracle import np_ref as R


def test_smf28_8192_grid_resolution():
    """The SMF-28 forward workload asks for Nt=2^13 at 10 ps; the reference's auto grid policy
    resolves it to Nt=8192, T=23.531 ps (physics_helpers.jl:138-185)."""
    fib = fiberhip.Fiber(length_m=0.5, power_w=0.1)
    g = fiberhip.resolve_grid(fib, fiberhip.Pulse(), fiberhip.Grid(nt=8192, time_window_ps=10.0))
    assert g.nt == 8192 and g.time_window_ps == 23.531 and g.policy == "exact"
    # default grid (test_numerical_kernels.jl:300-303)
    assert fiberhip.resolve_sampling_grid(fiberhip.Grid()) == fiberhip.Grid(nt=1024, time_window_ps=10.0,
                                                                             policy="exact")
    with pytest.raises(ValueError, match="use at least 21.178 ps"):
        fiberhip.resolve_sampling_grid(fiberhip.Grid(nt=8192, time_window_ps=12.0, policy="exact"))


@pytest.mark.parametrize("nt,window", [(256, 3.0), (1024, 10.0), (8192, 23.531)])
def test_setup_matches_oracle(nt, window):
    sim = fiberhip.get_disp_sim_params(1550e-9, 1, nt, window, 3)
    osim = R.sim_params(1550e-9, 1, nt, window, 3)
    assert np.array_equal(sim["ts"], osim["ts"]) and np.allclose(sim["ωs"], osim["ws"], rtol=1e-15)
    fib = fiberhip.get_disp_fiber_params_user_defined(0.5, sim, gamma_user=1.1e-3,
                                                      betas_user=[-2.17e-26, 1.2e-40])
    assert np.allclose(fib["Dω"], R.dispersion_user(osim, [-2.17e-26, 1.2e-40]), rtol=1e-15, atol=0)
    assert np.allclose(fib["hRω"], R.raman_fields(osim)["hR"], rtol=1e-15, atol=1e-300)
    assert np.allclose(fiberhip.self_steepening(sim), R.shock(osim), rtol=1e-15)
    _, u = fiberhip.get_initial_state([1.0], 0.1, 185e-15, 80.5e6, "sech_sq", sim)
    _, v = R.initial_state([1.0], 0.1, 185e-15, 80.5e6, "sech_sq", osim)
    assert np.array_equal(u, v)
    # sech^2 peak power 0.881374 P/(fwhm rep) (simulate_disp_mmf.jl:150)
    ut = np.fft.fft(u, axis=0)
    assert math.isclose(np.max(np.abs(ut) ** 2), 0.881374 * 0.1 / 185e-15 / 80.5e6, rel_tol=1e-12)


def test_band_mask_and_taylor_basis():
    sim = fiberhip.get_disp_sim_params(1550e-9, 1, 1024, 10.0, 3)
    m = fiberhip.frequency_band_mask(sim, (-16.0, -10.0))
    f = fiberhip.fftfreq(1024, 1 / sim["Δt"])
    assert m.sum() > 0 and np.all(f[m] >= -16.0) and np.all(f[m] <= -10.0)
    with pytest.raises(ValueError):
        fiberhip.frequency_band_mask(sim, (5.0, 5.0))
    b = fiberhip.taylor_phase_basis(sim, (2, 3, 4))
    osim = R.sim_params(1550e-9, 1, 1024, 10.0, 3)
    assert np.allclose(b, R.taylor_phase_basis(osim, (2, 3, 4)), rtol=1e-15, atol=0)
    for bad in ((), (-1, 2), (2, 2)):
        with pytest.raises(ValueError):
            fiberhip.taylor_phase_basis(sim, bad)


def test_problem_validation_errors():
    """fiber_field_problem mirrors the reference's ArgumentErrors."""
    sim = fiberhip.get_disp_sim_params(1550e-9, 1, 64, 3.0, 2)
    fib = fiberhip.get_disp_fiber_params_user_defined(0.01, sim, gamma_user=1.0, betas_user=[0.0])
    u = np.ones((64, 1), complex)
    fiberhip.fiber_field_problem(u, fib, sim)
    with pytest.raises(ValueError):
        fiberhip.fiber_field_problem(np.ones((32, 1)), fib, sim)
    bad = dict(fib)
    bad["γ"] = np.full((1, 1, 1, 1), 1.0 + 0j)
    with pytest.raises(ValueError, match="real-valued"):
        fiberhip.fiber_field_problem(u, bad, sim)
    with pytest.raises(ValueError):
        fiberhip.raman_fields(sim, fR=1.1)
    with pytest.raises(ValueError):
        fiberhip.raman_fields(sim, tau1=0.0)
    with pytest.raises(ValueError):
        fiberhip.get_disp_fiber_params_user_defined(0.01, sim, gamma_user=1.0)
    with pytest.raises(ValueError):
        fiberhip.get_disp_sim_params(1550e-9, 1, 63, 3.0, 2)
    with pytest.raises(ValueError):
        fiberhip.get_disp_sim_params(1550e-9, 1, 64, 3.0, 1)


def _declared_functions() -> list[str]:
    text = _lib.HEADER_PATH.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(fh_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_header_symbols():
    lib = fiberhip.lib()
    declared = _declared_functions()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_lib.exported_names()) == declared
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(fh_\w+)", out))
    assert set(declared) <= exported


def test_struct_layout_matches_header():
    """ProblemDesc: 6 int32, length_m, 4 pointers, one_m_fr, 2 int32, band_mask
    pointer, rtol, atol, max_steps, tape_segments, objective (+4 padding),
    objective_tau."""
    assert ctypes.sizeof(_lib.ProblemDesc) == 6 * 4 + 8 + 4 * 8 + 8 + 2 * 4 + 8 + 8 + 8 + 2 * 4 + 8 + 8
    assert _lib.ProblemDesc.objective.offset == 112 and _lib.ProblemDesc.objective_tau.offset == 120
    assert _lib.ProblemDesc.band_mask.offset == 80
    assert _lib.ProblemDesc.rtol.offset == 88 and _lib.ProblemDesc.max_steps.offset == 104
    assert _lib.ProblemDesc.tape_segments.offset == 108
    assert _lib.ProblemDesc.length_m.offset == 24
    assert ctypes.sizeof(_lib.Stats) == 32
    assert ctypes.sizeof(_lib.ServerConfig) == 8 + 4 * 4 + 3 * 8 + 2 * 4
    assert ctypes.sizeof(_lib.Request) == 4 * 4 + 8 + 4 * 8 + 8 + 8 + 4 * 8
    assert ctypes.sizeof(_lib.RequestResult) == 6 * 8 + 4 * 4


def test_last_error_starts_empty_and_bad_args_fail_without_gpu():
    lib = fiberhip.lib()
    assert lib.fh_last_error() is not None
    # NULL out-pointer is rejected before any device work
    assert lib.fh_context_create(0, None) == _lib.FH_EINVAL
    assert b"NULL" in lib.fh_last_error()
    assert lib.fh_plan_create(None, None, None) == _lib.FH_EINVAL


def test_grin_builder_invariants():
    """Analytic HG modes: unit norm, orthogonal; S_K fully symmetric and
    S_0000 = 1/(2 pi s^2) for the Gaussian fundamental; D referenced to
    mode 1 (beta0 = beta1 = 0 for mode 1) and equal for degenerate modes."""
    import math

    from fiberhip import grin

    phi, x, dx = grin.hg_modes(6)
    gram = np.einsum("abi,abj->ij", phi, phi)
    assert np.allclose(gram, np.eye(6), atol=1e-10)
    sk = grin.overlap_tensor(phi, dx * 1e-6)
    for perm in itertools.permutations(range(4)):
        assert np.allclose(sk, np.transpose(sk, perm), rtol=1e-12, atol=1e-12 * np.abs(sk).max())
    _, k, g = grin._profile(1.55, 25.0, 0.2)
    s2 = 1 / (k * g) * 1e-12
    assert math.isclose(sk[0, 0, 0, 0], 1 / (2 * math.pi * s2), rel_tol=1e-9)
    # independent restatement of compute_overlap_tensor on a coarse grid
    small = phi[::8, ::8, :3]
    ref = np.zeros((3,) * 4)
    for i, j, kk, l in itertools.product(range(3), repeat=4):
        ref[i, j, kk, l] = np.sum(small[:, :, i] * small[:, :, j] * small[:, :, kk] * small[:, :, l])
    assert np.allclose(grin.overlap_tensor(small, 1.0), ref, rtol=1e-12, atol=1e-15)
    sim = fiberhip.get_disp_sim_params(1550e-9, 6, 1024, 10.0, 2)
    D = grin.grin_dispersion(sim, 6)
    assert D[0, 0] == 0.0
    assert np.allclose(D[:, 1], D[:, 2]) and np.allclose(D[:, 3], D[:, 5])
    w = 2 * np.pi * fiberhip.fftfreq(1024, 1 / sim["Δt"]) * 1e12
    # mode 1 keeps only beta2 w^2 / 2
    beta2 = D[1, 0] / (w[1] ** 2 / 2)
    assert np.allclose(D[:, 0], beta2 * w ** 2 / 2, rtol=1e-9, atol=1e-12)


def test_grin6_launch():
    """10 nJ sech2 100 fs -> P_peak = 0.881374 E / fwhm = 88.137 kW, split by
    seeded Dirichlet weights."""
    from fiberhip import grin

    p = grin.grin_problem(6)
    ut = np.fft.fft(p.uω0, axis=0)
    assert math.isclose(np.max(np.sum(np.abs(ut) ** 2, axis=1)), 0.881374 * 10e-9 / 100e-15, rel_tol=1e-12)
    assert p.uω0.shape == (4096, 6)


def test_bench_kept_segment_budget():
    """bench.py spends one tape budget on both precisions: kept segments per
    gradient = budget / (batch * 4 * every * nt * 3 * element) - 1 (one region
    serves the recomputed segments), clamped to [0, segments]."""
    import importlib.util
    import pathlib

    path = pathlib.Path(__file__).resolve().parents[1] / "bench.py"
    spec = importlib.util.spec_from_file_location("bench_for_test", str(path))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    gib = 1 << 30
    # SMF-28 2 m gradient: 16384 samples, 2000 steps, checkpoint every 50 (40 segments), 256 gradients
    assert bench.kept_segments(16384, 2000, 50, "fp32", 256, 200 * gib) == 20
    assert bench.kept_segments(16384, 2000, 50, "fp64", 256, 200 * gib) == 9
    assert bench.kept_segments(16384, 2000, 50, "fp32", 256, 0) == 0
    assert bench.kept_segments(16384, 2000, 50, "fp32", 1, 200 * gib) == 40
    # the regions fit the budget: (kept + shared) * region bytes <= budget
    for prec, esz in (("fp32", 4), ("fp64", 8)):
        k = bench.kept_segments(16384, 2000, 50, prec, 256, 200 * gib)
        assert (k + 1) * 4 * 50 * 16384 * 3 * esz * 256 <= 200 * gib


def test_adjoint_model_default_tape_budget():
    """AdjointModel keeps as many tape segments as fit in its 4 GiB budget:
    every segment of the SMF-28 2 m gradient (N = 16384, 2000 steps, 40
    segments) in both precisions, and a bounded count for a long solve."""
    from conftest import smf_problem
    from fiberhip.adjoint import AdjointModel

    p = smf_problem(nt=16384, window_ps=47.061, length_m=2.0, power_w=0.2)
    assert AdjointModel.default_tape_segments(p, 2000, "fp64", 50) == 40
    assert AdjointModel.default_tape_segments(p, 2000, "fp32", 50) == 40
    # 20000 steps: 400 segments; 4 GiB holds 54 FP64 regions (78.6 MB) and 109 FP32 ones
    assert AdjointModel.default_tape_segments(p, 20000, "fp64", 50) == 53
    assert AdjointModel.default_tape_segments(p, 20000, "fp32", 50) == 108
    for prec, esz in (("fp64", 8), ("fp32", 4)):
        k = AdjointModel.default_tape_segments(p, 20000, prec, 50)
        assert (k + 1) * 4 * 50 * 16384 * 3 * esz <= AdjointModel.TAPE_BUDGET_BYTES


def _check_args(nt, modes=1, precision="fp64", steps=2000, scheme=None, checkpoint_every=50, max_steps=10000,
                tape_segments=0, band=True, max_batch=1):
    from oracle import np_ref as R

    sim = R.sim_params(1550e-9, modes, nt, 1.2 * nt / (2 * 193.41449), 3)
    rf = R.raman_fields(sim, 0.18)
    D = np.repeat(R.dispersion_user(sim, [-2.17e-26, 1.2e-40]), modes, axis=1)
    g = np.full((modes,) * 4, 1e-3)
    mask = np.zeros(nt, dtype=bool)
    mask[: nt // 8] = True
    return (D, g, rf["hR"], R.shock(sim), rf["one_m_fR"], 1.0, steps), dict(
        precision=precision, scheme=scheme, checkpoint_every=checkpoint_every, max_steps=max_steps,
        tape_segments=tape_segments, band_mask=mask if band else None, max_batch=max_batch)


def test_plan_check_rejects_gradients_beyond_32bit_regions():
    """fh_plan_check (host only): a forward plan is fine, but a gradient whose
    checkpoints or stage tape of one solve exceed 4 GiB is rejected with
    FH_EINVAL instead of addressing past the 32-bit buffer range."""
    from fiberhip.engine import Plan

    # FP64, 16384 samples, 40000 steps checkpointed every step: 40000 * 256 KiB = 10.5 GB per solve
    args, kw = _check_args(16384, steps=40000, checkpoint_every=1)
    assert Plan.check(*args, **kw) > 0
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EINVAL.*checkpoints of one solve exceed 4 GiB"):
        Plan.check(*args, gradient=True, **kw)
    # adaptive FP32 M = 10 at 32768 samples, max_steps 10000, every 5: 2000 checkpoints of 2.6 MB
    args, kw = _check_args(32768, modes=10, precision="fp32", scheme="tsit5", steps=None, checkpoint_every=5,
                           max_steps=10000)
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EINVAL.*checkpoints"):
        Plan.check(*args, gradient=True, **kw)
    # stage tape: each 2000-step region of the FP64 16384 tape holds 2.1 GB of u, seven regions pass 4 GiB
    args, kw = _check_args(16384, steps=20000, checkpoint_every=2000, tape_segments=6)
    with pytest.raises(fiberhip.FiberHIPError, match="stage tape of one solve exceeds 4 GiB"):
        Plan.check(*args, gradient=True, **kw)
    # the SMF-28 2 m gradient plan fits and its adjoint part is the layout's
    args, kw = _check_args(16384, steps=2000, checkpoint_every=50, tape_segments=40, max_batch=4)
    fwd, grad = Plan.check(*args, **kw), Plan.check(*args, gradient=True, **kw)
    ckpt = 40 * 16384 * 16
    tape = 40 * 4 * 50 * 16384 * (16 + 8)
    assert grad - fwd == 4 * (ckpt + tape + 2 * 16384 * 16 + 16384 * 8 + 8)


def test_plan_check_fp64_32768_single_mode_on_split_kernels():
    """One FP64 mode at 32768 samples runs on the multimode kernels with split
    transforms: its layout has their slice (7 fields forward, 13 for the
    gradient, two of them the split scratch) and a complex h tape per stage;
    five modes and 65536 samples have no kernel."""
    from fiberhip.engine import Plan

    nt = 32768
    args, kw = _check_args(nt, steps=100, checkpoint_every=50)
    fwd, grad = Plan.check(*args, **kw), Plan.check(*args, gradient=True, **kw)
    args0, kw0 = _check_args(nt, steps=100, checkpoint_every=50, band=False)
    assert fwd - Plan.check(*args0, **kw0) == nt
    slice_growth = (13 - 7) * 16 * nt
    ckpt = 2 * 16 * nt
    tape = 2 * (4 * 50 * 16 * nt)  # one shared region of u and of h
    assert grad - fwd == slice_growth + ckpt + tape + 2 * 16 * nt + 8 * nt + 8
    args, kw = _check_args(nt, steps=None, scheme="tsit5", max_steps=200)
    assert Plan.check(*args, gradient=True, **kw) > 0
    for bad_nt, bad_m in ((nt, 5), (65536, 1)):
        args, kw = _check_args(bad_nt, modes=bad_m, steps=10)
        with pytest.raises(fiberhip.FiberHIPError, match="FH_EUNSUPPORTED"):
            Plan.check(*args, **kw)


def test_plan_check_validates_like_plan_create():
    """The descriptor checks of fh_plan_create run on the host: unknown
    objectives, a negative temperature and a zero step count are rejected."""
    from fiberhip.engine import Plan

    args, kw = _check_args(1024, steps=100)
    holder = Plan.__new__(Plan)
    desc = holder._describe(*args, **kw)
    n = ctypes.c_int64()
    assert fiberhip.lib().fh_plan_check(ctypes.byref(desc), 0, ctypes.byref(n)) == _lib.FH_OK and n.value > 0
    desc.objective = 7
    assert fiberhip.lib().fh_plan_check(ctypes.byref(desc), 0, ctypes.byref(n)) == _lib.FH_EINVAL
    assert b"unknown objective" in fiberhip.lib().fh_last_error()
    desc.objective = _lib.FH_OBJ_WORST_MODE
    desc.objective_tau = -1.0
    assert fiberhip.lib().fh_plan_check(ctypes.byref(desc), 0, ctypes.byref(n)) == _lib.FH_EINVAL
    with pytest.raises(ValueError):
        Plan.check(*args, **kw, objective="nope")
    args, kw = _check_args(1024, steps=0)
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EINVAL"):
        Plan.check(*args, **kw)


def test_adjoint_model_default_tape_checkpoint_every_zero():
    """checkpoint_every = 0 follows the plan's rule (50, capped at the step count)."""
    from conftest import smf_problem
    from fiberhip.adjoint import AdjointModel

    p = smf_problem(nt=16384, window_ps=47.061, length_m=2.0, power_w=0.2)
    for steps in (2000, 20000, 30):
        for prec in ("fp64", "fp32"):
            assert AdjointModel.default_tape_segments(p, steps, prec, 0) == \
                AdjointModel.default_tape_segments(p, steps, prec, 50)


def test_mmf_costs_match_oracle():
    """Host objectives (fiberhip.adjoint.mmf_cost_*) against the oracle's restatement."""
    from oracle import np_ref as R

    rng = np.random.default_rng(4)
    u = rng.normal(size=(128, 6)) + 1j * rng.normal(size=(128, 6))
    mask = np.zeros(128, dtype=bool)
    mask[90:110] = True
    for kind, fn in (("sum", fiberhip.mmf_cost_sum), ("fundamental", fiberhip.mmf_cost_fundamental),
                     ("worst_mode", fiberhip.mmf_cost_worst_mode)):
        J, s = fn(u, mask)
        Jr, sr = R.mmf_costs(u, mask, kind)
        assert abs(J - Jr) <= 1e-15 * abs(Jr)
        assert np.allclose(s, sr, rtol=1e-13, atol=1e-18)


def test_adjoint_model_control_fields():
    """_single_mode_fields shapes: vector phase repeats over modes, scalars fill,
    non-positive amplitude and energy are rejected."""
    from conftest import smf_problem
    from fiberhip.adjoint import AdjointModel, _field_matrix, _gradient_for_shape

    a, shape = _field_matrix(np.arange(4.0), 4, 3, "phase")
    assert shape == "vector" and a.shape == (4, 3) and np.all(a[:, 2] == np.arange(4.0))
    assert _field_matrix(2.0, 4, 3, "amplitude")[1] == "scalar"
    g = np.arange(12.0).reshape(4, 3)
    assert np.all(_gradient_for_shape(g, "vector") == g.sum(axis=1))
    assert _gradient_for_shape(g, "scalar")[0] == g.sum()
    with pytest.raises(ValueError, match="vector length"):
        _field_matrix(np.zeros(5), 4, 3, "phase")
    m = AdjointModel.__new__(AdjointModel)
    m.problem = smf_problem(nt=256, window_ps=3.0, length_m=0.01)
    with pytest.raises(ValueError, match="strictly positive"):
        m.fields({"amplitude": 0.0})
    with pytest.raises(ValueError, match="energy control"):
        m.fields({"energy": -1.0})
    f = m.fields({"phase": np.zeros(256), "energy": 4.0})
    assert f["alpha"] == 2.0 and f["phase_shape"] == "vector"


def test_fd_grin_modes_match_hermite_gaussian_closed_forms():
    """The finite-difference GRIN setup (build_GRIN / solve_for_fiber_modes /
    get_params, fibers.jl:28-54, 91-132, 159-197) against the parabolic
    profile's closed forms: n_eff of the ten lowest modes within the grid error of
    beta_pq / k (2.5e-5 at 128 points), each degenerate group spanning the Hermite-Gaussian one
    (singular values of the projection > 0.99998), S_0000 and D(w) within the
    grid's O(dx^2) of the analytic builder, and the discretisation error
    shrinking like dx^2 from 101 to 128 points."""
    from fiberhip import grin
    from fiberhip.problem import get_disp_sim_params

    eps, x, dx = grin.build_grin(1.55, 128, 80.0, 25.0, 0.2, 2.0)
    assert eps.shape == (128, 128) and x[64] == 0.0
    _, v, neff = grin.solve_for_fiber_modes(1.55, 10, dx, eps)
    nco, k, g = grin._profile(1.55, 25.0, 0.2)
    hg_neff = np.array([math.sqrt(k * k * nco * nco - 2 * k * g * (p + q + 1)) / k for p, q in grin.MODE_ORDER])
    assert np.all(np.abs(neff - hg_neff) <= 2.5e-5)  # O(dx^2): 1.4e-6 (fundamental) to 1.8e-5 (p + q = 3)
    sim = get_disp_sim_params(1550e-9, 6, 1024, 4.0, 2)
    Dh = grin.grin_dispersion(sim, 6)
    errs = []
    for nx in (101, 128):
        _, D, gamma, phi, _ = grin.get_params(sim["f0"], nx, 80.0, 25.0, 0.2, 2.0, 6, 1024, sim["Δt"], 2)
        hg, _, dxx = grin.hg_modes(6, 1.55, 25.0, 0.2, nx, 80.0)
        f, h = phi.reshape(-1, 6), hg.reshape(-1, 6)
        for grp in ([0], [1, 2], [3, 4, 5]):
            assert np.linalg.svd(f[:, grp].T @ h[:, grp])[1].min() > 0.99998
        s_fd = grin.overlap_tensor(phi, dxx * 1e-6)[0, 0, 0, 0]
        s_hg = grin.overlap_tensor(hg, dxx * 1e-6)[0, 0, 0, 0]
        errs.append((abs(s_fd / s_hg - 1), np.abs(D - Dh).max() / np.abs(Dh).max()))
        # gamma is fully permutation symmetric and D is referenced to mode 1
        assert np.allclose(gamma, np.transpose(gamma, (1, 0, 2, 3)), rtol=1e-13)
        assert np.allclose(gamma, np.transpose(gamma, (0, 2, 1, 3)), rtol=1e-13)
        assert D[0, 0] == 0.0
    assert errs[1][0] < 2e-3 and errs[1][1] < 3e-3
    ratio = (128 / 101) ** 2
    assert errs[0][0] / errs[1][0] > 0.8 * ratio and errs[0][1] / errs[1][1] > 0.8 * ratio


def test_central_fd_weights():
    from fiberhip import grin

    assert np.allclose(grin.central_fd_weights(5, 1), [1 / 12, -2 / 3, 0, 2 / 3, -1 / 12], atol=1e-15)
    assert np.allclose(grin.central_fd_weights(5, 2), [-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12], atol=1e-14)
    assert np.allclose(grin.central_fd_weights(3, 1), [-0.5, 0, 0.5], atol=1e-15)


def test_bench_serving_report(tmp_path):
    """bench.py --serve-report combines the per-request records of both
    precisions: latency percentiles, rates, and the FP32-vs-FP64 mode-energy
    error over the requests both served."""
    import importlib.util
    import pathlib

    path = pathlib.Path(__file__).resolve().parents[1] / "bench.py"
    spec = importlib.util.spec_from_file_location("bench_for_test", str(path))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    n = 8
    e = np.ones((n, 10)) * np.nan
    e[:, :3] = 2.0
    kind = np.array(["forward", "gradient"] * 4)
    common = dict(nt=np.full(n, 1024), modes=np.array([1, 3, 6, 10] * 2), kind=kind, length=np.ones(n), rate=5.0)
    np.savez(tmp_path / "t_fp64.npz", ok=np.ones(n, bool), energy=e, latency_ms=np.arange(n, dtype=float) * 10,
             batch=np.ones(n), wall_s=4.0, **common)
    e32 = e.copy()
    e32[3, 1] = 2.0 * (1 + 3e-6)
    ok32 = np.ones(n, bool)
    ok32[5] = False
    np.savez(tmp_path / "t_fp32.npz", ok=ok32, energy=e32, latency_ms=np.arange(n, dtype=float), batch=np.ones(n) * 2,
             wall_s=1.0, **common)
    r = bench.serving_report(str(tmp_path / "t"))
    assert r["matched_requests"] == 7 and r["fp32"]["served"] == 7
    assert abs(r["fp32_vs_fp64_rel_mode_energy_err"]["max"] - 3e-6) < 1e-12
    assert r["fp32_vs_fp64_rel_mode_energy_err"]["within_1e-5"] == 7
    assert r["fp32"]["gradients_per_s"] == 3.0 and r["fp64"]["gradients_per_s"] == 1.0
    assert r["fp32_over_fp64_gradients_per_s"] == 3.0


def test_grin_setup_backends_are_validated():
    """solve_for_fiber_modes / overlap_tensor / grin_problem accept "host" and
    "device" only (the device path is fh_fd_modes / fh_overlap_tensor, GPU)."""
    from fiberhip import grin

    eps, _, dx = grin.build_grin(1.55, 16, 80.0, 25.0, 0.2, 2.0)
    with pytest.raises(ValueError):
        grin.solve_for_fiber_modes(1.55, 2, dx, eps, backend="cpu")
    with pytest.raises(ValueError):
        grin.overlap_tensor(np.ones((4, 4, 2)), 1e-6, backend="gpu")
    # the host backend of the overlap tensor equals the literal four-fold sum
    rng = np.random.default_rng(4)
    modes = rng.standard_normal((6, 6, 3))
    sk = grin.overlap_tensor(modes, 2e-6)
    lit = np.einsum("xyi,xyj,xyk,xyl->ijkl", modes, modes, modes, modes) / (2e-6) ** 2
    np.testing.assert_allclose(sk, lit, rtol=1e-12)
