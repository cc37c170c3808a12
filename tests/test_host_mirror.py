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
    pointer, rtol, atol, max_steps, tape_segments."""
    assert ctypes.sizeof(_lib.ProblemDesc) == 6 * 4 + 8 + 4 * 8 + 8 + 2 * 4 + 8 + 8 + 8 + 2 * 4
    assert _lib.ProblemDesc.band_mask.offset == 80
    assert _lib.ProblemDesc.rtol.offset == 88 and _lib.ProblemDesc.max_steps.offset == 104
    assert _lib.ProblemDesc.tape_segments.offset == 108
    assert _lib.ProblemDesc.length_m.offset == 24
    assert ctypes.sizeof(_lib.Stats) == 32
    assert ctypes.sizeof(_lib.ServerConfig) == 8 + 4 * 4 + 3 * 8 + 2 * 4
    assert ctypes.sizeof(_lib.Request) == 4 * 4 + 8 + 4 * 8 + 8 + 8 + 4 * 8
    assert ctypes.sizeof(_lib.RequestResult) == 6 * 8 + 2 * 4


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
