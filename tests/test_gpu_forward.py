"""GPU parity of the fused single-mode RK4IP path against the FP64 oracle.

All calls go through the C ABI (libfiberhip.so via ctypes).  The oracle is the
C restatement of _scalar_rk4_output (scalar_crosscheck.jl:8-68).
Tolerances: FP64 path 1e-10 relative (per-mode spectral energy and normwise
field), FP32-bulk path 1e-5 relative spectral energy.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import mode_energy_rel, oracle_fiber, rel, smf_problem

pytestmark = pytest.mark.gpu


def _oracle_out(problem, steps, launch=None):
    import oracle

    u0 = problem.uω0 if launch is None else launch
    return oracle.rk4_forward(oracle_fiber(problem), problem.fiber["L"], steps, u0)


def test_smf28_8192_fp64_matches_oracle(gpu_context):
    """SMF-28, 0.5 m, 0.1 W, 185 fs sech^2, Nt=8192 on the resolved
    23.531 ps window, 200 fixed RK4 steps; FP64 within 1e-10."""
    import fiberhip

    problem = smf_problem(nt=8192, window_ps=10.0, policy="auto_if_undersized")
    assert problem.nt == 8192 and abs(problem.sim["time_window"] - 23.531) < 1e-12
    print("case: smf28-8192 fp64 nt=8192 steps=200", flush=True)
    gpu = fiberhip.scalar_rk4_output(problem, 200, precision="fp64")
    ref = _oracle_out(problem, 200)
    e_rel = mode_energy_rel(gpu, ref)
    f_rel = rel(gpu, ref)
    print(f"smf28-8192 fp64: energy rel {e_rel}, field rel {f_rel:.3e}", flush=True)
    assert np.all(e_rel <= 1e-10)
    assert f_rel <= 1e-10


def test_smf28_8192_fp32_bulk_error(gpu_context):
    """FP32-bulk path on the SMF-28 8192-point forward workload within 1e-5 relative spectral energy."""
    import fiberhip

    problem = smf_problem(nt=8192, window_ps=10.0, policy="auto_if_undersized")
    print("case: smf28-8192 fp32 nt=8192 steps=200", flush=True)
    gpu = fiberhip.scalar_rk4_output(problem, 200, precision="fp32")
    ref = _oracle_out(problem, 200)
    e_rel = mode_energy_rel(gpu, ref)
    f_rel = rel(gpu, ref)
    print(f"smf28-8192 fp32: energy rel {e_rel}, field rel {f_rel:.3e}", flush=True)
    assert np.all(e_rel <= 1e-5)
    assert f_rel <= 1e-4


@pytest.mark.parametrize("nt", [64, 128, 256, 512, 1024, 2048, 4096, 16384])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sizes_match_oracle(gpu_context, nt, precision):
    import fiberhip

    window = max(3.0, 1.2 * nt / (2 * 193.41449))
    problem = smf_problem(nt=nt, window_ps=window, length_m=0.05, power_w=0.2, policy="exact")
    steps = 12
    print(f"case: sizes nt={nt} {precision}", flush=True)
    gpu = fiberhip.scalar_rk4_output(problem, steps, precision=precision)
    ref = _oracle_out(problem, steps)
    tol = 1e-11 if precision == "fp64" else 2e-5
    assert rel(gpu, ref) <= tol
    assert np.all(mode_energy_rel(gpu, ref) <= (1e-11 if precision == "fp64" else 1e-5))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_members_equal_single_solves(gpu_context, precision):
    """Each solve of a batch equals the same solve run alone (bitwise: same kernel)."""
    import fiberhip

    problem = smf_problem(nt=2048, window_ps=8.0, length_m=0.1, power_w=0.1, policy="exact")
    rng = np.random.default_rng(7)
    basis = fiberhip.taylor_phase_basis(problem, (2, 3))
    launches = []
    for _ in range(5):
        phi = basis @ np.array([rng.uniform(-1e4, 1e4), rng.uniform(-5e5, 5e5)])
        launches.append(problem.uω0 * np.exp(1j * phi)[:, None])
    plan = fiberhip.Plan.from_problem(problem, steps=40, precision=precision, max_batch=8)
    print(f"case: batch independence {precision}", flush=True)
    batch_out, stats = plan.forward(np.stack(launches))
    assert stats["rhs_evaluations"] == 5 * 160  # summed over the batch
    assert np.array_equal(plan.solve_stats(5), np.tile([40, 0, 160], (5, 1)))
    for i, l in enumerate(launches):
        alone, _ = plan.forward(l)
        assert np.array_equal(alone, batch_out[i])
    if precision == "fp64":
        for i in (0, 3):
            assert rel(batch_out[i], _oracle_out(problem, 40, launches[i])) <= 1e-11


def test_linear_dispersion_closed_form(gpu_context):
    """gamma = 0: output = cis(D L) u0 (test_scalar_crosscheck.jl:27-37, 2e-15)."""
    import fiberhip

    problem = smf_problem(nt=128, window_ps=4.0, length_m=0.02, power_w=0.02, policy="exact")
    fiber = dict(problem.fiber)
    fiber["γ"] = np.zeros((1, 1, 1, 1))
    lin = fiberhip.fiber_field_problem(problem.uω0, fiber, problem.sim)
    out = fiberhip.scalar_rk4_output(lin, 4)
    expected = np.exp(1j * fiber["Dω"] * fiber["L"]) * lin.uω0
    assert np.max(np.abs(out - expected)) <= 2e-15 + 2e-15 * np.max(np.abs(expected))


def test_cw_kerr_convergence(gpu_context):
    """CW Kerr: u0 cis(gamma P L); RK4 order > 3.5 and final error < 1e-4
    (test_scalar_crosscheck.jl:39-56)."""
    import fiberhip

    sim = fiberhip.get_disp_sim_params(1550e-9, 1, 64, 2.0, 2)
    fib = fiberhip.get_disp_fiber_params_user_defined(2.0, sim, fR=0.0, gamma_user=1.0, betas_user=[0.0])
    launch = np.fft.ifft(np.full((64, 1), np.sqrt(5.0), dtype=complex), axis=0)
    cw = fiberhip.fiber_field_problem(launch, fib, sim)
    exact = launch * np.exp(1j * 1.0 * 5.0 * 2.0)
    errs = [np.linalg.norm(fiberhip.scalar_rk4_output(cw, s) - exact) for s in (32, 64, 128)]
    assert np.log2(errs[0] / errs[1]) > 3.5
    assert np.log2(errs[1] / errs[2]) > 3.5
    assert errs[-1] / np.linalg.norm(exact) < 1e-4


def test_published_counterfactual_shift(gpu_context):
    """Published Delta C = C_on - C_off (counterfactual-raman-benchmark.md:75-79):
    -0.717087 / -0.103957 / -0.095847 THz; 2000 RK4 steps on the GPU FP64 path.
    Tolerance 2.5e-6 THz, inside the reference's 5.95e-6 numerical envelope."""
    import fiberhip

    on = smf_problem(nt=1024, window_ps=10.0, length_m=0.5, power_w=0.1, policy="exact")
    off = fiberhip.with_raman_fraction(on, 0.0)
    basis = fiberhip.taylor_phase_basis(on, (2, 3))
    plan_on = fiberhip.Plan.from_problem(on, steps=2000)
    plan_off = fiberhip.Plan.from_problem(off, steps=2000)
    for phi, published in (((0.0, 0.0), -0.717087), ((-6035.0, 0.0), -0.103957),
                           ((-6555.0, -350000.0), -0.095847)):
        launch = on.uω0 * np.exp(1j * (basis @ np.array(phi)))[:, None]
        c_on = fiberhip.centroid_thz(plan_on.forward(launch)[0], on.sim)
        c_off = fiberhip.centroid_thz(plan_off.forward(launch)[0], on.sim)
        print(f"Delta C {phi}: {c_on - c_off:.7f} (published {published})", flush=True)
        assert abs((c_on - c_off) - published) <= 2.5e-6


def test_energy_device_matches_host(gpu_context):
    """FP64 per-mode energy reduction on device equals the host sum."""
    import fiberhip

    problem = smf_problem(nt=1024, window_ps=10.0, length_m=0.1, policy="exact")
    for precision in ("fp64", "fp32"):
        plan = fiberhip.Plan.from_problem(problem, steps=40, precision=precision, max_batch=4)
        host = np.stack([problem.uω0 * (1.0 - 0.2 * i) for i in range(3)])
        dev = plan.to_device(host)
        e = plan.energy_device(3, dev.ptr)
        cast = host.astype(plan.field_dtype).astype(np.complex128)
        ref = np.sum(np.abs(cast) ** 2, axis=1)
        assert np.allclose(e, ref, rtol=1e-13, atol=0)
        # device forward on resident fields equals the host entry point
        out = fiberhip.DeviceArray(plan.context, dev.nbytes)
        plan.forward_device(3, dev.ptr, out.ptr)
        plan.synchronize()
        dev_out = plan.from_device(out, 3)
        host_out, _ = plan.forward(cast)
        assert np.array_equal(dev_out, host_out)
        ms = plan.time_forward_device(3, dev.ptr, out.ptr, 2)
        assert ms.shape == (2,) and np.all(ms > 0)


def test_error_codes(gpu_context):
    """Fail-closed checks mirror the reference's ArgumentErrors."""
    import fiberhip
    from fiberhip import _lib

    # 10 ps at Nt=8192 reaches negative absolute frequencies (helpers.jl:14-22)
    sim = fiberhip.get_disp_sim_params(1550e-9, 1, 8192, 10.0, 3, validate=False)
    fib = fiberhip.get_disp_fiber_params_user_defined(0.5, sim, gamma_user=1.1e-3, betas_user=[-2.17e-26, 1.2e-40])
    with pytest.raises(fiberhip.FiberHIPError) as err:
        fiberhip.Plan(fib["Dω"], fib["γ"], fib["hRω"], fiberhip.self_steepening(sim), fib["one_m_fR"], 0.5, 10)
    assert err.value.code == _lib.FH_EGRID
    problem = smf_problem(nt=256, window_ps=3.0, length_m=0.05, policy="exact")
    plan = fiberhip.Plan.from_problem(problem, steps=4, max_batch=2)
    bad = problem.uω0.copy()
    bad[3, 0] = np.nan
    with pytest.raises(fiberhip.FiberHIPError) as err:
        plan.forward(bad)
    assert err.value.code == _lib.FH_EINVAL
    with pytest.raises(fiberhip.FiberHIPError) as err:
        plan.forward(np.zeros_like(problem.uω0))
    assert err.value.code == _lib.FH_EINVAL
    with pytest.raises(fiberhip.FiberHIPError) as err:
        plan.forward(np.stack([problem.uω0] * 3))
    assert err.value.code == _lib.FH_EINVAL
    with pytest.raises(fiberhip.FiberHIPError) as err:
        fiberhip.Plan.from_problem(problem, steps=0)
    assert err.value.code == _lib.FH_EINVAL


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_nonfinite_solution_reports_enonfinite(gpu_context, precision):
    """A launch scaled until the Kerr term overflows: the fixed-step forward,
    the adaptive forward (which stops at the first non-finite error estimate)
    and the gradient all return FH_ENONFINITE, the reference's failed-retcode
    error (scalar_crosscheck.jl:66, propagation.jl:48-53)."""
    import fiberhip
    from fiberhip import _lib

    problem = smf_problem(nt=256, window_ps=3.0, length_m=0.05, policy="exact")
    u = problem.uω0 * (1e160 if precision == "fp64" else 1e18)
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
    print(f"case: non-finite forward/gradient {precision}", flush=True)
    for kw in ({"steps": 8}, {"scheme": "tsit5"}):
        plan = fiberhip.Plan.from_problem(problem, precision=precision, **kw)
        with pytest.raises(fiberhip.FiberHIPError) as err:
            plan.forward(u)
        assert err.value.code == _lib.FH_ENONFINITE, kw
    plan = fiberhip.Plan.from_problem(problem, steps=8, precision=precision, band_mask=mask)
    with pytest.raises(fiberhip.FiberHIPError) as err:
        plan.gradient(u)
    assert err.value.code == _lib.FH_ENONFINITE
