"""GPU parity of dense output (fh_forward_saveat) and of the FP64 field
metrics (fh_field_metrics).

Oracle: oracle/fiberoracle.c orc_tsit5_forward_save (the adaptive solver with
its continuous extension, pinned in tests/test_oracle_pins.py by the
interpolant's order conditions and the CW Kerr / linear closed forms at
interior positions) and oracle/np_ref.py field_metrics (_field_metrics,
propagation.jl:184-202, pinned against closed-form edge fractions).
FP64: saved spectra within 1e-10 of the oracle's, metrics within 1e-12.
FP32-bulk: saved spectral energies within 1e-5 of FP64.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import oracle_fiber, rel, smf_problem

pytestmark = pytest.mark.gpu

ZS = np.array([0.0, 0.0137, 0.1, 0.25, 0.4, 0.5])


def _launches(problem, n, seed):
    import fiberhip

    rng = np.random.default_rng(seed)
    basis = fiberhip.taylor_phase_basis(problem, (2, 3))
    return np.stack([problem.uω0 * np.exp(1j * (basis @ [rng.uniform(-1e4, 1e4), rng.uniform(-5e5, 5e5)]))[:, None]
                     for _ in range(n)])


def _np_sim(problem):
    return {"dt": problem.sim["Δt"], "ws": problem.sim["ωs"]}


KEYS = ("energy_pj", "peak_power_w", "temporal_edge_fraction", "spectral_edge_fraction", "photon_number")


def _check_metrics(dev, ref, rtol, edge_atol):
    """Relative agreement; edge fractions also absolutely (a centred pulse's
    edges hold roundoff-level power, ~1e-30 of the total)."""
    for key in KEYS:
        tol = rtol * abs(ref[key]) + (edge_atol if "edge" in key else 0.0)
        assert abs(dev[key] - ref[key]) <= tol, (key, dev[key], ref[key])
    assert np.allclose(dev["mode_energy_pj"], ref["mode_energy_pj"], rtol=rtol, atol=0)


def test_fp64_saveat_matches_oracle_and_endpoints(gpu_context):
    import fiberhip
    import oracle

    problem = smf_problem(nt=1024, window_ps=10.0, length_m=0.5, power_w=0.1)
    launches = _launches(problem, 3, 7)
    print("case: saveat fp64 nt=1024 batch=3", flush=True)
    plan = fiberhip.Plan.from_problem(problem, scheme="tsit5", rtol=1e-8, atol=1e-6, max_batch=3)
    out, st_save, saved = plan.forward_saveat(launches, ZS)
    assert saved.shape == (3, ZS.size, 1024, 1)
    fib = oracle_fiber(problem)
    for b in range(3):
        ref = oracle.tsit5_forward(fib, 0.5, launches[b], rtol=1e-8, atol=1e-6, saveat=ZS)
        errs = [rel(saved[b, j], ref.saved[j]) for j in range(ZS.size)]
        print("member", b, "max rel", max(errs), flush=True)
        assert max(errs) <= 1e-10
        assert np.array_equal(saved[b, 0], launches[b])
        assert np.array_equal(saved[b, -1], out[b])
        # each member keeps its own step sequence and matches its lone solve
        alone_out, _, alone = plan.forward_saveat(launches[b], ZS)
        assert np.array_equal(alone, saved[b]) and np.array_equal(alone_out, out[b])
    # forward() runs the plain kernel (no dense-output code, bit-identical to
    # the gradient's forward sweep); the dense-output kernel takes the same
    # steps and may differ in the last bits through its own FMA contraction
    plain, st_plain = plan.forward(launches)
    assert st_plain["accepted"] == st_save["accepted"] and st_plain["rejected"] == st_save["rejected"]
    assert max(rel(plain[b], out[b]) for b in range(3)) <= 1e-13


def test_fp32_saveat_within_1e5_of_fp64(gpu_context):
    import fiberhip

    problem = smf_problem(nt=8192, window_ps=23.531, length_m=1.0, power_w=0.1)
    launches = _launches(problem, 4, 9)
    zs = np.linspace(0.0, 1.0, 11)
    sv = {}
    for precision in ("fp64", "fp32"):
        print(f"case: saveat {precision} nt=8192 nz=11", flush=True)
        plan = fiberhip.Plan.from_problem(problem, scheme="tsit5", rtol=1e-6, atol=1e-6, precision=precision,
                                          max_batch=4)
        sv[precision] = plan.forward_saveat(launches, zs)[2]
    e64 = np.sum(np.abs(sv["fp64"]) ** 2, axis=2)
    e32 = np.sum(np.abs(sv["fp32"]) ** 2, axis=2)
    err = np.max(np.abs(e32 / e64 - 1))
    print("fp32 vs fp64 saved energy rel", err, flush=True)
    assert err <= 1e-5


def test_propagate_saveat_metrics_and_verify(gpu_context):
    import fiberhip
    from oracle import np_ref as R

    problem = smf_problem(nt=1024, window_ps=10.0, length_m=0.5, power_w=0.1)
    zs = [0.0, 0.125, 0.25, 0.375, 0.5]
    print("case: propagate saveat + metrics fp64", flush=True)
    res = fiberhip.propagate(problem, saveat=zs)
    assert res.spectra.shape == (1024, 1, 5) and res.solver == "Tsit5"
    sim = _np_sim(problem)
    dev = res.plan.field_metrics(res.saved, problem.sim)
    for j in range(5):
        _check_metrics(dev[j], R.field_metrics(res.saved[j], sim), 1e-12, 1e-15)
    ev = fiberhip.metrics(res)
    assert ev["saved_positions"] == 5 and ev["max_photon_number_drift"] < 1e-6
    assert fiberhip.verify(res)["pass_"]
    with pytest.raises(ValueError, match="sorted"):
        fiberhip.propagate(problem, saveat=[0.0, 0.3, 0.2, 0.5])
    with pytest.raises(ValueError, match="start at 0"):
        fiberhip.propagate(problem, saveat=[0.1, 0.5])
    with pytest.raises(ValueError, match="adaptive"):
        fiberhip.propagate(problem, saveat=zs, steps=200)


def test_field_metrics_multimode_and_fp32(gpu_context):
    """Random spectra, M = 3 (GRIN, fixed-step plan) in FP64 to 1e-12; the
    SMF-28 FP32-bulk plan's metrics within 1e-6 (FP32 transform, FP64 sums)."""
    import fiberhip
    from fiberhip import grin
    from oracle import np_ref as R

    rng = np.random.default_rng(11)
    gp = grin.grin_problem(3, nt=1024, time_window_ps=10.0)
    print("case: field metrics grin3 fp64", flush=True)
    plan = fiberhip.Plan.from_problem(gp, steps=10)
    spectra = rng.normal(size=(4, 1024, 3)) + 1j * rng.normal(size=(4, 1024, 3))
    spectra[1] *= np.exp(-np.linspace(-8, 8, 1024) ** 2)[:, None]  # a band-limited one
    dev = plan.field_metrics(spectra, gp.sim)
    sim = _np_sim(gp)
    for c in range(4):
        _check_metrics(dev[c], R.field_metrics(spectra[c], sim), 1e-12, 1e-15)

    problem = smf_problem(nt=8192, window_ps=23.531, length_m=1.0, power_w=0.1)
    print("case: field metrics smf28 fp32", flush=True)
    p32 = fiberhip.Plan.from_problem(problem, steps=10, precision="fp32")
    sp = _launches(problem, 2, 3)
    dev = p32.field_metrics(sp, problem.sim)
    sim = _np_sim(problem)
    for c in range(2):
        _check_metrics(dev[c], R.field_metrics(sp[c], sim), 1e-6, 1e-9)


def test_saveat_errors(gpu_context):
    import fiberhip

    problem = smf_problem(nt=1024, window_ps=10.0, length_m=0.5, power_w=0.1)
    fixed = fiberhip.Plan.from_problem(problem, steps=50)
    print("case: saveat error paths", flush=True)
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EUNSUPPORTED"):
        fixed.forward_saveat(problem.uω0, ZS)
    ada = fiberhip.Plan.from_problem(problem, scheme="tsit5")
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EINVAL"):
        ada.forward_saveat(problem.uω0, [0.0, 0.3, 0.2])
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EINVAL"):
        ada.forward_saveat(problem.uω0, [0.0, 0.6])
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EINVAL"):
        ada.field_metrics(np.zeros((1024, 1)), problem.sim)
