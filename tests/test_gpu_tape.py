"""Kept forward solves (fh_forward_tape / fh_adjoint / fh_tape_free) and the
device objectives, through the C ABI.

The reference's adjoint contract reuses the forward solution it cached
(_cached_forward!, physics_models.jl:799-819, 1032-1047).  Checks:
* forward_tape + adjoint(seed) returns the bits of the fused fh_gradient with
  the same seed (output, lam0, phase gradient), for the fixed-step and
  adaptive kernels, single mode and multimode, FP64 and FP32, with and
  without kept tape segments; a second sweep over the same tape repeats them;
* the physical gradients of physics_models.jl:1042-1047 (amplitude, energy)
  equal their definitions from lam0;
* the device objectives (mmf_cost_sum / _fundamental / _worst_mode) give the
  cost and seed of the oracle's restatement (FP64 1e-12), and the fused
  fh_gradient with a non-sum objective equals the tape path;
* a multimode amplitude / energy gradient against central differences.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel, smf_problem

pytestmark = pytest.mark.gpu


def _launches(problem, batch, seed=5):
    import fiberhip

    rng = np.random.default_rng(seed)
    basis = fiberhip.taylor_phase_basis(problem, (2, 3))
    coef = np.stack([rng.uniform(-5e3, 5e3, batch), rng.uniform(-1e5, 1e5, batch)], axis=1)
    return problem.uω0[None] * np.exp(1j * (coef @ basis.T))[:, :, None]


def _grin(m, nt=512):
    from fiberhip import grin

    return grin.grin_problem(m, nt=nt, time_window_ps=max(3.0, 1.2 * nt / (2 * 193.41449)), length_m=0.004,
                             energy_j=2e-9, nx=96, window_um=60.0)


def _seeds(u, rng):
    return (rng.normal(size=u.shape) + 1j * rng.normal(size=u.shape)) * np.abs(u).max()


def _mode_sum(g):
    """sum over modes in the order the fused kernels accumulate (0.0 + m0 + m1 ...)."""
    out = np.zeros(g.shape[:-1])
    for m in range(g.shape[-1]):
        out = out + g[..., m]
    return out


CASES = [
    # name, problem builder, plan kwargs
    ("smf-fixed-fp64", lambda: smf_problem(nt=1024, window_ps=10.0, length_m=0.05), dict(steps=60, tape_segments=0)),
    ("smf-fixed-fp64-kept", lambda: smf_problem(nt=1024, window_ps=10.0, length_m=0.05),
     dict(steps=60, tape_segments=2)),
    ("smf-fixed-fp32-kept", lambda: smf_problem(nt=2048, window_ps=10.0, length_m=0.05),
     dict(steps=60, tape_segments=1, precision="fp32")),
    ("smf-tsit5-fp64-kept", lambda: smf_problem(nt=1024, window_ps=10.0, length_m=0.05),
     dict(scheme="tsit5", rtol=1e-7, atol=1e-7, tape_segments=1)),
    ("smf-tsit5-fp32", lambda: smf_problem(nt=1024, window_ps=10.0, length_m=0.05),
     dict(scheme="tsit5", rtol=1e-6, atol=1e-6, precision="fp32")),
    ("grin3-fixed-fp64-kept", lambda: _grin(3), dict(steps=40, tape_segments=1)),
    ("grin6-fixed-fp32", lambda: _grin(6), dict(steps=40, precision="fp32")),
    ("grin3-tsit5-fp64-kept", lambda: _grin(3), dict(scheme="tsit5", rtol=1e-7, atol=1e-7, tape_segments=1)),
]


@pytest.mark.parametrize("name,build,kw", CASES, ids=[c[0] for c in CASES])
def test_tape_matches_fused_gradient_bitwise(gpu_context, name, build, kw):
    import fiberhip

    problem = build()
    kw = dict(kw)
    prec = kw.pop("precision", "fp64")
    print(f"case: tape {name}", flush=True)
    plan = fiberhip.Plan.from_problem(problem, precision=prec, max_batch=3, checkpoint_every=16, **kw)
    u0 = _launches(problem, 3)
    seeds = _seeds(u0, np.random.default_rng(1))
    fused = plan.gradient(u0, seed=seeds)
    tape = plan.forward_tape(u0)
    assert np.array_equal(tape.output, fused["uL"])
    r1 = tape.adjoint(seeds)
    assert np.array_equal(r1["lam0"], fused["lam0"])
    assert np.array_equal(_mode_sum(r1["grad_phase"]), fused["grad"])
    assert np.all(np.isnan(r1["cost"]))
    # a second sweep (single mode with kept tape rebuilds it first) repeats the bits
    r2 = tape.adjoint(seeds)
    assert np.array_equal(r2["lam0"], r1["lam0"]) and np.array_equal(r2["grad_phase"], r1["grad_phase"])
    # physical gradients from lam0 (physics_models.jl:1042-1047), FP64 on the device
    # (bounds scaled by |lam0| |u0|: the device may fuse the two products)
    lam, u = r1["lam0"], plan.from_device(plan.to_device(u0), 3) if prec == "fp32" else u0
    scale = np.abs(lam) * np.abs(u)
    assert np.all(np.abs(r1["grad_amplitude"] - 2 * np.real(np.conj(lam) * u)) <= 1e-14 * scale)
    assert np.all(np.abs(r1["grad_phase"] - 2 * np.real(np.conj(lam) * 1j * u)) <= 1e-14 * scale)
    assert np.all(np.abs(r1["energy_dot"] - np.real(np.sum(np.conj(lam) * u, axis=(1, 2))))
                  <= 1e-13 * np.sum(scale, axis=(1, 2)))
    tape.free()


@pytest.mark.parametrize("kind", ["band_sum", "fundamental", "worst_mode"])
def test_device_objectives_match_oracle(gpu_context, kind):
    import fiberhip
    from oracle import np_ref as R

    problem = _grin(3)
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -4.0))
    tau = 20.0
    print(f"case: device objective {kind} GRIN M=3 fp64", flush=True)
    plan = fiberhip.Plan.from_problem(problem, steps=40, max_batch=2, band_mask=mask, checkpoint_every=16,
                                      objective=kind, objective_tau=tau)
    u0 = _launches(problem, 2, seed=9)
    tape = plan.forward_tape(u0)
    dev = tape.adjoint(None)
    okind = {"band_sum": "sum"}.get(kind, kind)
    for b in range(2):
        J, seed = R.mmf_costs(tape.output[b], mask, okind, tau)
        assert abs(dev["cost"][b] - J) <= 1e-13 * abs(J)
        host = tape.adjoint(np.stack([seed if c == b else R.mmf_costs(tape.output[c], mask, okind, tau)[1]
                                      for c in range(2)]))
        assert rel(dev["lam0"][b], host["lam0"][b]) <= 1e-12
    # the fused entry point with this objective runs the same sweeps
    fused = plan.gradient(u0)
    assert np.allclose(fused["cost"], dev["cost"], rtol=1e-14, atol=0)
    assert rel(fused["lam0"], dev["lam0"]) <= 1e-13
    tape.free()


def test_amplitude_and_energy_gradients_vs_central_differences(gpu_context):
    """The spectral-shaper model's amplitude and energy gradients
    (physics_models.jl:1043-1046) against central differences of the band cost
    (FP64, GRIN M = 3, fixed step)."""
    import fiberhip

    problem = _grin(3)
    obj = fiberhip.mmf_objective(problem, "worst_mode", band=(-16.0, -4.0), tau=20.0)
    model = fiberhip.fiber_model(problem, steps=40, precision="fp64", checkpoint_every=16, objective=obj)
    rng = np.random.default_rng(3)
    ctrl = {"phase": rng.uniform(-0.3, 0.3, problem.nt), "amplitude": rng.uniform(0.8, 1.2, problem.nt),
            "energy": 1.3}
    print("case: amplitude/energy gradient FD, GRIN M=3 fp64", flush=True)
    final = model.forward(ctrl)
    g = model.physical_gradient(ctrl, obj.terminal_adjoint(final))
    cost = lambda c: obj.cost(model.forward(c))  # noqa: E731
    h = 1e-5
    ep, em = dict(ctrl, energy=1.3 + h), dict(ctrl, energy=1.3 - h)
    fd = (cost(ep) - cost(em)) / (2 * h)
    assert abs(g["energy"][0] - fd) <= 1e-6 * abs(fd) + 1e-14
    for k in (3, problem.nt // 2 + 7, problem.nt - 20):
        ap, am = ctrl["amplitude"].copy(), ctrl["amplitude"].copy()
        ap[k] += h
        am[k] -= h
        fd = (cost(dict(ctrl, amplitude=ap)) - cost(dict(ctrl, amplitude=am))) / (2 * h)
        assert abs(g["amplitude"][k] - fd) <= 1e-5 * abs(fd) + 1e-12
        pp, pm = ctrl["phase"].copy(), ctrl["phase"].copy()
        pp[k] += h
        pm[k] -= h
        fd = (cost(dict(ctrl, phase=pp)) - cost(dict(ctrl, phase=pm))) / (2 * h)
        assert abs(g["phase"][k] - fd) <= 1e-5 * abs(fd) + 1e-12
    model.close()


def test_plan_check_bytes_match_device(gpu_context):
    """fh_plan_check's host-side layout equals what a plan holds, before and
    after its first gradient."""
    import fiberhip

    problem = smf_problem(nt=2048, window_ps=10.0, length_m=0.05)
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -4.0))
    from fiberhip.problem import self_steepening

    f = problem.fiber
    args = (f["Dω"], f["γ"], f["hRω"], self_steepening(problem.sim), f["one_m_fR"], f["L"], 80)
    kw = dict(precision="fp32", max_batch=4, band_mask=mask, checkpoint_every=16, tape_segments=2)
    plan = fiberhip.Plan(*args, **kw)
    assert plan.device_bytes == fiberhip.Plan.check(*args, **kw)
    plan.gradient(_launches(problem, 4))
    assert plan.device_bytes == fiberhip.Plan.check(*args, gradient=True, **kw)
    for case in ((3, "fp64"), (6, "fp32")):
        gp = _grin(case[0])
        g = gp.fiber
        a = (g["Dω"], g["γ"], g["hRω"], self_steepening(gp.sim), g["one_m_fR"], g["L"], 40)
        k = dict(precision=case[1], max_batch=2, band_mask=fiberhip.frequency_band_mask(gp, (-16.0, -4.0)),
                 checkpoint_every=16)
        p = fiberhip.Plan(*a, **k)
        p.gradient(_launches(gp, 2))
        assert p.device_bytes == fiberhip.Plan.check(*a, gradient=True, **k)
