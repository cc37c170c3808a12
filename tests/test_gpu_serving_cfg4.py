"""The request server on the large classes of the serving trace: GRIN fibers
with M in {6, 10} HG modes at Nt in {8192, 16384, 32768} (carrier-margin
windows), peak powers 1-20 kW, short fibers, forward and forward+adjoint
requests submitted together.

* every served result equals the plan API's solve of the same request bit for
  bit (output, and for gradients the band-objective phase gradient and cost);
* two forward requests solved again by the oracle at their (short) length:
  FP32-bulk within 1e-5 relative spectral energy per mode, and the field
  within 1e-4 (fields move mostly in phase);
* the FP64 server serves the 16384- and 32768-sample multimode classes (the
  latter on split transforms), each result equal to the FP64 plan API's bit
  for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import mode_energy_rel, oracle_fiber, rel

pytestmark = pytest.mark.gpu

WINDOW_PS = {8192: 23.531, 16384: 47.061, 32768: 94.122}
STEP_M = 1e-3


def _problem(cache, nt, m):
    from fiberhip import grin

    if (nt, m) not in cache:
        cache[(nt, m)] = grin.grin_problem(m, nt=nt, time_window_ps=WINDOW_PS[nt], length_m=1.0)
    return cache[(nt, m)]


def _requests(seed=21):
    """(problem, kind, length, launch): uneven classes, lengths and peak powers."""
    import fiberhip

    rng = np.random.default_rng(seed)
    cache, out = {}, []
    classes = [(8192, 6), (8192, 10), (16384, 6), (16384, 10), (32768, 6), (32768, 10)]
    for i in range(12):
        nt, m = classes[i % len(classes)]
        p = _problem(cache, nt, m)
        ut = np.fft.fft(p.uω0, axis=0)
        peak = np.exp(rng.uniform(np.log(1e3), np.log(2e4)))
        basis = fiberhip.taylor_phase_basis(p, (2, 3))
        ph = basis @ np.array([rng.uniform(-1e4, 1e4), rng.uniform(-5e5, 5e5)])
        u0 = p.uω0 * np.sqrt(peak / np.max(np.sum(np.abs(ut) ** 2, axis=1))) * np.exp(1j * ph)[:, None]
        kind = "gradient" if (i // len(classes)) % 2 else "forward"
        out.append((p, kind, float(rng.uniform(0.004, 0.012)), u0))
    return out


def _submit(srv, req):
    import fiberhip

    p, kind, length, u0 = req
    mask = fiberhip.frequency_band_mask(p, (-16.0, -10.0)) if kind == "gradient" else None
    return srv.submit(p, kind=kind, u0=u0, length_m=length, band_mask=mask)


def test_large_classes_match_plan_api_and_oracle(gpu_context):
    import fiberhip
    import oracle

    reqs = _requests()
    print("case: serving large classes fp32 (12 requests, Nt up to 32768, M 6 and 10)", flush=True)
    with fiberhip.Server(lanes=4, max_batch=4, precision="fp32", multimode_step_m=STEP_M,
                         mm_grad_max_steps=64) as srv:
        ids = [_submit(srv, r) for r in reqs]
        got = [srv.wait(i) for i in ids]
    for req, res in zip(reqs, got):
        p, kind, length, u0 = req
        steps = max(16, int(np.ceil(length / STEP_M)))
        assert res["retcode"] == 0 and res["accepted"] == steps
        fib = dict(p.fiber)
        fib["L"] = length
        q = fiberhip.fiber_field_problem(p.uω0, fib, p.sim)
        print(f"case: plan API nt={p.nt} M={p.modes} {kind} steps={steps}", flush=True)
        mask = fiberhip.frequency_band_mask(p, (-16.0, -10.0))
        plan = fiberhip.Plan.from_problem(q, steps=steps, precision="fp32", band_mask=mask, checkpoint_every=8)
        if kind == "gradient":
            ref = plan.gradient(u0)
            assert np.array_equal(ref["uL"], res["uL"])
            assert np.array_equal(ref["grad"], res["grad"]) and ref["cost"] == res["cost"]
        else:
            assert np.array_equal(plan.forward(u0)[0], res["uL"])
        plan.close()
    # two forward requests against the oracle
    for i in (1, 4):
        p, kind, length, u0 = reqs[i]
        assert kind == "forward"
        steps = max(16, int(np.ceil(length / STEP_M)))
        print(f"case: oracle nt={p.nt} M={p.modes} steps={steps}", flush=True)
        ref = oracle.rk4_forward(oracle_fiber(p), length, steps, u0)
        e = mode_energy_rel(got[i]["uL"], ref)
        print(f"fp32 served vs oracle: per-mode energy {e.max():.2e}, field {rel(got[i]['uL'], ref):.2e}", flush=True)
        assert np.all(e <= 1e-5)
        assert rel(got[i]["uL"], ref) <= 1e-4


def test_fp64_server_covers_large_multimode_classes(gpu_context):
    import fiberhip

    reqs = [r for r in _requests() if r[0].nt == 16384][:2] + [r for r in _requests() if r[0].nt == 32768][:2]
    print("case: serving fp64 16384 and 32768 multimode", flush=True)
    with fiberhip.Server(lanes=2, max_batch=2, precision="fp64", multimode_step_m=STEP_M,
                         mm_grad_max_steps=64) as srv:
        ids = [_submit(srv, r) for r in reqs]
        got = [srv.wait(i) for i in ids]
    for req, res in zip(reqs, got):
        p, kind, length, u0 = req
        steps = max(16, int(np.ceil(length / STEP_M)))
        fib = dict(p.fiber)
        fib["L"] = length
        q = fiberhip.fiber_field_problem(p.uω0, fib, p.sim)
        plan = fiberhip.Plan.from_problem(q, steps=steps, precision="fp64",
                                          band_mask=fiberhip.frequency_band_mask(p, (-16.0, -10.0)))
        ref = plan.gradient(u0)["uL"] if kind == "gradient" else plan.forward(u0)[0]
        assert np.array_equal(ref, res["uL"])
