"""GPU tests of the request server (fh_server_*): continuous batching of
uneven requests, each result bitwise equal to the same request solved alone
and to the plan API's solve of the same problem (also with kept stage tape in
the mixed server's gradient plans); memory-budget admission;
unsupported request classes rejected at submission."""
from __future__ import annotations

import threading

import numpy as np
import pytest

from conftest import smf_problem

pytestmark = pytest.mark.gpu


def _requests():
    """A mixed list of (problem, kind, length, launch): single-mode SMF-28 at
    three grid sizes and 3-mode GRIN, each forward and band-objective
    gradient, with uneven lengths and launch phases."""
    import fiberhip
    from fiberhip import grin

    rng = np.random.default_rng(11)
    smf = {nt: smf_problem(nt=nt, window_ps={512: 5.0, 1024: 10.0, 2048: 10.0}[nt], length_m=0.3, power_w=0.1)
           for nt in (512, 1024, 2048)}
    mm = grin.grin_problem(3, nt=512, time_window_ps=4.0, length_m=0.01, energy_j=1e-9, nx=96, window_um=60.0)
    out = []
    for i in range(24):
        if i % 4 == 3:
            kind = "gradient" if i % 8 == 7 else "forward"
            out.append((mm, kind, float(rng.uniform(0.002, 0.01)), mm.uω0 * np.exp(1j * rng.uniform(0, 6.28))))
            continue
        nt = (512, 1024, 2048)[i % 3]
        p = smf[nt]
        basis = fiberhip.taylor_phase_basis(p, (2, 3))
        ph = basis @ np.array([rng.uniform(-1e4, 1e4), rng.uniform(-5e5, 5e5)])
        kind = "gradient" if i % 2 else "forward"
        out.append((p, kind, float(rng.uniform(0.05, 0.4)), p.uω0 * np.exp(1j * ph)[:, None]))
    return out


def _submit(server, req):
    import fiberhip

    p, kind, length, u0 = req
    mask = fiberhip.frequency_band_mask(p, (-16.0, -10.0)) if kind == "gradient" else None
    return server.submit(p, kind=kind, u0=u0, length_m=length, band_mask=mask)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_results_same_alone_and_in_the_mix(gpu_context, precision):
    import fiberhip

    reqs = _requests()
    print(f"case: serving mix {precision}", flush=True)
    # the mixed server keeps the first tape segment of its single-mode gradients,
    # the per-request servers and the plan API recompute: results are bitwise equal
    with fiberhip.Server(lanes=3, max_batch=8, precision=precision, multimode_step_m=2.5e-4,
                         tape_segments=1) as srv:
        ids = [None] * len(reqs)

        def feed(start):
            for i in range(start, len(reqs), 2):
                ids[i] = _submit(srv, reqs[i])

        threads = [threading.Thread(target=feed, args=(s,)) for s in (0, 1)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        mixed = [srv.wait(i) for i in ids]
    assert max(r["batch"] for r in mixed) > 1
    for r in mixed:
        assert r["retcode"] == 0 and r["latency_ms"] >= r["queue_ms"] >= 0
    print("case: serving alone", flush=True)
    for req, got in zip(reqs, mixed):
        with fiberhip.Server(lanes=1, max_batch=1, precision=precision, multimode_step_m=2.5e-4) as srv:
            alone = srv.wait(_submit(srv, req))
        assert alone["batch"] == 1
        assert np.array_equal(alone["uL"], got["uL"])
        if req[1] == "gradient":
            assert np.array_equal(alone["grad"], got["grad"]) and alone["cost"] == got["cost"]
    # the same solves through the plan API
    for req, got in list(zip(reqs, mixed))[:8]:
        p, kind, length, u0 = req
        fib = dict(p.fiber)
        fib["L"] = length
        q = fiberhip.fiber_field_problem(p.uω0, fib, p.sim)
        if p.modes == 1:
            mask = fiberhip.frequency_band_mask(p, (-16.0, -10.0)) if kind == "gradient" else None
            plan = fiberhip.Plan.from_problem(q, scheme="tsit5", rtol=1e-6, atol=1e-6, precision=precision,
                                              band_mask=mask)
            ref = plan.gradient(u0)["uL"] if kind == "gradient" else plan.forward(u0)[0]
        else:
            steps = max(16, int(np.ceil(length / 2.5e-4)))
            assert got["accepted"] == steps
            plan = fiberhip.Plan.from_problem(q, steps=steps, precision=precision,
                                              band_mask=fiberhip.frequency_band_mask(p, (-16.0, -10.0)))
            ref = plan.gradient(u0)["uL"] if kind == "gradient" else plan.forward(u0)[0]
        assert np.array_equal(ref, got["uL"])


def test_memory_budget_admission_and_rejections(gpu_context):
    import fiberhip
    from fiberhip import grin

    reqs = [r for r in _requests() if r[0].modes == 1]
    # the largest plan of the mix (2048-point gradient); the tight budget holds one such plan only
    p0 = max((r[0] for r in reqs), key=lambda p: p.nt)
    one = fiberhip.Plan.from_problem(p0, scheme="tsit5", rtol=1e-6, atol=1e-6, precision="fp32", max_batch=4,
                                     band_mask=fiberhip.frequency_band_mask(p0, (-16.0, -10.0)))
    one.gradient(p0.uω0)
    budget = int(1.3 * one.device_bytes)
    print(f"case: serving budget {budget}", flush=True)
    with fiberhip.Server(lanes=2, max_batch=4, precision="fp32", memory_budget_bytes=8 * budget) as srv:
        got = [srv.wait(_submit(srv, r)) for r in reqs[:6]]
    with fiberhip.Server(lanes=2, max_batch=4, precision="fp32", memory_budget_bytes=budget) as tight:
        ids = [_submit(tight, r) for r in reqs[:6]]
        small = [tight.wait(i) for i in ids]
        for a, b in zip(got, small):
            assert np.array_equal(a["uL"], b["uL"])
        mm = grin.grin_problem(3, nt=512, time_window_ps=4.0, length_m=0.01, energy_j=1e-9, nx=96, window_um=60.0)
        with pytest.raises(fiberhip.FiberHIPError, match="FH_EUNSUPPORTED"):  # beyond the step capacity
            tight.submit(mm, kind="gradient", band_mask=np.ones(512, dtype=bool), length_m=100.0)
        five = grin.grin_problem(5, nt=512, time_window_ps=4.0, length_m=0.01, energy_j=1e-9, nx=96, window_um=60.0)
        with pytest.raises(fiberhip.FiberHIPError, match="FH_EUNSUPPORTED"):  # no kernel for M = 5
            tight.submit(five)
    # a short multimode gradient still runs at least 16 steps: beyond a 10-step capacity
    with fiberhip.Server(lanes=1, max_batch=2, precision="fp32", multimode_step_m=1e-4, mm_grad_max_steps=10) as cap:
        with pytest.raises(fiberhip.FiberHIPError, match="FH_EUNSUPPORTED"):
            cap.submit(mm, kind="gradient", band_mask=np.ones(512, dtype=bool), length_m=5e-4)
    with fiberhip.Server(lanes=1, max_batch=4, precision="fp32", memory_budget_bytes=1 << 20) as starved:
        with pytest.raises(fiberhip.FiberHIPError, match="FH_ENOMEM"):
            starved.wait(_submit(starved, reqs[1]))


def _shared_a_launch(batches) -> bool:
    """Given the launch sizes the good requests report, whether one other
    request rode in one of their launches.  Launches take queued requests in
    order, so a launch of k requests that held only good ones is reported k
    times; one that held the failed request too, k - 1 times."""
    from collections import Counter

    return any(n % k for k, n in Counter(batches).items())


def test_non_finite_request_is_rerun_alone(gpu_context):
    """A multimode request whose FP32 fields overflow (launch x 1e30: |u|^2
    leaves FP32's range in the first Kerr term) fails with FH_ENONFINITE.  If
    it shared a launch it is re-run once, alone, before it is reported
    (attempts 2); the requests launched with it are unaffected -- bitwise equal
    to the same requests solved alone -- and the class is served again after
    the failed launch's plan was released."""
    import fiberhip
    from fiberhip import grin

    mm = grin.grin_problem(3, nt=512, time_window_ps=4.0, length_m=0.01, energy_j=1e-9, nx=96, window_um=60.0)
    launches = [mm.uω0 * np.exp(1j * 0.3), mm.uω0 * 1e30, mm.uω0 * np.exp(1j * 1.1), mm.uω0 * np.exp(1j * 2.0)]
    print("case: serving non-finite rerun fp32", flush=True)
    with fiberhip.Server(lanes=1, max_batch=8, precision="fp32", multimode_step_m=2.5e-4) as srv:
        ids = [srv.submit(mm, kind="forward", u0=u, length_m=0.004) for u in launches]
        good = [srv.wait(ids[i]) for i in (0, 2, 3)]
        with pytest.raises(fiberhip.FiberHIPError, match="FH_ENONFINITE") as err:
            srv.wait(ids[1])
        # the plan of the failed launch was released: the class is still served
        again = srv.wait(srv.submit(mm, kind="forward", u0=launches[0], length_m=0.004))
        assert again["retcode"] == 0 and np.array_equal(again["uL"], good[0]["uL"])
    bad = err.value.result
    assert bad["attempts"] == (2 if _shared_a_launch([g["batch"] for g in good]) else 1)
    assert bad["batch"] == 1
    for g, u in zip(good, [launches[i] for i in (0, 2, 3)]):
        assert g["retcode"] == 0 and g["attempts"] == 1
        with fiberhip.Server(lanes=1, max_batch=1, precision="fp32", multimode_step_m=2.5e-4) as one:
            alone = one.wait(one.submit(mm, kind="forward", u0=u, length_m=0.004))
        assert np.array_equal(alone["uL"], g["uL"])
