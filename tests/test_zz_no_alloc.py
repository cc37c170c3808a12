"""fh_gradient_device_noalloc against fh_gradient_device, through the C ABI.

The allocation-free entry point keeps its buffers in the library's fixed
per-GPU scratch array; its results must equal the allocating entry point's
bit for bit (same kernels, same arithmetic).  Two calls in a row, the second
with a smaller batch (the array keeps the first call's contents, so a stale
status word or slice would show), over the fixed-step and adaptive kernels,
single mode and multimode (including a device objective that needs a seed
buffer), FP64 and FP32; and a call whose buffers would not fit is rejected.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import smf_problem

pytestmark = pytest.mark.gpu


def _launches(problem, batch, seed):
    import fiberhip

    rng = np.random.default_rng(seed)
    basis = fiberhip.taylor_phase_basis(problem, (2, 3))
    coef = np.stack([rng.uniform(-5e3, 5e3, batch), rng.uniform(-1e5, 1e5, batch)], axis=1)
    return problem.uω0[None] * np.exp(1j * (coef @ basis.T))[:, :, None]


def _grin(m, nt=512):
    from fiberhip import grin

    return grin.grin_problem(m, nt=nt, time_window_ps=max(3.0, 1.2 * nt / (2 * 193.41449)), length_m=0.004,
                             energy_j=2e-9, nx=96, window_um=60.0)


def _run(plan, u0, noalloc, seed=None):
    import fiberhip

    batch = u0.shape[0]
    d_u0 = plan.to_device(u0)
    d_seed = plan.to_device(seed) if seed is not None else None
    d_uL = fiberhip.DeviceArray(plan.context, d_u0.nbytes)
    d_lam = fiberhip.DeviceArray(plan.context, d_u0.nbytes)
    d_grad = fiberhip.DeviceArray(plan.context, 8 * batch * plan.nt)
    d_cost = fiberhip.DeviceArray(plan.context, 8 * batch)
    plan.gradient_device(batch, d_u0.ptr, d_seed.ptr if d_seed else 0, d_uL.ptr, d_lam.ptr, d_grad.ptr, d_cost.ptr,
                         noalloc=noalloc)
    plan.synchronize()
    return dict(uL=plan.from_device(d_uL, batch), lam0=plan.from_device(d_lam, batch),
                grad=d_grad.to_host(np.float64, (batch, plan.nt)), cost=d_cost.to_host(np.float64, (batch,)))


CASES = [
    ("smf-fixed-fp64", lambda: smf_problem(nt=1024, window_ps=10.0, length_m=0.05), dict(steps=60), False),
    ("smf-fixed-fp32-kept", lambda: smf_problem(nt=2048, window_ps=10.0, length_m=0.05),
     dict(steps=60, tape_segments=2, precision="fp32"), False),
    ("smf-tsit5-fp64", lambda: smf_problem(nt=1024, window_ps=10.0, length_m=0.05),
     dict(scheme="tsit5", rtol=1e-7, atol=1e-7, max_steps=400), True),
    ("grin3-fixed-fp64-worst", lambda: _grin(3), dict(steps=40, objective="worst_mode", objective_tau=20.0), False),
    ("grin6-tsit5-fp32", lambda: _grin(6), dict(scheme="tsit5", rtol=1e-6, atol=1e-6, precision="fp32",
                                                max_steps=400), False),
]


@pytest.mark.parametrize("name,build,kw,with_seed", CASES, ids=[c[0] for c in CASES])
def test_noalloc_gradient_matches_allocating_entry_point(gpu_context, name, build, kw, with_seed):
    import fiberhip

    problem = build()
    kw = dict(kw)
    prec = kw.pop("precision", "fp64")
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -4.0))
    plan = fiberhip.Plan.from_problem(problem, precision=prec, max_batch=4, band_mask=mask, checkpoint_every=16,
                                      **kw)
    for call, batch in enumerate((4, 2)):
        print(f"case: noalloc {name} call {call} batch {batch}", flush=True)
        u0 = _launches(problem, batch, seed=10 + call)
        seed = None
        if with_seed:
            rng = np.random.default_rng(call)
            seed = (rng.normal(size=u0.shape) + 1j * rng.normal(size=u0.shape)) * np.abs(u0).max()
        a = _run(plan, u0, noalloc=True, seed=seed)
        b = _run(plan, u0, noalloc=False, seed=seed)
        for k in ("uL", "lam0", "grad"):
            assert np.array_equal(a[k], b[k]), k
        if seed is None:
            assert np.array_equal(a["cost"], b["cost"])
        assert np.all(np.isfinite(a["grad"]))


def test_noalloc_rejects_beyond_scratch(gpu_context):
    """The SMF-28 2 m gradient with every segment kept needs 3.2 GB per
    solve: far above the fixed array, so the call fails before any launch."""
    import fiberhip

    problem = smf_problem(nt=16384, window_ps=47.061, length_m=2.0, power_w=0.2)
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
    plan = fiberhip.Plan.from_problem(problem, steps=2000, precision="fp64", max_batch=1, band_mask=mask,
                                      tape_segments=40)
    d = fiberhip.DeviceArray(plan.context, 16 * problem.nt)
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EINVAL.*bytes of scratch"):
        plan.gradient_device(1, d.ptr, 0, d.ptr, d.ptr, d.ptr, d.ptr, noalloc=True)
