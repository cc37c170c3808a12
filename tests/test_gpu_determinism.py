"""Bitwise reproducibility of every shipped kernel shape.

A solve's result may depend only on its own inputs: identical launches in one
batch return identical bits, a launch solved alone returns the same bits as
inside a mixed batch, and a repeated call returns the same bits again.  The
matrix covers each (precision, grid size) instantiation of the single-mode
propagation and gradient kernels and the multimode kernel at the largest
FP64 grid, at lengths where the dynamics are strongly nonlinear (Raman
self-frequency shift over 2 m), so that any lane reading a stale value shows
up in the output bits.

The FP64 multimode gradients (fixed step and adaptive, 512 and 1024 threads,
split transforms at 32768 samples, two modes) are checked over 8 identical
members and two launches, and against the oracle's discrete adjoint: built
with the compiler's default memory waits these kernels returned members that
differed by up to 1e-3 in lam0 from launch to launch.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import smf_problem

pytestmark = pytest.mark.gpu

WINDOW_PS = {1024: 10.0, 2048: 10.0, 4096: 12.0, 8192: 23.531, 16384: 47.061}
CASES = [(1024, 0.5, 100), (4096, 0.5, 100), (8192, 2.0, 200), (16384, 2.0, 200)]


def _identical(a: np.ndarray) -> int:
    """Number of batch members whose bits differ from member 0."""
    return sum(not np.array_equal(a[b], a[0]) for b in range(1, a.shape[0]))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("nt,length,steps", CASES)
def test_forward_identical_members(gpu_context, precision, nt, length, steps):
    import fiberhip

    print(f"case: forward {precision} nt={nt} L={length} steps={steps}", flush=True)
    p = smf_problem(nt=nt, window_ps=WINDOW_PS[nt], length_m=length, power_w=0.2)
    plan = fiberhip.Plan.from_problem(p, steps=steps, precision=precision, max_batch=16)
    first = None
    for _ in range(2):
        out, _ = plan.forward(np.stack([p.uω0] * 16))
        assert _identical(out) == 0
        if first is None:
            first = out[0].copy()
        assert np.array_equal(out[0], first)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("nt,length,steps", CASES)
def test_gradient_identical_members(gpu_context, precision, nt, length, steps):
    import fiberhip

    print(f"case: gradient {precision} nt={nt} L={length} steps={steps}", flush=True)
    p = smf_problem(nt=nt, window_ps=WINDOW_PS[nt], length_m=length, power_w=0.2)
    mask = fiberhip.frequency_band_mask(p, (-16.0, -10.0))
    plan = fiberhip.Plan.from_problem(p, steps=steps, precision=precision, max_batch=16, band_mask=mask,
                                      checkpoint_every=50)
    first = None
    for _ in range(2):
        res = plan.gradient(np.stack([p.uω0] * 16))
        for key in ("uL", "lam0", "grad"):
            assert _identical(res[key]) == 0, key
        assert _identical(res["cost"][:, None]) == 0
        if first is None:
            first = res["lam0"][0].copy()
        assert np.array_equal(res["lam0"][0], first)


@pytest.mark.parametrize("nt", [8192, 16384])
def test_fp64_gradient_alone_equals_mixed_batch(gpu_context, nt):
    """Distinct launches (random Taylor phases) solved in one batch and one at a time."""
    import fiberhip

    print(f"case: fp64 gradient alone vs mixed nt={nt}", flush=True)
    p = smf_problem(nt=nt, window_ps=WINDOW_PS[nt], length_m=2.0, power_w=0.2)
    mask = fiberhip.frequency_band_mask(p, (-16.0, -10.0))
    rng = np.random.default_rng(5)
    basis = fiberhip.taylor_phase_basis(p, (2, 3))
    launches = np.stack([p.uω0 * np.exp(1j * (basis @ np.array([rng.uniform(-1e4, 1e4),
                                                                   rng.uniform(-5e5, 5e5)])))[:, None]
                         for _ in range(6)])
    plan = fiberhip.Plan.from_problem(p, steps=200, precision="fp64", max_batch=6, band_mask=mask,
                                      checkpoint_every=50)
    mixed = plan.gradient(launches)
    for b in range(6):
        alone = plan.gradient(launches[b])
        assert np.array_equal(alone["lam0"], mixed["lam0"][b]), b
        assert np.array_equal(alone["grad"], mixed["grad"][b]), b


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("modes,nt", [(6, 4096), (3, 8192)])
def test_multimode_identical_members(gpu_context, precision, modes, nt):
    import fiberhip
    from fiberhip import grin

    print(f"case: multimode {precision} M={modes} nt={nt}", flush=True)
    window = 12.0 if nt == 4096 else 24.0
    p = grin.grin_problem(modes, nt=nt, time_window_ps=window, length_m=0.05)
    plan = fiberhip.Plan.from_problem(p, steps=100, precision=precision, max_batch=16)
    for _ in range(2):
        out, _ = plan.forward(np.stack([p.uω0] * 16))
        assert _identical(out) == 0


FP64_MM_GRADIENTS = [("fixed", 2, 4096), ("fixed", 6, 4096), ("fixed", 3, 8192), ("fixed", 3, 16384),
                     ("fixed", 1, 32768), ("fixed", 3, 32768), ("adaptive", 3, 8192), ("adaptive", 3, 16384),
                     ("adaptive", 1, 32768)]


@pytest.mark.parametrize("scheme,modes,nt", FP64_MM_GRADIENTS)
def test_fp64_multimode_gradient_members_and_oracle(gpu_context, scheme, modes, nt):
    import fiberhip
    import oracle
    from conftest import oracle_fiber, rel
    from fiberhip import grin
    from oracle import np_ref as R

    window = {4096: 12.0, 8192: 23.531, 16384: 47.061, 32768: 94.122}[nt]
    if modes == 1:
        p = smf_problem(nt=nt, window_ps=window, length_m=0.3, power_w=0.2)
        L, steps, every = 0.3, 60, 20
    else:
        p = grin.grin_problem(modes, nt=nt, time_window_ps=window, length_m=0.002, energy_j=2e-9, nx=96,
                              window_um=60.0)
        L, steps, every = 0.002, 8, 4
    mask = fiberhip.frequency_band_mask(p, (-16.0, -10.0))
    print(f"case: fp64 multimode gradient {scheme} M={modes} nt={nt}", flush=True)
    kw = dict(scheme="tsit5", rtol=1e-7, atol=1e-6) if scheme == "adaptive" else dict(steps=steps)
    plan = fiberhip.Plan.from_problem(p, precision="fp64", band_mask=mask, checkpoint_every=every, max_batch=8, **kw)
    first = None
    for _ in range(2):
        res = plan.gradient(np.stack([p.uω0] * 8))
        assert _identical(res["lam0"]) == 0 and _identical(res["grad"]) == 0
        first = res if first is None else first
        assert np.array_equal(res["lam0"], first["lam0"])
    fib = oracle_fiber(p)
    if scheme == "adaptive":
        hs = plan.step_sizes(1)[0]
        uL = oracle.tsit5_replay(fib, L, hs, p.uω0)
        J, seed = R.band_cost(uL, mask)
        lam0 = oracle.tsit5_adjoint(fib, L, hs, p.uω0, seed, every=50)
    else:
        uL = oracle.rk4_forward(fib, L, steps, p.uω0)
        J, seed = R.band_cost(uL, mask)
        lam0 = oracle.rk4_adjoint(fib, L, steps, p.uω0, seed, every=every)
    err = rel(res["lam0"][0], lam0)
    print(f"lam0 rel {err:.2e}", flush=True)
    assert err <= 1e-10 and rel(res["uL"][0], uL) <= 1e-10
