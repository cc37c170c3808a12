"""The SMF-28 2 m gradient at its full size (Nt = 16384, 47.061 ps, 0.2 W,
2000 RK4IP steps, checkpoint every 50, band -16..-10 THz, GDD/TOD launch
phase of seed 2), checked three ways:

* against the oracle's discrete adjoint of the same map at full size
  (tests/golden/smf28_2m_gradients.npz, written by make_smf28_2m_gradients.py):
  FP64 within 1e-10 relative (normwise);
* against the reference's continuous adjoint (adjoint_disp_mmf!,
  sensitivity_disp_mmf.jl:32-79, restated in the oracle, integrated backward
  with RK4 at 2000 and 4000 steps and Richardson-extrapolated): the discrete
  adjoint of the 2000-step map differs from the continuous one by the map's
  own truncation (the 2000- and 4000-step discrete gradients differ by 2e-5),
  bound 1e-4 normwise (the reference's own adjoint-vs-FD worst case is
  2.73e-5, counterfactual-raman-benchmark.md:90-91);
* against central differences of the FP64 forward solve (h = 1e-4 rad) on 32
  random bins of the gradient's support (seed 3), all 64 perturbed forwards
  in one batched launch: FP64 gradient within 1e-6 per bin, FP32-bulk
  gradient within 5e-4 normwise over the 32 bins.
The FP32-bulk gradient's normwise error against FP64 over all bins is bounded
by 5e-4 here: its error comes from FP32 rounding inside each right-hand side
(transforms), amplified by the cancellation in 2 Re(conj(lam0) i u0).
"""
from __future__ import annotations

import pathlib
import sys

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def test_smf28_2m_gradient_full_size(gpu_context):
    import fiberhip
    from oracle import np_ref as R

    sys.path.insert(0, str(GOLDEN))
    from make_smf28_2m_gradients import STEPS, setup

    fx = np.load(GOLDEN / "smf28_2m_gradients.npz")
    problem, coef, phase, mask = setup()
    assert np.array_equal(coef, fx["coef"])
    launch = problem.uω0 * np.exp(1j * phase)[:, None]
    res = {}
    for prec in ("fp64", "fp32"):
        print(f"case: SMF-28 2 m gradient {prec} nt=16384 steps={STEPS}", flush=True)
        plan = fiberhip.Plan.from_problem(problem, steps=STEPS, precision=prec, band_mask=mask, checkpoint_every=50,
                                          tape_segments=40)
        res[prec] = plan.gradient(launch)
        plan.close()
    g64, g32 = res["fp64"]["grad"], res["fp32"]["grad"]
    e_disc = rel(g64, fx["grad_discrete"])
    e_cont = rel(g64, fx["grad_continuous"])
    e_32 = rel(g32, g64)
    print(f"fp64 vs oracle discrete {e_disc:.3e}; vs continuous adjoint {e_cont:.3e}; fp32 vs fp64 {e_32:.3e}",
          flush=True)
    assert e_disc <= 1e-10
    assert abs(res["fp64"]["cost"] - float(fx["cost"])) <= 1e-12 * float(fx["cost"])
    assert e_cont <= 1e-4
    assert e_32 <= 5e-4
    assert abs(res["fp32"]["cost"] / res["fp64"]["cost"] - 1) <= 1e-5

    # central differences on 32 bins of the gradient's support
    rng = np.random.default_rng(3)
    support = np.nonzero(np.abs(g64) > 1e-3 * np.abs(g64).max())[0]
    bins = np.sort(rng.choice(support, size=32, replace=False))
    h = 1e-4
    perturbed = []
    for k in bins:
        for sgn in (1, -1):
            ph = phase.copy()
            ph[k] += sgn * h
            perturbed.append(problem.uω0 * np.exp(1j * ph)[:, None])
    print("case: 64 perturbed FP64 forwards in one launch", flush=True)
    fwd = fiberhip.Plan.from_problem(problem, steps=STEPS, precision="fp64", max_batch=64)
    outs, _ = fwd.forward(np.stack(perturbed))
    fd = np.array([(R.band_cost(outs[2 * j], mask)[0] - R.band_cost(outs[2 * j + 1], mask)[0]) / (2 * h)
                   for j in range(len(bins))])
    e64 = np.abs(g64[bins] - fd) / np.abs(fd)
    e32 = rel(g32[bins], fd)
    print(f"FD on 32 bins: fp64 max per-bin rel {e64.max():.3e}; fp32 normwise {e32:.3e}", flush=True)
    assert np.all(e64 <= 1e-6)
    assert e32 <= 5e-4
