"""Writes smf28_2m_gradients.npz: the oracle's phase gradients of the SMF-28
2 m band objective at its full size (Nt = 16384, T = 47.061 ps, 0.2 W, 2000
RK4IP steps, band -16..-10 THz, launch phase GDD/TOD drawn with seed 2).

* grad_discrete: the oracle's discrete adjoint of the fixed-step map
  (oracle/fiberoracle.c: orc_rk4_adjoint, checkpoint every 50);
* grad_continuous_2000 / _4000: the reference's continuous adjoint
  (adjoint_disp_mmf!, sensitivity_disp_mmf.jl:32-79, restated in
  orc_adjoint_rhs) integrated backward with RK4 over 2000 and 4000 steps
  (oracle.continuous_adjoint), the solve_adjoint_disp_mmf path of
  spectral_shaper_model (physics_models.jl:1032-1042), each started from the
  output of the forward trajectory it runs over (at half its step), and
  grad_continuous: their Richardson extrapolation c4000 + (c4000 - c2000)/15
  (RK4's error falls as h^4).
Run from the repository root:  python tests/golden/make_smf28_2m_gradients.py
(about fifteen minutes on one core, 3 GB of memory)."""
from __future__ import annotations

import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "fiber-raman-suppression_amd"), str(ROOT)]

import numpy as np  # noqa: E402

import fiberhip  # noqa: E402
import oracle  # noqa: E402
from oracle import np_ref as R  # noqa: E402

STEPS = 2000


def setup():
    fib = fiberhip.Fiber(length_m=2.0, power_w=0.2, preset="SMF28", beta_order=3)
    problem = fiberhip.fiber_problem(fib, grid=fiberhip.Grid(nt=16384, time_window_ps=47.061, policy="exact"),
                                     raman_threshold_thz=None)
    rng = np.random.default_rng(2)
    coef = np.array([rng.uniform(-1e4, 1e4), rng.uniform(-5e5, 5e5)])
    phase = fiberhip.taylor_phase_basis(problem, (2, 3)) @ coef
    mask = fiberhip.frequency_band_mask(problem, (-16.0, -10.0))
    return problem, coef, phase, mask


def main():
    problem, coef, phase, mask = setup()
    f = problem.fiber
    fib = oracle.Fiber(f["Dω"], f["γ"], f["hRω"], fiberhip.self_steepening(problem.sim), f["one_m_fR"])
    u0 = problem.uω0 * np.exp(1j * phase)[:, None]
    t = time.time()
    uL = oracle.rk4_forward(fib, f["L"], STEPS, u0)
    J, seed = R.band_cost(uL, mask)
    lam_d = oracle.rk4_adjoint(fib, f["L"], STEPS, u0, seed, every=50)
    print(f"discrete adjoint {time.time() - t:.0f} s, J = {J:.12f}", flush=True)
    g_d = R.phase_gradient(lam_d, u0)
    g_c = {}
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    for steps in (STEPS, 2 * STEPS):
        # the continuous adjoint runs backward over a forward trajectory at half
        # its step; its terminal condition comes from that same trajectory
        t = time.time()
        uL_fine = oracle.rk4_forward(fib, f["L"], 2 * steps, u0)
        _, seed_fine = R.band_cost(uL_fine, mask)
        print(f"forward at {2 * steps} steps vs {STEPS}: output field {rel(uL_fine, uL):.3e}", flush=True)
        g_c[steps] = R.phase_gradient(oracle.continuous_adjoint(fib, f["L"], steps, u0, seed_fine), u0)
        print(f"continuous adjoint, {steps} steps: {time.time() - t:.0f} s", flush=True)
    g_x = g_c[2 * STEPS] + (g_c[2 * STEPS] - g_c[STEPS]) / 15
    print(f"c{STEPS} vs c{2 * STEPS}: {rel(g_c[STEPS], g_c[2 * STEPS]):.3e}; extrapolated vs discrete "
          f"{rel(g_x, g_d):.3e}", flush=True)
    np.savez(pathlib.Path(__file__).with_name("smf28_2m_gradients.npz"), coef=coef, cost=J, grad_discrete=g_d,
             grad_continuous=g_x, grad_continuous_2000=g_c[STEPS], grad_continuous_4000=g_c[2 * STEPS])


if __name__ == "__main__":
    main()
