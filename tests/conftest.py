"""Shared test setup: import paths, the `gpu` marker, small problem builders."""
from __future__ import annotations

import pathlib
import sys

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
PKG = ROOT / "fiber-raman-suppression_amd"
for p in (str(PKG), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfiberhip on cuda:0)")


@pytest.fixture(scope="session")
def gpu_context():
    import fiberhip

    print("gpu_context: creating fh_context on device 0", flush=True)
    return fiberhip.default_context()


def smf_problem(nt=1024, window_ps=10.0, length_m=0.5, power_w=0.1, preset="SMF28", fR=None,
                beta_order=3, policy="exact"):
    import fiberhip

    fib = fiberhip.Fiber(length_m=length_m, power_w=power_w, preset=preset, beta_order=beta_order,
                         raman_fraction=fR)
    grid = fiberhip.Grid(nt=nt, time_window_ps=window_ps, policy=policy)
    return fiberhip.fiber_problem(fib, grid=grid, raman_threshold_thz=None)


def oracle_fiber(problem):
    import fiberhip
    import oracle

    f = problem.fiber
    return oracle.Fiber(f["Dω"], f["γ"], f["hRω"], fiberhip.self_steepening(problem.sim), f["one_m_fR"])


def rel(a, b) -> float:
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def mode_energy_rel(a, b) -> np.ndarray:
    ea = np.sum(np.abs(np.asarray(a)) ** 2, axis=0)
    eb = np.sum(np.abs(np.asarray(b)) ** 2, axis=0)
    return np.abs(ea - eb) / eb


def cw_gamma(m: int, a: float, b: float) -> np.ndarray:
    """gamma[i,j,k,l] with `a` on (i,i,i,i), `b` on every arrangement of
    (i,i,j,j) for i != j and zero elsewhere (fully permutation-symmetric)."""
    import itertools

    g = np.zeros((m,) * 4)
    for i in range(m):
        g[i, i, i, i] = a
        for j in range(m):
            if j != i:
                for t in set(itertools.permutations((i, i, j, j))):
                    g[t] = b
    return g


def cw_multimode_case(m: int, pattern: str, nt: int = 256, window_ps: float = 3.0, k0: int = 5,
                      phase: float = 2.0, a: float = 1e-2, b: float = 4e-3, fR: float = 0.18,
                      length: float = 0.5):
    """Continuous-wave launch with a closed-form multimode solution.

    Every mode carries power P at the single frequency bin k0, where all modes
    share the dispersion D0 (the other bins get mode-dependent D and hold no
    power).  |u(t)|^2 is then constant in time, so the Raman convolution sees
    only DC, where hR = fR: the Raman and Kerr parts add up to the full
    instantaneous term dK_ij = sum_kl gamma_ijkl Re(a_k conj a_l), and the
    shock factor acts as the scalar s = shock[k0].  With gamma from cw_gamma:
      * "phase": all modes in phase, eta_i = (a + 3 (M-1) b) P a_i;
      * "quad":  the first M/2 modes in phase, the rest a quarter period
        behind (Re(a_i conj a_j) = 0 across the groups),
        eta_i = (a + (2M - 3) b) P a_i.
    Either way every mode keeps its power and turns at the same rate, so the
    lab-frame output is u_i(L) = u_i(0) cis(D0 L + s c P L) with c the
    bracket above.  P is chosen so that s c P L = `phase`.

    Returns (D, gamma, hR, shock, one_m_fR, u0, expected, length)."""
    from oracle import np_ref as R

    if pattern == "quad" and m % 2:
        raise ValueError("the quadrature pattern needs an even mode count")
    sim = R.sim_params(1550e-9, m, nt, window_ps, 3)
    rf = R.raman_fields(sim, fR)
    S = R.shock(sim)
    base = R.dispersion_user(sim, [-2.17e-26, 1.2e-40])[:, 0]
    D = np.stack([base * (1 + 0.05 * k) + 3.0 * k for k in range(m)], axis=1)
    D[k0, :] = base[k0]
    c = a + (3 * (m - 1) if pattern == "phase" else 2 * m - 3) * b
    s = S[k0]
    P = phase / (s * c * length)
    theta = 0.3 + (np.pi / 2) * (np.arange(m) >= m // 2) * (pattern == "quad")
    u0 = np.zeros((nt, m), dtype=np.complex128)
    u0[k0, :] = np.sqrt(P) * np.exp(1j * theta)
    expected = u0 * np.exp(1j * (D[k0, 0] * length + s * c * P * length))
    return D, cw_gamma(m, a, b), rf["hR"], S, rf["one_m_fR"], u0, expected, length
