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
