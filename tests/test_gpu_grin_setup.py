"""GRIN modal setup on the GPU: the finite-difference mode solve
(fh_fd_modes, solve_for_fiber_modes of fibers.jl:159-197) and the overlap
tensor (fh_overlap_tensor, compute_overlap_tensor of fibers.jl:214-222)
against the host restatement in fiberhip/grin.py, which
test_host_mirror.py pins to the parabolic profile's Hermite-Gaussian closed
forms.

Bars: eigenvalues within 1e-12 relative of scipy's shift-invert Lanczos;
each degenerate group spans the same subspace (singular values of the
projection within 1e-9 of 1); residual ||A v - beta^2 v|| / beta^2 <= 1e-9
with A assembled independently on the host; repeated solves bitwise equal;
overlap tensor within 1e-13 (relative to its largest entry) of the host sum;
get_params' betas and D within what the 1e-12 eigenvalue agreement allows
through its finite differences, and gamma of the device modes equal to the
host tensor of the same modes.
"""
from __future__ import annotations

import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _operator(eps, wavelength_um, dx):
    import scipy.sparse as sp

    nx, ny = eps.shape
    k2 = (2 * math.pi / wavelength_um) ** 2
    idx = np.arange(nx * ny).reshape(nx, ny, order="F")
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [eps.reshape(-1, order="F") * k2 - 4 / dx ** 2]
    for a, b in ((idx[1:, :], idx[:-1, :]), (idx[:, 1:], idx[:, :-1])):
        for r, c in ((a, b), (b, a)):
            rows.append(r.ravel())
            cols.append(c.ravel())
            vals.append(np.full(r.size, 1 / dx ** 2))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nx * ny,) * 2)


def _groups(d, rel=1e-9):
    out, cur = [], [0]
    for j in range(1, len(d)):
        if abs(d[j] - d[cur[-1]]) <= rel * abs(d[j]):
            cur.append(j)
        else:
            out.append(cur)
            cur = [j]
    return out + [cur]


@pytest.mark.parametrize("nx,nmodes", [(96, 10), (128, 10), (64, 6)])
def test_fd_modes_match_the_host_eigensolver(gpu_context, nx, nmodes):
    from fiberhip import grin

    print(f"case: fd modes nx={nx} M={nmodes}", flush=True)
    eps, _, dx = grin.build_grin(1.55, nx, 80.0, 25.0, 0.2, 2.0)
    t0 = time.perf_counter()
    dh, vh, nh = grin.solve_for_fiber_modes(1.55, nmodes, dx, eps)
    t1 = time.perf_counter()
    dd, vd, nd = grin.solve_for_fiber_modes(1.55, nmodes, dx, eps, backend="device")
    t2 = time.perf_counter()
    print(f"host {1e3 * (t1 - t0):.0f} ms, device {1e3 * (t2 - t1):.0f} ms", flush=True)
    assert np.max(np.abs(dd / dh - 1)) <= 1e-12
    assert np.max(np.abs(nd / nh - 1)) <= 1e-12
    A = _operator(eps, 1.55, dx)
    res = np.linalg.norm(A @ vd - vd * dd, axis=0) / np.abs(dd)
    assert res.max() <= 1e-9
    assert np.allclose(np.linalg.norm(vd, axis=0), 1.0, rtol=0, atol=1e-14)
    for grp in _groups(dh):
        sv = np.linalg.svd(vh[:, grp].T @ vd[:, grp], compute_uv=False)
        assert sv.min() >= 1 - 1e-9, (grp, sv)
    # a non-degenerate mode is fixed up to sign: largest entry positive
    big = np.argmax(np.abs(vd[:, 0]))
    assert vd[big, 0] > 0
    hs = np.sign(vh[np.argmax(np.abs(vh[:, 0])), 0])
    assert np.max(np.abs(vd[:, 0] - hs * vh[:, 0])) <= 1e-9


def test_fd_modes_are_reproducible(gpu_context):
    from fiberhip import grin

    eps, _, dx = grin.build_grin(1.55, 96, 80.0, 25.0, 0.2, 2.0)
    a = grin.solve_for_fiber_modes(1.55, 10, dx, eps, backend="device")
    b = grin.solve_for_fiber_modes(1.55, 10, dx, eps, backend="device")
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("m", [1, 6, 10, 16])
def test_overlap_tensor_matches_the_host_sum(gpu_context, m):
    from fiberhip import grin

    print(f"case: overlap tensor M={m}", flush=True)
    rng = np.random.default_rng(m)
    modes = rng.standard_normal((80, 80, m))
    host = grin.overlap_tensor(modes, 0.5e-6)
    dev = grin.overlap_tensor(modes, 0.5e-6, backend="device")
    assert np.max(np.abs(dev - host)) <= 1e-13 * np.max(np.abs(host))
    # the HG modes: S_0000 = 1 / (2 pi s^2) up to the grid's quadrature
    hg, _, dx = grin.hg_modes(6, 1.55, 25.0, 0.2, 128, 80.0)
    assert np.max(np.abs(grin.overlap_tensor(hg, dx * 1e-6, backend="device") - grin.overlap_tensor(hg, dx * 1e-6))) \
        <= 1e-13 * grin.overlap_tensor(hg, dx * 1e-6)[0, 0, 0, 0]


def test_get_params_on_the_device(gpu_context):
    from fiberhip import grin
    from fiberhip.problem import get_disp_sim_params

    sim = get_disp_sim_params(1550e-9, 6, 1024, 4.0, 2)
    print("case: get_params device nx=101", flush=True)
    host = grin.get_params(sim["f0"], 101, 80.0, 25.0, 0.2, 2.0, 6, 1024, sim["Δt"], 2)
    dev = grin.get_params(sim["f0"], 101, 80.0, 25.0, 0.2, 2.0, 6, 1024, sim["Δt"], 2, backend="device")
    betas_h, D_h, g_h, phi_h, _ = host
    betas_d, D_d, g_d, phi_d, _ = dev
    # D comes from beta(f) by central differences over 1 THz steps: eigenvalues
    # that agree to 1e-12 give beta to 5e-13 relative, which the n-th difference
    # amplifies by sum|w_n| / (2 pi df)^n and D by |w|^n / n! (twice, with the
    # referencing to mode 1)
    beta_abs = 2 * math.pi / 1.55e-6 * grin.silica_index(1.55)
    bounds = [2 * 1e-12 * beta_abs] + [2 * np.abs(grin.central_fd_weights(5, n)).sum() * 1e-12 * beta_abs
                                       / (2 * math.pi * 1e12) ** n for n in (1, 2)]
    for n in range(3):
        assert np.max(np.abs(betas_d[n] - betas_h[n])) <= bounds[n], n
    w_max = math.pi / (sim["Δt"] * 1e-12)
    assert np.max(np.abs(D_d - D_h)) <= sum(w_max ** n / math.factorial(n) * bounds[n] for n in range(3))
    # gamma depends on the basis inside degenerate groups: compare the device
    # tensor with the host sum over the same device modes, and the fundamental
    assert np.max(np.abs(g_d - grin.overlap_tensor(phi_d, 80.0 / 101 * 1e-6) * g_h[0, 0, 0, 0]
                         / grin.overlap_tensor(phi_h, 80.0 / 101 * 1e-6)[0, 0, 0, 0])) <= 1e-12 * g_h[0, 0, 0, 0]
    assert abs(g_d[0, 0, 0, 0] / g_h[0, 0, 0, 0] - 1) <= 1e-9


def test_invalid_arguments_are_rejected(gpu_context):
    import fiberhip
    from fiberhip import grin

    eps, _, dx = grin.build_grin(1.55, 32, 80.0, 25.0, 0.2, 2.0)
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EINVAL"):
        grin.solve_for_fiber_modes(1.55, 41, dx, eps, backend="device")
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EINVAL"):
        grin.solve_for_fiber_modes(1.55, 4, dx, np.full((2, 2), 2.0), backend="device")
    with pytest.raises(fiberhip.FiberHIPError, match="FH_EINVAL"):
        grin.overlap_tensor(np.ones((4, 4, 17)), 1e-6, backend="device")


def test_grin_problem_with_device_setup_propagates_like_the_host_one(gpu_context):
    """grin_problem(mode_solver="fd", setup_backend="device"): the same fiber
    as the host setup up to the basis inside degenerate groups; the fundamental
    mode's launch is basis-independent, so its FP64 propagation agrees."""
    import fiberhip
    from fiberhip import grin

    print("case: grin_problem fd device M=3 nt=512", flush=True)
    kw = dict(nt=512, time_window_ps=4.0, length_m=0.02, energy_j=1e-9, nx=81, window_um=60.0, launch=[1, 0, 0],
              mode_solver="fd")
    ph, pd = grin.grin_problem(3, **kw), grin.grin_problem(3, setup_backend="device", **kw)
    out = []
    for p in (ph, pd):
        plan = fiberhip.Plan.from_problem(p, steps=40, precision="fp64")
        out.append(np.abs(plan.forward(p.uω0)[0][:, 0]) ** 2)
        plan.close()
    assert np.max(np.abs(out[1] - out[0])) <= 1e-9 * np.max(out[0])
