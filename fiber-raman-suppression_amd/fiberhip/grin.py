"""Parabolic GRIN multimode fibers: analytic Hermite-Gaussian modes, or the
reference's finite-difference mode solve (end of this file).

Builds the operators that get_disp_fiber_params / get_params produce
(fibers.jl:91-132): the Kerr tensor gamma = S_K n2 w0 / c from the overlap
integral of unit-norm sampled modes (compute_overlap_tensor, fibers.jl:214-222)
and the dispersion D(w) from finite differences of beta(f) on a symmetric
stencil, referenced to mode 1 (fibers.jl:98-118).

The reference finds its modes with a 5-point finite-difference eigensolve
(Arpack, fibers.jl:159-197).  Here the modes of the parabolic profile are the
harmonic-oscillator (Hermite-Gaussian) solutions of the same scalar wave
equation with eps(r) ~ n_co^2 - g^2 r^2, g^2 = 2 n_co (n_co - n_cl) / a^2:
    psi_pq ~ H_p(x/s) H_q(y/s) exp(-(x^2+y^2)/(2 s^2)),  s^2 = 1/(k g)
    beta_pq^2 = k^2 n_co^2 - 2 k g (p + q + 1).
"""
from __future__ import annotations

import math

import numpy as np

from .problem import (
    FiberFieldProblem,
    fftfreq,
    fiber_field_problem,
    get_disp_sim_params,
    get_initial_state,
    raman_fields,
)

N2_SILICA = 2.3e-20  # m^2/W (fibers.jl:125)
MODE_ORDER = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3)]
# 5-point central finite-difference weights for the first and second derivative
FD5 = {1: np.array([1 / 12, -2 / 3, 0.0, 2 / 3, -1 / 12]),
       2: np.array([-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12])}


def silica_index(wavelength_um: float) -> float:
    """Sellmeier index of fused silica (coefficients of build_GRIN, fibers.jl:30-38)."""
    l2 = wavelength_um ** 2
    return math.sqrt(1 + 0.6961663 * l2 / (l2 - 0.0684043 ** 2) + 0.4079426 * l2 / (l2 - 0.1162414 ** 2)
                     + 0.8974794 * l2 / (l2 - 9.896161 ** 2))


def _profile(wavelength_um: float, radius_um: float, core_na: float):
    nco = silica_index(wavelength_um)
    ncl = math.sqrt(nco ** 2 - core_na ** 2)
    k = 2 * math.pi / wavelength_um
    g = math.sqrt(2 * nco * (nco - ncl)) / radius_um
    return nco, k, g


def hermite(n: int, x: np.ndarray) -> np.ndarray:
    """Physicists' Hermite polynomial by recurrence."""
    h0, h1 = np.ones_like(x), 2 * x
    if n == 0:
        return h0
    for k in range(1, n):
        h0, h1 = h1, 2 * x * h1 - 2 * k * h0
    return h1


def hg_modes(modes: int, wavelength_um=1.55, radius_um=25.0, core_na=0.2, nx=256, window_um=80.0):
    """Unit-norm sampled modes (nx, nx, M) on the grid x = (-nx/2 .. nx/2-1) dx (fibers.jl:45-47)."""
    _, k, g = _profile(wavelength_um, radius_um, core_na)
    s = 1 / math.sqrt(k * g)
    dx = window_um / nx
    x = np.arange(-nx / 2, nx / 2) * dx
    X, Y = np.meshgrid(x, x, indexing="ij")
    out = np.empty((nx, nx, modes))
    for i, (p, q) in enumerate(MODE_ORDER[:modes]):
        psi = hermite(p, X / s) * hermite(q, Y / s) * np.exp(-(X ** 2 + Y ** 2) / (2 * s * s))
        out[:, :, i] = psi / math.sqrt(np.sum(psi ** 2))
    return out, x, dx


def overlap_tensor(modes: np.ndarray, dx_si: float, backend: str = "host") -> np.ndarray:
    """S_K[i,j,k,l] = sum_mn phi_i phi_j phi_k phi_l / dx^2 (fibers.jl:214-222).
    backend "device": FP64 on the GPU (fh_overlap_tensor)."""
    nx = modes.shape[0]
    f = modes.reshape(nx * nx, -1)
    m = f.shape[1]
    if backend == "device":
        return _overlap_tensor_device(f, 1.0 / dx_si ** 2)
    if backend != "host":
        raise ValueError('backend must be "host" or "device"')
    pair = np.einsum("ri,rj->rij", f, f).reshape(nx * nx, m * m)
    return (pair.T @ pair).reshape(m, m, m, m) / dx_si ** 2


def beta_pq(f_thz: float, p: int, q: int, radius_um=25.0, core_na=0.2) -> float:
    """Propagation constant [1/m] of HG mode (p, q) at frequency f."""
    lam_um = 2.99792458e8 / (f_thz * 1e12) * 1e6
    nco, k, g = _profile(lam_um, radius_um, core_na)
    return math.sqrt(k * k * nco * nco - 2 * k * g * (p + q + 1)) * 1e6


def grin_dispersion(sim: dict, modes: int, beta_order=2, delta_f_thz=1.0, radius_um=25.0, core_na=0.2):
    """D(w) (nt, M): beta(f) on a 2*beta_order+1 point stencil, central finite
    differences, beta0/beta1 referenced to mode 1 (fibers.jl:92-118)."""
    if beta_order != 2:
        raise ValueError("the analytic GRIN builder uses beta_order = 2 (5-point stencil)")
    f0 = sim["f0"]
    offsets = np.arange(-2, 3)
    bf = np.array([[beta_pq(f0 + delta_f_thz * o, p, q, radius_um, core_na) for (p, q) in MODE_ORDER[:modes]]
                   for o in offsets])  # (5, M)
    d = {n: FD5[n] @ bf / (2 * math.pi * delta_f_thz / 1e-12) ** n for n in (1, 2)}
    betas = np.stack([bf[2] - bf[2, 0], d[1] - d[1][0], d[2]], axis=0)  # (3, M)
    w = 2 * math.pi * fftfreq(int(sim["Nt"]), 1 / sim["Δt"]) * 1e12
    cols = np.stack([w ** n / math.factorial(n) for n in range(3)], axis=1)
    return cols @ betas


def grin_problem(modes: int = 6, nt: int = 4096, time_window_ps: float = 12.0, length_m: float = 0.3,
                 energy_j: float = 10e-9, fwhm_s: float = 100e-15, seed: int = 1, nx: int = 256,
                 window_um: float = 80.0, radius_um=25.0, core_na=0.2, wavelength_m=1550e-9, fR=0.18,
                 launch=None, mode_solver: str = "hg", setup_backend: str = "host") -> FiberFieldProblem:
    """Default problem: parabolic GRIN (50 um core, NA 0.2), M modes on an
    nx^2 grid, sech^2 pulse of energy_j split over modes by Dirichlet(1)
    weights with uniform phases (seeded), initial_modes normalised
    (physics_models.jl:289-297).  mode_solver "hg" uses the analytic
    Hermite-Gaussian modes and beta_pq; "fd" the reference's procedure
    (get_params: finite-difference eigensolve of the index profile).
    setup_backend "device" runs the mode solves and the overlap tensor on the
    GPU (fh_fd_modes, fh_overlap_tensor)."""
    sim = get_disp_sim_params(wavelength_m, modes, nt, time_window_ps, 2)
    w0 = 2 * math.pi * sim["f0"] * 1e12
    if mode_solver == "fd":
        _, D, gamma, phi, _ = get_params(sim["f0"], nx, window_um, radius_um, core_na, 2.0, modes, nt, sim["Δt"], 2,
                                         backend=setup_backend)
    elif mode_solver == "hg":
        phi, _, dx = hg_modes(modes, wavelength_m * 1e6, radius_um, core_na, nx, window_um)
        sk = overlap_tensor(phi, dx * 1e-6, setup_backend)
        gamma = sk * N2_SILICA * w0 / 2.99792458e8
        D = grin_dispersion(sim, modes, radius_um=radius_um, core_na=core_na)
    else:
        raise ValueError('mode_solver must be "hg" or "fd"')
    if launch is None:
        rng = np.random.default_rng(seed)
        w = rng.dirichlet(np.ones(modes))
        theta = rng.uniform(0, 2 * math.pi, modes)
        launch = np.sqrt(w) * np.exp(1j * theta)
    launch = np.asarray(launch, complex)
    launch = launch / np.linalg.norm(launch)
    rep = 1.0  # energy per pulse = P_avg / rep_rate
    _, uw0 = get_initial_state(launch, energy_j * rep, fwhm_s, rep, "sech_sq", sim)
    fiber = {"ϕ": phi, "Dω": D, "γ": gamma, "L": float(length_m), "zsave": None, "x": None,
             "gain_parameters": 0.0}
    fiber.update(raman_fields(sim, fR))
    return fiber_field_problem(uw0, fiber, sim, preset="GRIN_50_HG" if mode_solver == "hg" else "GRIN_50_FD")


# ------------------------------------------------ finite-difference modes
#
# The reference's own GRIN setup (fibers.jl:28-54, 91-132, 159-197): the index
# profile on an nx^2 grid, the scalar wave equation discretised with the
# 5-point stencil, its most-bound eigenpairs, beta(f) on a symmetric frequency
# stencil differentiated with central finite-difference weights, and the
# overlap tensor of the modes at f0.  The eigenpairs are found by shift-invert
# Lanczos about k^2 n_co^2 (above every bound mode), which returns the same
# eigenpairs as the reference's largest-real-part Arnoldi; within a degenerate
# group the basis is arbitrary, as there.


def build_grin(wavelength_um: float, nx: int, window_um: float, radius_um: float, core_na: float,
               alpha: float):
    """build_GRIN, fibers.jl:28-54: eps = n(r)^2 with n = max(n_cl, n_co - (n_co - n_cl)(r/a)^alpha)
    on x = (-nx/2 .. nx/2-1) dx; eps[i, j] at (x[j], y[i]) (meshgrid, helpers.jl:8-12)."""
    nco = silica_index(wavelength_um)
    ncl = math.sqrt(nco ** 2 - core_na ** 2)
    dx = window_um / nx
    x = np.arange(-nx / 2, nx / 2) * dx
    X = np.tile(x[None, :], (nx, 1))
    Y = np.tile(x[:, None], (1, nx))
    eps = np.maximum(ncl, nco - (nco - ncl) * (np.sqrt(X ** 2 + Y ** 2) / radius_um) ** alpha) ** 2
    return eps, x, dx


def solve_for_fiber_modes(wavelength_um: float, nmodes: int, dx: float, eps: np.ndarray, backend: str = "host"):
    """solve_for_fiber_modes, fibers.jl:159-197: the nmodes largest eigenpairs of
    A = laplacian_5pt + eps (2 pi / lambda)^2 (zero field outside the grid).
    Returns beta^2 [um^-2] (descending), unit-norm modes (nx*nx, nmodes) in
    column-major grid order, and n_eff = lambda sqrt(beta^2) / (2 pi).
    backend "host": scipy's shift-invert Lanczos; "device": Chebyshev-filtered
    subspace iteration in FP64 on the GPU (fh_fd_modes)."""
    if backend == "device":
        d, v = _fd_modes_device(eps, 2 * math.pi / wavelength_um, dx, nmodes)
        return d, v, wavelength_um * np.sqrt(d) / (2 * math.pi)
    if backend != "host":
        raise ValueError('backend must be "host" or "device"')
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    nx, ny = eps.shape
    k2 = (2 * math.pi / wavelength_um) ** 2
    c = 1.0 / dx ** 2
    # column-major index i + nx j: neighbours along the first grid index differ by 1, along the second by nx
    d0 = eps.reshape(-1, order="F") * k2 - 4 * c
    ex = np.full(nx * ny - 1, c)
    ex[np.arange(1, nx * ny) % nx == 0] = 0.0  # no coupling across columns
    ey = np.full(nx * ny - nx, c)
    A = sp.diags([d0, ex, ex, ey, ey], [0, 1, -1, nx, -nx], format="csc")
    sigma = float(np.max(eps)) * k2
    d, v = sla.eigsh(A, k=nmodes, sigma=sigma, which="LM", tol=1e-14)
    order = np.argsort(d)[::-1]
    d, v = d[order], v[:, order]
    v = v / np.sqrt(np.sum(v ** 2, axis=0))
    return d, v, wavelength_um * np.sqrt(d) / (2 * math.pi)


def _fd_modes_device(eps: np.ndarray, k0: float, dx: float, nmodes: int, tol: float = 1e-10, max_iter: int = 200):
    """fh_fd_modes: (beta^2 descending, modes (nx*ny, nmodes) column-major grid order)."""
    import ctypes

    from ._lib import check, lib
    from .engine import default_context

    nx, ny = eps.shape
    e = np.ascontiguousarray(np.asarray(eps, dtype=np.float64).reshape(-1, order="F"))
    d = np.empty(nmodes)
    v = np.empty((nmodes, nx * ny))
    its = ctypes.c_int32()
    dp = ctypes.POINTER(ctypes.c_double)
    check(lib().fh_fd_modes(default_context().handle, nx, ny, float(dx), float(k0), e.ctypes.data_as(dp), int(nmodes),
                            float(tol), int(max_iter), d.ctypes.data_as(dp), v.ctypes.data_as(dp), ctypes.byref(its)))
    return d, v.T


def _overlap_tensor_device(f: np.ndarray, scale: float) -> np.ndarray:
    """fh_overlap_tensor of the modes f (points, M)."""
    import ctypes

    from ._lib import check, lib
    from .engine import default_context

    points, m = f.shape
    phi = np.ascontiguousarray(np.asarray(f, dtype=np.float64).T)
    sk = np.empty(m ** 4)
    dp = ctypes.POINTER(ctypes.c_double)
    check(lib().fh_overlap_tensor(default_context().handle, int(points), int(m), phi.ctypes.data_as(dp),
                                  float(scale), sk.ctypes.data_as(dp)))
    return sk.reshape(m, m, m, m, order="F")


def central_fd_weights(points: int, order: int) -> np.ndarray:
    """Weights w on the offsets -(points-1)/2 .. (points-1)/2 with
    sum_j w_j o_j^m / m! = [m == order] for m < points (central_fdm(points, order).coefs)."""
    half = (points - 1) // 2
    o = np.arange(-half, half + 1, dtype=float)
    V = np.array([o ** m / math.factorial(m) for m in range(points)])
    rhs = np.zeros(points)
    rhs[order] = 1.0
    return np.linalg.solve(V, rhs)


def get_params(f0_thz: float, nx: int, window_um: float, radius_um: float, core_na: float, alpha: float, modes: int,
               nt: int, dt_ps: float, beta_order: int, delta_f_thz: float = 1.0, backend: str = "host"):
    """get_params, fibers.jl:91-132: (betas (beta_order+1, M), D (nt, M) rad/m,
    gamma (M,M,M,M) 1/(W m), modes (nx, nx, M), x) from finite-difference modes.
    backend "device" runs the mode solves and the overlap tensor on the GPU."""
    c0 = 2.99792458e8
    points = 2 * beta_order + 1
    half = (points - 1) // 2
    beta_f = np.zeros((points, modes))
    for i, off in enumerate(range(-half, half + 1)):
        lam_um = c0 / ((f0_thz + delta_f_thz * off) * 1e12) * 1e6
        eps, _, dx = build_grin(lam_um, nx, window_um, radius_um, core_na, alpha)
        _, _, neff = solve_for_fiber_modes(lam_um, modes, dx, eps, backend)
        beta_f[i] = 2 * math.pi * neff / (lam_um * 1e-6)
    dbeta = np.stack([central_fd_weights(points, n) @ beta_f / (2 * math.pi * delta_f_thz / 1e-12) ** n
                      for n in range(1, beta_order + 1)])
    betas = np.vstack([beta_f[half] - beta_f[half, 0], dbeta[0] - dbeta[0, 0], dbeta[1:]])
    w = 2 * math.pi * fftfreq(nt, 1 / dt_ps) * 1e12
    D = np.stack([w ** n / math.factorial(n) for n in range(beta_order + 1)], axis=1) @ betas
    lam0_um = c0 / f0_thz / 1e12 * 1e6
    eps, x, dx = build_grin(lam0_um, nx, window_um, radius_um, core_na, alpha)
    _, phi, _ = solve_for_fiber_modes(lam0_um, modes, dx, eps, backend)
    phi = phi.reshape(nx, nx, modes, order="F")
    sk = overlap_tensor(phi, dx * 1e-6, backend)
    gamma = sk * N2_SILICA * 2 * math.pi * f0_thz * 1e12 / c0
    return betas, D, gamma, phi, x
