"""numpy FP64 restatement of FiberLab's setup helpers and GMMNLSE hot path.

TEST INFRASTRUCTURE ONLY.  This module is the oracle: only tests/,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import
it, and only to check or time against libfiberhip.  The product never
imports it.

Every function cites the reference (Julia) file:line it restates.  numpy's
pocketfft stands in for FFTW; the C restatement in fiberoracle.c uses its
own radix-2 FFT, so the two restatements cross-check each other.

Arrays are (nt, m): time/frequency index first, mode second, FFT order.
"""
from __future__ import annotations

import math

import numpy as np

C0 = 2.99792458e8  # m/s, helpers.jl:69
PLANCK = 6.62607e-34  # J s, helpers.jl:70

# physics_helpers.jl:1-26
SINGLE_MODE_PRESETS = {
    "SMF28": dict(gamma=1.1e-3, betas=[-2.17e-26, 1.2e-40], fR=0.18),
    "SMF28_beta2_only": dict(gamma=1.1e-3, betas=[-2.17e-26], fR=0.18),
    "HNLF": dict(gamma=10.0e-3, betas=[-0.5e-26, 1.0e-40], fR=0.18),
    "HNLF_zero_disp": dict(gamma=10.0e-3, betas=[-0.1e-26, 3.0e-40], fR=0.18),
}


def fftfreq(n: int, fs: float) -> np.ndarray:
    """FFTW.fftfreq(n, fs): sample frequencies in raw FFT order."""
    return np.fft.fftfreq(n, d=1.0 / fs)


def validate_positive_frequency_grid(lambda0: float, nt: int, window_ps: float) -> None:
    """helpers.jl:14-22."""
    f0 = C0 / lambda0 / 1e12
    minimum = nt / (2 * f0)
    if not window_ps > minimum:
        safe = (math.floor(minimum * 1000) + 1) / 1000
        raise ValueError(
            f"simulation bandwidth reaches nonpositive absolute optical frequencies: "
            f"Nt={nt} requires time_window > {minimum:.9g} ps (use at least {safe:.3f} ps)")


def sim_params(lambda0: float, m: int, nt: int, window_ps: float, beta_order: int,
               validate: bool = True) -> dict:
    """get_disp_sim_params, helpers.jl:56-84."""
    if not (nt >= 4 and nt & (nt - 1) == 0):
        raise ValueError("Nt must be a power of two >= 4")
    if beta_order < 2:
        raise ValueError("beta_order must be >= 2")
    f0 = C0 / lambda0 / 1e12
    if validate:
        validate_positive_frequency_grid(lambda0, nt, window_ps)
    dt = window_ps / nt
    ts = 1e-12 * np.array([-window_ps / 2 + i * dt for i in range(nt)])
    fs = f0 + np.fft.fftshift(fftfreq(nt, 1 / dt))
    w0 = 2 * np.pi * f0
    ws = 2 * np.pi * fs
    return dict(lambda0=lambda0, f0=f0, M=m, Nt=nt, time_window=window_ps, dt=dt, ts=ts,
                fs=fs, w0=w0, ws=ws, beta_order=beta_order)


def raman_fields(sim: dict, fR: float = 0.18, tau1: float = 12.2, tau2: float = 32.0) -> dict:
    """_single_oscillator_raman_fields, helpers.jl:97-130."""
    nt, dt = sim["Nt"], sim["dt"]
    omega = 2 * np.pi * fftfreq(nt, 1 / dt)
    decay, resonance = 1e3 / tau2, 1e3 / tau1
    response = fR * (decay ** 2 + resonance ** 2) / ((decay + 1j * omega) ** 2 + resonance ** 2)
    centered = np.exp(1j * (-omega * nt * dt / 2)) * response
    centered[nt // 2] = centered[nt // 2].real
    return {"hR": centered, "one_m_fR": 1 - fR}


def dispersion_user(sim: dict, betas) -> np.ndarray:
    """D(w) of get_disp_fiber_params_user_defined, helpers.jl:265-273, shape (nt,1)."""
    order = sim["beta_order"]
    if len(betas) > order - 1:
        raise ValueError("betas_user length must be <= beta_order-1")
    beta = np.concatenate([[0.0, 0.0], np.asarray(betas, float), np.zeros(order - 1 - len(betas))])
    w = 2 * np.pi * fftfreq(sim["Nt"], 1 / sim["dt"]) * 1e12
    cols = np.stack([w ** n / math.factorial(n) for n in range(order + 1)], axis=1)
    return (cols @ beta).reshape(-1, 1)


def shock(sim: dict) -> np.ndarray:
    """selfsteep = fftshift(ws/w0), simulate_disp_mmf.jl:89."""
    return np.fft.fftshift(sim["ws"] / sim["w0"])


def initial_state(modes, p_avg: float, fwhm: float, rep: float, form: str, sim: dict):
    """get_initial_state, simulate_disp_mmf.jl:140-159 -> (ut0, uw0)."""
    ts = sim["ts"]
    if form == "gauss":
        sigma = fwhm / 1.66511
        base = np.exp(-ts ** 2 / 2 / sigma ** 2)
        p_peak = 0.939437 * p_avg / fwhm / rep
    elif form == "sech_sq":
        tau = fwhm / 1.7627
        base = 1 / np.cosh(-ts / tau)
        p_peak = 0.881374 * p_avg / fwhm / rep
    else:
        raise ValueError("unsupported pulse_form")
    u = base.astype(complex)[:, None] * np.asarray(modes, complex)[None, :].conj()
    u = u / np.sqrt(np.max(np.sum(np.abs(u) ** 2, axis=1)))
    ut0 = u * np.sqrt(p_peak)
    return ut0, np.fft.ifft(ut0, axis=0)


def taylor_phase_basis(sim: dict, orders, scales=None) -> np.ndarray:
    """taylor_phase_basis, design_api.jl:61-101."""
    orders = tuple(orders)
    scales = np.ones(len(orders)) if scales is None else np.asarray(scales, float)
    omega_fs = 2 * np.pi * fftfreq(sim["Nt"], 1 / sim["dt"]) * 1e-3
    return np.stack([s * omega_fs ** o / math.factorial(o) for o, s in zip(orders, scales)], axis=1)


def frequency_band_mask(sim: dict, lower: float, upper: float) -> np.ndarray:
    """frequency_band_mask, scientific_metrics.jl:190-201 (inclusive)."""
    f = fftfreq(sim["Nt"], 1 / sim["dt"])
    mask = (lower <= f) & (f <= upper)
    if not mask.any():
        raise ValueError("band selects no bins")
    return mask


def band_cost(u: np.ndarray, mask: np.ndarray):
    """_field_spectral_band_cost, physics_helpers.jl:222-232 -> (J, Wirtinger seed)."""
    total = np.sum(np.abs(u) ** 2)
    band = np.sum(np.abs(u[mask, :]) ** 2)
    cost = band / total
    return float(cost), u * (mask[:, None] - cost) / total


def centroid_thz(u: np.ndarray, sim: dict) -> float:
    """_scalar_reference_centroid_thz, scalar_crosscheck.jl:88-98."""
    f = fftfreq(sim["Nt"], 1 / sim["dt"])
    e = np.sum(np.abs(u) ** 2, axis=1)
    return float(np.sum(f * e) / np.sum(e))


def edge_fraction(power: np.ndarray) -> float:
    """_edge_fraction, propagation.jl:176-182."""
    k = max(1, int(math.floor(0.05 * len(power))))
    return float((power[:k].sum() + power[-k:].sum()) / power.sum())


def field_metrics(spectrum: np.ndarray, sim: dict) -> dict:
    """_field_metrics and _photon_number, propagation.jl:184-202."""
    temporal = np.fft.fft(spectrum, axis=0)
    tp = np.sum(np.abs(temporal) ** 2, axis=1)
    mode_energy = np.sum(np.abs(temporal) ** 2, axis=0) * sim["dt"]
    sp = np.fft.fftshift(np.sum(np.abs(spectrum) ** 2, axis=1))
    omega = np.fft.fftshift(np.abs(sim["ws"]))
    photons = sim["dt"] * np.sum(np.abs(spectrum) ** 2 / omega[:, None])
    return dict(energy_pj=float(mode_energy.sum()), mode_energy_pj=mode_energy,
                peak_power_w=float(tp.max()), temporal_edge_fraction=edge_fraction(tp),
                spectral_edge_fraction=edge_fraction(sp), photon_number=float(photons))


# ------------------------------------------------------------ hot path


def rhs(z: float, x: np.ndarray, D: np.ndarray, gamma: np.ndarray, hR: np.ndarray,
        shock_: np.ndarray, one_m_fR: float) -> np.ndarray:
    """disp_mmf!, simulate_disp_mmf.jl:25-60 (gamma indexed [i,j,k,l])."""
    u = np.fft.fft(np.exp(1j * D * z) * x, axis=0)
    v, w = u.real, u.imag
    dK = np.einsum("ijkl,tk,tl->tij", gamma, v, v) + np.einsum("ijkl,tk,tl->tij", gamma, w, w)
    etaK = one_m_fR * (np.einsum("tij,tj->ti", dK, v) + 1j * np.einsum("tij,tj->ti", dK, w))
    conv = np.fft.ifft(hR[:, None, None] * np.fft.fft(dK.astype(complex), axis=0), axis=0)
    dR = np.fft.fftshift(conv, axes=0).real
    etaR = np.einsum("tij,tj->ti", dR, v) + 1j * np.einsum("tij,tj->ti", dR, w)
    eta = np.fft.ifft(etaK + etaR, axis=0) * shock_[:, None]
    return 1j * np.exp(-1j * D * z) * eta


def rk4(x0: np.ndarray, length: float, steps: int, D, gamma, hR, shock_, one_m_fR) -> np.ndarray:
    """_scalar_rk4_output, scalar_crosscheck.jl:50-65 (any M); returns the lab spectrum."""
    x = x0.copy()
    h = length / steps
    for i in range(steps):
        z = i * h
        k1 = rhs(z, x, D, gamma, hR, shock_, one_m_fR)
        k2 = rhs(z + h / 2, x + h / 2 * k1, D, gamma, hR, shock_, one_m_fR)
        k3 = rhs(z + h / 2, x + h / 2 * k2, D, gamma, hR, shock_, one_m_fR)
        k4 = rhs(z + h, x + h * k3, D, gamma, hR, shock_, one_m_fR)
        x = x + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return np.exp(1j * D * length) * x


def rhs_vjp(z, x, gy, D, gamma, hR, shock_, one_m_fR):
    """Pullback of rhs() for a real objective (real-linear map), derived
    from the forward operation order of disp_mmf!; R^T uses conj(hR) with the
    sign alternation as adjoint_disp_mmf! does (sensitivity_disp_mmf.jl:96,127,241)."""
    nt = x.shape[0]
    P = np.exp(1j * D * z)
    u = np.fft.fft(P * x, axis=0)
    v, w = u.real, u.imag
    dK = np.einsum("ijkl,tk,tl->tij", gamma, v, v) + np.einsum("ijkl,tk,tl->tij", gamma, w, w)
    sigma = np.where(np.arange(nt) % 2 == 0, 1.0, -1.0)
    H = hR * sigma
    dR = np.fft.ifft(H[:, None, None] * np.fft.fft(dK, axis=0), axis=0).real
    h = one_m_fR * dK + dR
    G = np.fft.fft(-1j * P * shock_[:, None] * gy, axis=0) / nt
    gu = np.einsum("tij,ti->tj", h, G)
    A = np.real(np.conj(G)[:, :, None] * u[:, None, :])
    B = one_m_fR * A + np.fft.ifft(np.conj(H)[:, None, None] * np.fft.fft(A, axis=0), axis=0).real
    C = np.einsum("tij,ijkl->tkl", B, gamma)
    gu = gu + np.einsum("tkl,tl->tk", C + np.transpose(C, (0, 2, 1)), u)
    return np.conj(P) * (nt * np.fft.ifft(gu, axis=0))


def rk4_adjoint(x0, lam_out, length, steps, D, gamma, hR, shock_, one_m_fR):
    """Discrete adjoint of rk4(): Wirtinger seed on the lab output -> seed on x0."""
    h = length / steps
    args = (D, gamma, hR, shock_, one_m_fR)
    xs = [x0.copy()]
    x = x0.copy()
    for i in range(steps):
        z = i * h
        k1 = rhs(z, x, *args)
        k2 = rhs(z + h / 2, x + h / 2 * k1, *args)
        k3 = rhs(z + h / 2, x + h / 2 * k2, *args)
        k4 = rhs(z + h, x + h * k3, *args)
        x = x + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        xs.append(x)
    lam = np.exp(-1j * D * length) * lam_out
    for i in range(steps - 1, -1, -1):
        z, x = i * h, xs[i]
        k1 = rhs(z, x, *args)
        y2 = x + h / 2 * k1
        k2 = rhs(z + h / 2, y2, *args)
        y3 = x + h / 2 * k2
        k3 = rhs(z + h / 2, y3, *args)
        y4 = x + h * k3
        g4 = rhs_vjp(z + h, y4, h / 6 * lam, *args)
        g3 = rhs_vjp(z + h / 2, y3, h / 3 * lam + h * g4, *args)
        g2 = rhs_vjp(z + h / 2, y2, h / 3 * lam + h / 2 * g3, *args)
        g1 = rhs_vjp(z, x, h / 6 * lam + h / 2 * g2, *args)
        lam = lam + g1 + g2 + g3 + g4
    return lam


def phase_gradient(lam0: np.ndarray, u0_shaped: np.ndarray) -> np.ndarray:
    """g_phi = 2 Re(conj(lam0) i u0), physics_models.jl:1042 (summed over modes)."""
    return np.sum(2 * np.real(np.conj(lam0) * 1j * u0_shaped), axis=1)


# ------------------------------------------------------ multimode costs


def mmf_costs(u: np.ndarray, mask: np.ndarray, kind: str, tau: float = 50.0):
    """mmf_cost_sum / mmf_cost_fundamental / mmf_cost_worst_mode (mmf_cost.jl:
    49-62, 81-98, 124-166) -> (J, dJ/dconj(u)), written mode by mode."""
    nt, m = u.shape
    e_tot = [float(np.vdot(u[:, k], u[:, k]).real) for k in range(m)]
    e_band = [float(np.vdot(u[mask, k], u[mask, k]).real) for k in range(m)]
    ind = mask.astype(float)
    seed = np.zeros((nt, m), dtype=complex)
    if kind == "sum":
        J = sum(e_band) / sum(e_tot)
        for k in range(m):
            seed[:, k] = u[:, k] * (ind - J) / sum(e_tot)
        return J, seed
    if kind == "fundamental":
        J = e_band[0] / e_tot[0]
        seed[:, 0] = u[:, 0] * (ind - J) / e_tot[0]
        return J, seed
    if kind == "worst_mode":
        act = [k for k in range(m) if e_tot[k] > 0]
        r = {k: e_band[k] / e_tot[k] for k in act}
        rmax = max(r.values())
        z = sum(math.exp(tau * (r[k] - rmax)) for k in act)
        J = rmax + (math.log(z) - math.log(len(act))) / tau
        for k in act:
            w = math.exp(tau * (r[k] - rmax)) / z
            seed[:, k] = w * u[:, k] * (ind - r[k]) / e_tot[k]
        return J, seed
    raise ValueError(kind)
