"""Pump + signal propagation with gain (the reference's YDFA amplifier model).

Host mirror of ``src/gain_simulation/gain.jl`` (YDFAParams, cross sections,
PSD conversion, steady-state rate equations) and of the entry point
``solve_disp_gain_smf`` (``src/simulation/simulate_disp_gain_mmf.jl:200-239``).
The right-hand side ``disp_gain_smf!`` (:21-63) and its Tsit5 integration run
in libfiberhip (``fh_forward_gain``); this module folds the gain model into
the per-bin and scalar coefficients the kernel evaluates:

* ``W12 = sum_k w12[k] |u_k|^2``, ``W21 = sum_k w21[k] |u_k|^2`` -- the signal
  transition rates of ``calculate_gain_YDFA`` (gain.jl:154-155) with the PSD
  of ``psd_from_uω`` (:115-120) folded in;
* ``R12 = r12 Pp``, ``R21 = r21 Pp`` (:151-152);
* ``n2 = (R12 + W12) / (R12 + R21 + W12 + W21 + inv_tau)`` (:157);
* gain per bin ``g_k = ge[k] n2 - ga[k] n1`` (:160), pump gain
  ``gP = gpe n2 - gpa n1`` (:162).

Arrays are in FFT order: the reference evaluates the gain on the shifted
(ascending-frequency) grid and ``ifftshift``s it back (simulate_disp_gain_mmf.jl:98-100).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, lib

H_PLANCK = 6.62607015e-34  # gain.jl:151
C0 = 2.99792458e8


def _marcuse_overlap(core_radius: float, na: float, wavelength: float):
    """Normalised frequency, Marcuse mode-field radius and overlap factor (gain.jl:55-61)."""
    v = 2 * math.pi * core_radius * na / wavelength
    w = core_radius * (0.616 + 1.66 / (v ** 1.5) + 0.987 / (v ** 6))
    return v, w, 1 - math.exp(-2 * core_radius ** 2 / w ** 2)


@dataclass
class YDFAParams:
    """YDFAParams (gain.jl:30-75), with the defaults of the reference.  The
    derived fields (area, pump frequency, frequency grid in Hz, bin width and
    overlap factors) are computed from the given values at construction, as
    the reference's keyword defaults are."""

    fs: np.ndarray            # signal frequency grid [THz], ascending (sim["fs"])
    Δt: float                 # time step [ps]
    Nt: int
    L: float = 2.5
    core_radius: float = 2.5e-6
    NA: float = 0.13
    rho: float = 1e25
    tau21: float = 1.4e-3
    λp: float = 976e-9
    σap: float = 2.5e-24
    σep: float = 2.5e-24
    λs0: float = 1030e-9
    pulse_rep_rate: float = 80e6
    σas: Optional[np.ndarray] = None
    σes: Optional[np.ndarray] = None
    A: float = field(init=False)
    νp: float = field(init=False)
    νs0: float = field(init=False)
    νs: np.ndarray = field(init=False)
    dt: float = field(init=False)
    df: float = field(init=False)
    V_p: float = field(init=False)
    w_p: float = field(init=False)
    Gamma_p: float = field(init=False)
    V_s: float = field(init=False)
    w_s: float = field(init=False)
    Gamma_s: float = field(init=False)

    def __post_init__(self):
        self.fs = np.asarray(self.fs, dtype=np.float64)
        self.A = math.pi * self.core_radius ** 2
        self.νp = C0 / self.λp
        self.νs0 = C0 / self.λs0
        self.νs = self.fs * 1e12
        self.dt = self.Δt * 1e-12
        self.df = 1 / (self.Nt * self.dt)
        self.V_p, self.w_p, self.Gamma_p = _marcuse_overlap(self.core_radius, self.NA, self.λp)
        self.V_s, self.w_s, self.Gamma_s = _marcuse_overlap(self.core_radius, self.NA, self.λs0)
        if self.σas is None:
            self.σas = np.zeros(self.fs.size)
        if self.σes is None:
            self.σes = np.zeros(self.fs.size)


def get_ydfa_cross_sections(fs, absorption_wavelength_nm, absorption_intensity, emission_wavelength_nm,
                            emission_intensity, scale: float = 1e-27) -> dict:
    """get_ydfa_cross_sections (gain.jl:86-112): linear interpolation of the
    Yb absorption and emission spectra (wavelength in nm, intensity times
    `scale` in m^2) onto the grid's wavelengths c0 / (fs 1e12), flat outside
    the tabulated range.  The caller passes the tables (the reference reads
    them from Yb_absorption.npz / Yb_emission.npz beside gain.jl)."""
    lam_abs = np.asarray(absorption_wavelength_nm, dtype=np.float64) * 1e-9
    lam_em = np.asarray(emission_wavelength_nm, dtype=np.float64) * 1e-9
    s_abs = np.asarray(absorption_intensity, dtype=np.float64) * scale
    s_em = np.asarray(emission_intensity, dtype=np.float64) * scale
    for lam in (lam_abs, lam_em):
        if not np.all(np.diff(lam) > 0):
            raise ValueError("cross-section wavelengths must be strictly increasing")
    target = C0 / (np.asarray(fs, dtype=np.float64) * 1e12)
    return {"lambda": target, "sigma_as": np.interp(target, lam_abs, s_abs),
            "sigma_es": np.interp(target, lam_em, s_em)}


def get_YDFAParams(sim: dict, cross_sections: dict, **overrides) -> YDFAParams:
    """get_YDFAParams (gain.jl:175-182): parameters on the simulation grid with
    the cross sections of `cross_sections` (get_ydfa_cross_sections)."""
    p = YDFAParams(fs=sim["fs"], Δt=sim["Δt"], Nt=sim["Nt"], **overrides)
    p.σas = np.asarray(cross_sections["sigma_as"], dtype=np.float64).copy()
    p.σes = np.asarray(cross_sections["sigma_es"], dtype=np.float64).copy()
    return p


def psd_from_uω(uω, p: YDFAParams) -> np.ndarray:
    """psd_from_uω (gain.jl:115-120): physical PSD [W/Hz] on the shifted grid."""
    return np.fft.fftshift(np.abs(np.asarray(uω) * p.Nt * p.dt) ** 2, axes=0) * p.pulse_rep_rate


def calculate_gain_YDFA(Pp, Ps_vec, p: YDFAParams):
    """calculate_gain_YDFA (gain.jl:148-167) on the shifted grid -> (g(ν) [1/m], gP [1/m])."""
    h = H_PLANCK
    R12 = (p.Gamma_p * p.σap * Pp) / (h * p.νp * p.A)
    R21 = (p.Gamma_p * p.σep * Pp) / (h * p.νp * p.A)
    ps = np.asarray(Ps_vec).reshape(-1)
    W12 = np.sum((p.Gamma_s * p.σas * ps * p.df) / (h * p.νs * p.A))
    W21 = np.sum((p.Gamma_s * p.σes * ps * p.df) / (h * p.νs * p.A))
    n2 = (R12 + W12) / (R12 + R21 + W12 + W21 + 1 / p.tau21)
    n1 = 1 - n2
    g = p.Gamma_s * (p.σes * n2 - p.σas * n1) * p.rho
    gP = p.Gamma_p * (p.σep * n2 - p.σap * n1) * p.rho
    return g, gP


@dataclass
class GainCoefficients:
    """The gain model as the kernel evaluates it (arrays in FFT order)."""

    w12: np.ndarray
    w21: np.ndarray
    ga: np.ndarray
    ge: np.ndarray
    r12: float
    r21: float
    inv_tau: float
    gpa: float
    gpe: float

    def descriptor(self) -> "_lib.GainDesc":
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (self.w12, self.w21, self.ga, self.ge)]
        self._keep = arrs
        dp = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in arrs]
        return _lib.GainDesc(dp[0], dp[1], dp[2], dp[3], self.r12, self.r21, self.inv_tau, self.gpa, self.gpe)


def ydfa_coefficients(p: YDFAParams) -> GainCoefficients:
    """Fold calculate_gain_YDFA and psd_from_uω into per-bin coefficients on
    |u_k|^2 of the field in FFT order: Ps_k = |u_k Nt dt|^2 rep, and
    W12 = sum Gamma_s sigma_as Ps df / (h nu A) over the shifted grid."""
    h = H_PLANCK
    unshift = np.fft.ifftshift
    psd_scale = (p.Nt * p.dt) ** 2 * p.pulse_rep_rate
    w12 = unshift(p.Gamma_s * p.σas * psd_scale * p.df / (h * p.νs * p.A))
    w21 = unshift(p.Gamma_s * p.σes * psd_scale * p.df / (h * p.νs * p.A))
    return GainCoefficients(w12=w12, w21=w21, ga=unshift(p.Gamma_s * p.σas * p.rho),
                            ge=unshift(p.Gamma_s * p.σes * p.rho),
                            r12=p.Gamma_p * p.σap / (h * p.νp * p.A), r21=p.Gamma_p * p.σep / (h * p.νp * p.A),
                            inv_tau=1 / p.tau21, gpa=p.Gamma_p * p.σap * p.rho, gpe=p.Gamma_p * p.σep * p.rho)


def constant_coefficients(gain: float, nt: int) -> GainCoefficients:
    """compute_gain(uω, pGain::Number, Pp) (simulate_disp_gain_mmf.jl:76-80):
    g = pGain on every bin and gP = -1 (n2 = 0: every rate zero, inv_tau = 1)."""
    z = np.zeros(nt)
    return GainCoefficients(w12=z, w21=z, ga=np.full(nt, -float(gain)), ge=z, r12=0.0, r21=0.0, inv_tau=1.0,
                            gpa=1.0, gpe=0.0)


def coefficients(gain_parameters, nt: int) -> GainCoefficients:
    """fiber["gain_parameters"]: a number (constant gain), YDFAParams, or
    already-folded GainCoefficients."""
    if isinstance(gain_parameters, GainCoefficients):
        return gain_parameters
    if isinstance(gain_parameters, YDFAParams):
        if gain_parameters.Nt != nt:
            raise ValueError("YDFAParams grid does not match the simulation grid")
        return ydfa_coefficients(gain_parameters)
    return constant_coefficients(float(gain_parameters), nt)


@dataclass
class GainSolution:
    """Endpoint of solve_disp_gain_smf: the lab-frame output spectrum
    cis(D L) x(L), the pump power Pp(L) and the solver's counters."""

    output_spectrum: np.ndarray
    pump_power: np.ndarray
    stats: dict
    retcode: str = "Success"


def solve_disp_gain_smf(uω0, fiber: dict, sim: dict, pump_power=0.0, precision: str = "fp64",
                        rtol: float = 1e-5, atol: float = 1e-6, max_steps: int = 100000, context=None) -> GainSolution:
    """solve_disp_gain_smf (simulate_disp_gain_mmf.jl:200-239) on the GPU:
    Tsit5 (reltol 1e-5, OrdinaryDiffEq's default abstol 1e-6) on the state
    [Pp; ũω] from 0 to fiber["L"].  `uω0` is (nt, 1) or a batch (B, nt, 1)
    of launches; `pump_power` a scalar or one per launch.  Single mode only,
    as the reference requires."""
    from .engine import Plan, _dp
    from .problem import self_steepening

    if sim["M"] != 1:
        raise ValueError(f"disp_gain_smf is single mode only, requires M = 1 (got M = {sim['M']})")
    u = np.asarray(uω0, dtype=np.complex128)
    single = u.ndim == 2
    if single:
        u = u[None]
    B, nt = u.shape[0], u.shape[1]
    if u.shape != (B, sim["Nt"], 1):
        raise ValueError("uω0 must have shape (Nt, 1) or (batch, Nt, 1)")
    pump = np.broadcast_to(np.asarray(pump_power, dtype=np.float64), (B,)).copy()
    coef = coefficients(fiber.get("gain_parameters", 0.0), nt)
    plan = Plan(fiber["Dω"], fiber["γ"], fiber["hRω"], self_steepening(sim), fiber["one_m_fR"], fiber["L"], None,
                precision=precision, max_batch=B, scheme="tsit5", rtol=rtol, atol=atol, max_steps=max_steps,
                context=context)
    try:
        u0 = np.ascontiguousarray(u[:, :, 0])
        out = np.empty_like(u0)
        pumpL = np.empty(B)
        st = _lib.Stats()
        desc = coef.descriptor()
        check(lib().fh_forward_gain(plan.handle, B, ctypes.byref(desc), _dp(pump), _dp(u0.view(np.float64)),
                                    _dp(out.view(np.float64)), _dp(pumpL), ctypes.byref(st)))
        stats = dict(accepted=st.accepted, rejected=st.rejected, rhs_evals=st.rhs_evals)
    finally:
        plan.close()
    o = out[:, :, None]
    return GainSolution(o[0] if single else o, pumpL[0] if single else pumpL, stats)
