"""Problem construction: the FiberLab surface that feeds the propagation seam.

Host-side mirror of the reference's notebook API for single-mode problems
(``Fiber``/``Pulse``/``Grid``/``fiber_problem``) and of the low-level builders
it calls.  Each function cites the Julia it follows; argument meaning,
defaults and error behaviour (``ValueError`` where Julia throws
``ArgumentError``) are kept.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional, Sequence

import numpy as np

C0 = 2.99792458e8  # m/s (helpers.jl:69)
PLANCK = 6.62607e-34  # J s (helpers.jl:70)
SILICA_RAMAN_FRACTION = 0.18  # helpers.jl:86
SILICA_RAMAN_TAU1_FS = 12.2
SILICA_RAMAN_TAU2_FS = 32.0

# physics_helpers.jl:1-26
SINGLE_MODE_FIBER_PRESETS = {
    "SMF28": dict(name="SMF-28", gamma=1.1e-3, betas=(-2.17e-26, 1.2e-40), fR=0.18),
    "SMF28_beta2_only": dict(name="SMF-28 beta2 only", gamma=1.1e-3, betas=(-2.17e-26,), fR=0.18),
    "HNLF": dict(name="HNLF", gamma=10.0e-3, betas=(-0.5e-26, 1.0e-40), fR=0.18),
    "HNLF_zero_disp": dict(name="HNLF zero dispersion", gamma=10.0e-3, betas=(-0.1e-26, 3.0e-40),
                           fR=0.18),
}


@dataclasses.dataclass(frozen=True)
class Fiber:
    """api.jl:13-40."""
    length_m: float
    power_w: float
    regime: str = "single_mode"
    preset: str = "SMF28"
    beta_order: int = 3
    raman_fraction: Optional[float] = None

    def __post_init__(self):
        if self.raman_fraction is not None and not (0 <= self.raman_fraction <= 1):
            raise ValueError("fiber raman_fraction must be finite and lie in [0, 1]")


@dataclasses.dataclass(frozen=True)
class Pulse:
    """api.jl:46-60 defaults: 185 fs sech^2 at 80.5 MHz."""
    fwhm_s: float = 185e-15
    rep_rate_hz: float = 80.5e6
    shape: str = "sech_sq"


@dataclasses.dataclass(frozen=True)
class Grid:
    """Sampling request; policy "exact"/"fixed" validates, "auto_if_undersized" grows."""
    nt: int = 1024
    time_window_ps: float = 10.0
    policy: str = "auto_if_undersized"


def fftfreq(n: int, fs: float) -> np.ndarray:
    """FFTW.fftfreq(n, fs) in raw FFT order."""
    k = np.arange(n)
    k = np.where(k < (n + 1) // 2, k, k - n)
    return k * (fs / n)


# --------------------------------------------------------------- grids


def validate_positive_frequency_grid(wavelength_m: float, nt: int, time_window_ps: float) -> None:
    """helpers.jl:14-22."""
    f0 = C0 / wavelength_m / 1e12
    minimum = nt / (2 * f0)
    if not time_window_ps > minimum:
        safe = (math.floor(minimum * 1000) + 1) / 1000
        raise ValueError(
            "simulation bandwidth reaches nonpositive absolute optical frequencies: "
            f"Nt={nt} at λ0={wavelength_m * 1e9:.4g} nm requires time_window > {minimum:.9g} ps "
            f"(use at least {safe:.3f} ps); increase the window or use fewer samples")


def _peak_power_from_average(power_w, fwhm_s, rep_rate_hz, shape):
    """physics_helpers.jl:79-84."""
    factor = {"sech_sq": 0.881374, "gauss": 0.939437, "gaussian": 0.939437}.get(shape)
    if factor is None:
        raise ValueError(f"unsupported pulse shape `{shape}`")
    return factor * power_w / (fwhm_s * rep_rate_hz)


def _recommended_time_window_ps(length_m, pulse: Pulse, gamma, beta2, power_w):
    """physics_helpers.jl:86-95."""
    p_peak = _peak_power_from_average(power_w, pulse.fwhm_s, pulse.rep_rate_hz, pulse.shape)
    walkoff_ps = abs(beta2) * length_m * (2 * math.pi * 13e12) * 1e12
    t0 = pulse.fwhm_s / 1.763
    spm_angular_width = 0.86 * (gamma * p_peak * length_m) / t0
    spm_ps = abs(beta2) * length_m * spm_angular_width * 1e12
    return max(5.0, math.ceil(2.0 * (walkoff_ps + spm_ps + 0.5)))


def _nt_for_window_ps(time_window_ps, max_time_step_ps=0.0105):
    """physics_helpers.jl:97-103."""
    nt = 1
    while nt < math.ceil(time_window_ps / max_time_step_ps):
        nt <<= 1
    return nt


def resolve_sampling_grid(grid: Grid, wavelength_m=1550e-9, minimum_time_window_ps=0.0,
                          max_time_step_ps=0.0105, minimum_frequency_fraction=0.1) -> Grid:
    """physics_helpers.jl:138-185."""
    nt, window = int(grid.nt), float(grid.time_window_ps)
    if not (nt >= 4 and nt & (nt - 1) == 0):
        raise ValueError("grid nt must be a power of two ≥ 4")
    if not (math.isfinite(window) and window > 0):
        raise ValueError("grid time_window_ps must be positive and finite")
    if grid.policy in ("exact", "fixed"):
        validate_positive_frequency_grid(wavelength_m, nt, window)
        return Grid(nt=nt, time_window_ps=window, policy="exact")
    if grid.policy != "auto_if_undersized":
        raise ValueError(f"unsupported grid policy `{grid.policy}`")
    f0 = C0 / wavelength_m / 1e12

    def min_margin(n):
        return n / (2 * f0) / (1 - minimum_frequency_fraction)

    window = max(window, minimum_time_window_ps)
    nt = max(nt, _nt_for_window_ps(window, max_time_step_ps))
    if window <= min_margin(nt):
        minimum = min_margin(nt)
        rounded = (math.floor(minimum * 1000) + 1) / 1000
        window = rounded if rounded / nt <= max_time_step_ps else np.nextafter(minimum, np.inf)
    validate_positive_frequency_grid(wavelength_m, nt, window)
    return Grid(nt=nt, time_window_ps=float(window), policy="exact")


def resolve_grid(fiber: Fiber, pulse: Pulse, grid: Grid, wavelength_m=1550e-9) -> Grid:
    """physics_helpers.jl:186-201 (single-mode presets)."""
    preset = SINGLE_MODE_FIBER_PRESETS[fiber.preset]
    recommended = 0.0
    if grid.policy == "auto_if_undersized":
        recommended = _recommended_time_window_ps(fiber.length_m, pulse, preset["gamma"],
                                                  preset["betas"][0], fiber.power_w)
    return resolve_sampling_grid(grid, wavelength_m=wavelength_m, minimum_time_window_ps=recommended)


def get_disp_sim_params(wavelength_m: float, m: int, nt: int, time_window_ps: float, beta_order: int,
                        validate: bool = True) -> dict:
    """get_disp_sim_params, helpers.jl:56-84.  ``validate=False`` skips the
    positive-frequency validator (low-level variant only)."""
    if not (math.isfinite(wavelength_m) and wavelength_m > 0):
        raise ValueError("λ0 must be positive and finite")
    if m < 1:
        raise ValueError("M must be a positive integer")
    if not (nt >= 4 and nt & (nt - 1) == 0):
        raise ValueError("Nt must be a power of two ≥ 4")
    if not (math.isfinite(time_window_ps) and time_window_ps > 0):
        raise ValueError("time_window must be positive and finite")
    if beta_order < 2:
        raise ValueError("β_order must be an integer ≥ 2")
    f0 = C0 / wavelength_m / 1e12
    if validate:
        validate_positive_frequency_grid(wavelength_m, nt, time_window_ps)
    dt = time_window_ps / nt
    ts = 1e-12 * np.array([-time_window_ps / 2 + i * dt for i in range(nt)])
    fs = f0 + np.fft.fftshift(fftfreq(nt, 1 / dt))
    w0 = 2 * math.pi * f0
    return {"λ0": wavelength_m, "f0": f0, "M": m, "Nt": nt, "time_window": time_window_ps, "Δt": dt,
            "ts": ts, "fs": fs, "ω0": w0, "ωs": 2 * math.pi * fs, "c0": C0, "h": PLANCK,
            "ε": 1e-12 * dt / (PLANCK * 1e12 * f0), "β_order": beta_order}


def raman_fields(sim: dict, fR=SILICA_RAMAN_FRACTION, tau1=SILICA_RAMAN_TAU1_FS,
                 tau2=SILICA_RAMAN_TAU2_FS) -> dict:
    """_single_oscillator_raman_fields, helpers.jl:97-130."""
    fraction, t1, t2 = float(fR), float(tau1), float(tau2)
    if not (math.isfinite(fraction) and 0 <= fraction <= 1):
        raise ValueError("fR must be finite and lie in [0, 1]")
    if not (math.isfinite(t1) and t1 > 0):
        raise ValueError("τ1 must be positive and finite")
    if not (math.isfinite(t2) and t2 > 0):
        raise ValueError("τ2 must be positive and finite")
    nt, dt = int(sim["Nt"]), float(sim["Δt"])
    omega = 2 * math.pi * fftfreq(nt, 1 / dt)
    decay, resonance = 1e3 / t2, 1e3 / t1
    with np.errstate(all="ignore"):
        response = fraction * (decay ** 2 + resonance ** 2) / ((decay + 1j * omega) ** 2 + resonance ** 2)
        centered = np.exp(1j * (-omega * nt * dt / 2)) * response
    centered[nt // 2] = centered[nt // 2].real
    if not (math.isfinite(decay) and math.isfinite(resonance) and np.all(np.isfinite(centered))):
        raise ValueError("Raman response is not finite for the requested parameters")
    return {"hRω": centered, "one_m_fR": 1 - fraction,
            "raman_response_model": "blow_wood_single_damped_oscillator_v1",
            "raman_fraction": fraction, "raman_tau1_fs": t1, "raman_tau2_fs": t2}


def get_disp_fiber_params_user_defined(length_m: float, sim: dict, fR=SILICA_RAMAN_FRACTION,
                                       tau1=SILICA_RAMAN_TAU1_FS, tau2=SILICA_RAMAN_TAU2_FS,
                                       gamma_user=None, betas_user=None) -> dict:
    """helpers.jl:245-282."""
    if gamma_user is None or betas_user is None:
        raise ValueError("gamma_user and betas_user must both be provided")
    order = int(sim["β_order"])
    if len(betas_user) > order - 1:
        raise ValueError(f"betas_user length must be ≤ β_order-1 ({order - 1}); got {len(betas_user)}")
    raman = raman_fields(sim, fR, tau1, tau2)
    beta = np.concatenate([[0.0, 0.0], np.asarray(betas_user, float), np.zeros(order - 1 - len(betas_user))])
    w = 2 * math.pi * fftfreq(int(sim["Nt"]), 1 / sim["Δt"]) * 1e12
    cols = np.stack([w ** n / math.factorial(n) for n in range(order + 1)], axis=1)
    fiber = {"ϕ": None, "Dω": (cols @ beta).reshape(-1, 1), "γ": np.full((1, 1, 1, 1), float(gamma_user)),
             "L": float(length_m), "zsave": None, "x": None, "gain_parameters": 0.0}
    fiber.update(raman)
    return fiber


def get_initial_state(modes: Sequence[complex], p_avg: float, fwhm_s: float, rep_rate_hz: float,
                      pulse_form: str, sim: dict):
    """get_initial_state, simulate_disp_mmf.jl:140-159 -> (ut0, uω0)."""
    ts = sim["ts"]
    if pulse_form == "gauss":
        sigma = fwhm_s / 1.66511
        base = np.exp(-ts ** 2 / 2 / sigma ** 2)
        p_peak = 0.939437 * p_avg / fwhm_s / rep_rate_hz
    elif pulse_form == "sech_sq":
        tau = fwhm_s / 1.7627
        with np.errstate(over="ignore"):  # cosh overflows to inf far out in wide windows: sech -> 0 there
            base = 1 / np.cosh(-ts / tau)
        p_peak = 0.881374 * p_avg / fwhm_s / rep_rate_hz
    else:
        raise ValueError(f'unsupported pulse_form={pulse_form!r}; expected "gauss" or "sech_sq"')
    u = base.astype(complex)[:, None] * np.conj(np.asarray(modes, complex))[None, :]
    u = u / math.sqrt(np.max(np.sum(np.abs(u) ** 2, axis=1)))
    ut0 = u * math.sqrt(p_peak)
    return ut0, np.fft.ifft(ut0, axis=0)


def self_steepening(sim: dict) -> np.ndarray:
    """selfsteep = fftshift(ωs/ω0), simulate_disp_mmf.jl:89."""
    return np.fft.fftshift(sim["ωs"] / sim["ω0"])


def taylor_phase_basis(problem_or_sim, orders=(2, 3, 4), coefficient_scales_fs=None) -> np.ndarray:
    """taylor_phase_basis, design_api.jl:61-101: columns scale*Ω^n/n!, Ω in rad/fs."""
    sim = problem_or_sim.sim if isinstance(problem_or_sim, FiberFieldProblem) else problem_or_sim
    orders = (orders,) if isinstance(orders, int) else tuple(orders)
    if not orders:
        raise ValueError("taylor_phase_basis requires at least one order")
    if any((not isinstance(o, (int, np.integer))) or isinstance(o, bool) for o in orders):
        raise ValueError("Taylor phase orders must be integers")
    if any(o < 0 for o in orders):
        raise ValueError("Taylor phase orders must be non-negative")
    if len(set(orders)) != len(orders):
        raise ValueError("Taylor phase orders must be unique")
    if coefficient_scales_fs is None:
        scales = np.ones(len(orders))
    else:
        scales = np.atleast_1d(np.asarray(coefficient_scales_fs, float))
        if len(scales) != len(orders):
            raise ValueError("coefficient_scales_fs must provide one scale per Taylor order")
        if not np.all(np.isfinite(scales) & (scales > 0)):
            raise ValueError("coefficient_scales_fs values must be positive and finite")
    omega_fs = 2 * math.pi * fftfreq(int(sim["Nt"]), 1 / sim["Δt"]) * 1e-3
    return np.stack([s * omega_fs ** o / math.factorial(o) for o, s in zip(orders, scales)], axis=1)


def frequency_band_mask(problem_or_sim, interval) -> np.ndarray:
    """frequency_band_mask, scientific_metrics.jl:190-201 (inclusive, FFT order)."""
    sim = problem_or_sim.sim if isinstance(problem_or_sim, FiberFieldProblem) else problem_or_sim
    lower, upper = (float(x) for x in interval)
    if not (math.isfinite(lower) and math.isfinite(upper) and lower < upper):
        raise ValueError("frequency band must contain finite lower < upper bounds")
    f = fftfreq(int(sim["Nt"]), 1 / sim["Δt"])
    mask = (lower <= f) & (f <= upper)
    if not mask.any():
        raise ValueError(f"frequency band [{lower}, {upper}] THz selects no bins on this grid")
    return mask


# ------------------------------------------------------------- problems


@dataclasses.dataclass
class FiberFieldProblem:
    """physics_models.jl:9-17: launch spectrum, fiber dict, sim dict, band mask."""
    uω0: np.ndarray
    fiber: dict
    sim: dict
    band_mask: Optional[np.ndarray] = None
    preset: str = "custom"

    @property
    def nt(self) -> int:
        return int(self.sim["Nt"])

    @property
    def modes(self) -> int:
        return int(self.sim["M"])


def fiber_field_problem(uω0, fiber: dict, sim: dict, band_mask=None, raman_threshold_thz=None,
                        preset="custom") -> FiberFieldProblem:
    """physics_models.jl:459-577 (the checks the propagation seam relies on)."""
    u = np.asarray(uω0, dtype=np.complex128)
    if u.ndim == 1:
        u = u[:, None]
    nt, m = int(sim["Nt"]), int(sim["M"])
    if u.shape != (nt, m):
        raise ValueError(f"uω0 has size {u.shape}, expected {(nt, m)}")
    if not np.all(np.isfinite(u)):
        raise ValueError("uω0 must be finite")
    D = np.asarray(fiber["Dω"], dtype=np.float64)
    if D.shape != (nt, m):
        raise ValueError(f"Dω has size {D.shape}, expected {(nt, m)}")
    gamma = np.asarray(fiber["γ"])
    if not np.isrealobj(gamma):
        raise ValueError("gamma tensor must use real-valued storage")
    if gamma.shape != (m,) * 4 or not np.all(np.isfinite(gamma)):
        raise ValueError("gamma tensor must be finite with shape (M, M, M, M)")
    if band_mask is None and raman_threshold_thz is not None:
        band_mask = fftfreq(nt, 1 / sim["Δt"]) < float(raman_threshold_thz)
        if not band_mask.any():
            raise ValueError("raman_threshold_thz selects no frequency bins")
    return FiberFieldProblem(u, fiber, sim, None if band_mask is None else np.asarray(band_mask, bool),
                             preset)


def fiber_problem(fiber: Fiber, pulse: Pulse = Pulse(), grid: Grid = Grid(), wavelength_m=1550e-9,
                  raman_threshold_thz=-5.0, band_mask=None) -> FiberFieldProblem:
    """fiber_problem for single-mode presets, physics_models.jl:330-360 with
    _single_mode_low_level_setup (:115-146)."""
    if fiber.regime != "single_mode":
        raise ValueError("this host mirror builds single-mode preset problems; pass explicit operators "
                         "through fiber_field_problem for multimode fibers")
    if fiber.preset not in SINGLE_MODE_FIBER_PRESETS:
        raise ValueError(f"unknown single-mode fiber preset `{fiber.preset}`")
    preset = SINGLE_MODE_FIBER_PRESETS[fiber.preset]
    resolved = resolve_grid(fiber, pulse, grid, wavelength_m)
    sim = get_disp_sim_params(wavelength_m, 1, resolved.nt, resolved.time_window_ps, fiber.beta_order)
    fR = preset["fR"] if fiber.raman_fraction is None else fiber.raman_fraction
    fib = get_disp_fiber_params_user_defined(fiber.length_m, sim, fR=fR, gamma_user=preset["gamma"],
                                             betas_user=preset["betas"])
    shape = "gauss" if pulse.shape in ("gauss", "gaussian") else pulse.shape
    _, uw0 = get_initial_state([1.0], fiber.power_w, pulse.fwhm_s, pulse.rep_rate_hz, shape, sim)
    return fiber_field_problem(uw0, fib, sim, band_mask=band_mask,
                               raman_threshold_thz=raman_threshold_thz if band_mask is None else None,
                               preset=fiber.preset)


def with_launch(problem: FiberFieldProblem, launch) -> FiberFieldProblem:
    """physics_models.jl:667: same fiber and grid, new launch spectrum."""
    return fiber_field_problem(launch, problem.fiber, problem.sim, band_mask=problem.band_mask,
                               preset=problem.preset)


def with_raman_fraction(problem: FiberFieldProblem, fraction: float) -> FiberFieldProblem:
    """physics_models.jl:616: rebuild hR/one_m_fR with a new fraction."""
    fiber = dict(problem.fiber)
    fiber.update(raman_fields(problem.sim, fraction, fiber.get("raman_tau1_fs", SILICA_RAMAN_TAU1_FS),
                              fiber.get("raman_tau2_fs", SILICA_RAMAN_TAU2_FS)))
    return FiberFieldProblem(problem.uω0, fiber, problem.sim, problem.band_mask, problem.preset)
