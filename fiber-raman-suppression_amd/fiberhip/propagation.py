"""propagate / metrics / verify over libfiberhip (propagation.jl mirror)."""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

from .engine import Plan
from .problem import FiberFieldProblem, fftfreq


@dataclasses.dataclass(frozen=True)
class PropagationResult:
    """propagation.jl:5-23 (spectra at z = 0 and z = L, lab frame)."""
    problem: FiberFieldProblem
    z_m: tuple
    input_spectrum: np.ndarray
    output_spectrum: np.ndarray
    accepted_steps: int
    rhs_evaluations: int
    precision: str
    solver: str = "RK4IP_fixed"

    @property
    def spectra(self) -> np.ndarray:
        return np.stack([self.input_spectrum, self.output_spectrum], axis=2)


def _forward_input(problem: FiberFieldProblem, input_spectrum) -> np.ndarray:
    """propagation.jl:25-32."""
    field = np.asarray(input_spectrum, dtype=np.complex128)
    if field.ndim == 1:
        field = field[:, None]
    if field.shape != problem.uω0.shape:
        raise ValueError(f"input spectrum size {field.shape} does not match {problem.uω0.shape}")
    if not np.all(np.isfinite(field)):
        raise ValueError("input spectrum must be finite")
    if not np.linalg.norm(field) > 0:
        raise ValueError("input spectrum must be nonzero")
    return field


def propagate(problem: FiberFieldProblem, steps: int = 2000, precision: str = "fp64",
              plan: Optional[Plan] = None, input_spectrum=None) -> PropagationResult:
    """propagate (propagation.jl:89-123) on the GPU engine.  The integrator is
    the fixed-step global-frame RK4IP of _scalar_rk4_output; ``steps`` sets
    its resolution (2000 steps tracks the reference's tight Tsit5 to ~6e-10,
    counterfactual-raman-benchmark.md:96-101)."""
    launch = _forward_input(problem, problem.uω0 if input_spectrum is None else input_spectrum)
    plan = plan or Plan.from_problem(problem, steps=steps, precision=precision)
    out, stats = plan.forward(launch)
    return PropagationResult(problem, (0.0, float(problem.fiber["L"])), launch, out,
                             stats["accepted"], stats["rhs_evaluations"], precision)


def scalar_rk4_output(problem: FiberFieldProblem, steps: int, precision: str = "fp64") -> np.ndarray:
    """_scalar_rk4_output (scalar_crosscheck.jl:8-68) on the GPU engine."""
    if isinstance(steps, bool) or not isinstance(steps, (int, np.integer)) or steps <= 0:
        raise ValueError("steps must be a positive integer")
    if problem.modes != 1:
        raise ValueError("the fixed-step propagation cross-check supports scalar problems only")
    if problem.fiber.get("gain_parameters", 0.0) not in (0, 0.0):
        raise ValueError("the fixed-step propagation cross-check does not support gain")
    return propagate(problem, steps=int(steps), precision=precision).output_spectrum


def _edge_fraction(power: np.ndarray) -> float:
    """propagation.jl:176-182."""
    k = max(1, int(math.floor(0.05 * len(power))))
    total = power.sum()
    if not total > 0:
        raise ValueError("field power must be nonzero")
    return float((power[:k].sum() + power[-k:].sum()) / total)


def field_metrics(spectrum: np.ndarray, sim: dict) -> dict:
    """_field_metrics / _photon_number, propagation.jl:184-202 (FP64)."""
    temporal = np.fft.fft(spectrum, axis=0)
    tp = np.sum(np.abs(temporal) ** 2, axis=1)
    mode_energy = np.sum(np.abs(temporal) ** 2, axis=0) * sim["Δt"]
    sp = np.fft.fftshift(np.sum(np.abs(spectrum) ** 2, axis=1))
    omega = np.fft.fftshift(np.abs(sim["ωs"]))
    return dict(energy_pj=float(mode_energy.sum()), mode_energy_pj=mode_energy, peak_power_w=float(tp.max()),
                temporal_edge_fraction=_edge_fraction(tp), spectral_edge_fraction=_edge_fraction(sp),
                photon_number=float(sim["Δt"] * np.sum(np.abs(spectrum) ** 2 / omega[:, None])))


def metrics(result: PropagationResult) -> dict:
    """_propagation_metrics, propagation.jl:204-237."""
    sim = result.problem.sim
    samples = [field_metrics(result.input_spectrum, sim), field_metrics(result.output_spectrum, sim)]
    e0, p0 = samples[0]["energy_pj"], samples[0]["photon_number"]
    return dict(
        samples=result.problem.nt, modes=result.problem.modes, saved_positions=2,
        input_energy_pj=e0, output_energy_pj=samples[1]["energy_pj"],
        input_mode_energy_pj=samples[0]["mode_energy_pj"], output_mode_energy_pj=samples[1]["mode_energy_pj"],
        input_peak_power_w=samples[0]["peak_power_w"], output_peak_power_w=samples[1]["peak_power_w"],
        max_temporal_edge_fraction=max(s["temporal_edge_fraction"] for s in samples),
        max_spectral_edge_fraction=max(s["spectral_edge_fraction"] for s in samples),
        max_relative_energy_change=max(abs(s["energy_pj"] / e0 - 1) for s in samples),
        max_photon_number_drift=max(abs(s["photon_number"] / p0 - 1) for s in samples),
        accepted_steps=result.accepted_steps, rhs_evaluations=result.rhs_evaluations)


def verify(result: PropagationResult, temporal_edge_limit=1e-3, spectral_edge_limit=1e-3,
           photon_drift_limit: Optional[float] = 1e-4, energy_drift_limit: Optional[float] = None) -> dict:
    """verify, propagation.jl:244-304 (gates on the FP64 evidence)."""
    for name, v in (("temporal_edge_limit", temporal_edge_limit), ("spectral_edge_limit", spectral_edge_limit)):
        if not (math.isfinite(v) and 0 <= v <= 1):
            raise ValueError(f"{name} must be finite and lie in [0, 1]")
    ev = metrics(result)
    finite = bool(np.all(np.isfinite(result.output_spectrum)))
    t_ok = ev["max_temporal_edge_fraction"] <= temporal_edge_limit
    s_ok = ev["max_spectral_edge_fraction"] <= spectral_edge_limit
    p_ok = None if photon_drift_limit is None else ev["max_photon_number_drift"] <= photon_drift_limit
    e_ok = None if energy_drift_limit is None else ev["max_relative_energy_change"] <= energy_drift_limit
    return dict(pass_=finite and t_ok and s_ok and p_ok is not False and e_ok is not False, finite_fields=finite,
                temporal_edges_contained=t_ok, spectral_edges_contained=s_ok, photon_number_conserved=p_ok,
                energy_change_acceptable=e_ok, **{k: ev[k] for k in (
                    "max_temporal_edge_fraction", "max_spectral_edge_fraction", "max_photon_number_drift",
                    "max_relative_energy_change")})


def centroid_thz(spectrum: np.ndarray, sim: dict) -> float:
    """_scalar_reference_centroid_thz, scalar_crosscheck.jl:88-98."""
    f = fftfreq(int(sim["Nt"]), 1 / sim["Δt"])
    e = np.sum(np.abs(spectrum) ** 2, axis=1)
    return float(np.sum(f * e) / np.sum(e))
