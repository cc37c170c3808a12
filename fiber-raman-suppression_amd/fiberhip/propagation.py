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
    """propagation.jl:5-23: lab-frame spectra at the saved positions z_m
    (first 0, last L), shape (nt, m, nz) as the reference stores them."""
    problem: FiberFieldProblem
    z_m: tuple
    saved: np.ndarray  # (nz, nt, m)
    accepted_steps: int
    rhs_evaluations: int
    precision: str
    solver: str
    plan: Plan

    @property
    def spectra(self) -> np.ndarray:
        return np.transpose(self.saved, (1, 2, 0))

    @property
    def input_spectrum(self) -> np.ndarray:
        return self.saved[0]

    @property
    def output_spectrum(self) -> np.ndarray:
        return self.saved[-1]


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


def _save_positions(saveat, length_m: float) -> np.ndarray:
    """propagation.jl:67-77."""
    if saveat is None:
        return np.array([0.0, length_m])
    if isinstance(saveat, (str, bytes)) or not hasattr(saveat, "__len__"):
        raise ValueError("saveat must be a vector of distances")
    z = np.asarray(saveat, dtype=np.float64).ravel()
    if z.size < 2:
        raise ValueError("saveat must include 0 and the fiber length")
    if not np.all(np.isfinite(z)):
        raise ValueError("saveat distances must be finite")
    if not np.all(np.diff(z) > 0):
        raise ValueError("saveat distances must be sorted and unique")
    if not (z[0] == 0.0 and z[-1] == length_m):
        raise ValueError(f"saveat must start at 0 and end at the fiber length {length_m}")
    return z


def propagate(problem: FiberFieldProblem, saveat=None, steps: Optional[int] = None, precision: str = "fp64",
              plan: Optional[Plan] = None, input_spectrum=None) -> PropagationResult:
    """propagate (propagation.jl:89-123) on the GPU engine.

    Without ``steps`` the integrator is the reference's adaptive Tsit5 with
    the fiber's reltol/abstol (solve_disp_mmf, simulate_disp_mmf.jl:183-207)
    and ``saveat`` samples its continuous extension, as _lab_spectrum samples
    the ODE solution (any M).  With ``steps`` it is the fixed-step
    global-frame RK4IP of _scalar_rk4_output, which stores the input and
    output only."""
    length = float(problem.fiber["L"])
    z_m = _save_positions(saveat, length)
    launch = _forward_input(problem, problem.uω0 if input_spectrum is None else input_spectrum)
    if steps is None:
        plan = plan or Plan.from_problem(problem, precision=precision, scheme="tsit5")
        out, stats, saved = plan.forward_saveat(launch, z_m)
        solver = "Tsit5"
    else:
        if z_m.size != 2:
            raise ValueError("saveat needs the adaptive scheme (steps=None)")
        plan = plan or Plan.from_problem(problem, steps=steps, precision=precision)
        out, stats = plan.forward(launch)
        saved = np.stack([launch, out])
        solver = "RK4IP_fixed"
    if not np.allclose(saved[-1], out, rtol=1e-12, atol=0):
        raise RuntimeError("stored output does not match the solver endpoint")
    return PropagationResult(problem, tuple(float(z) for z in z_m), saved, stats["accepted"],
                             stats["rhs_evaluations"], precision, solver, plan)


def scalar_rk4_output(problem: FiberFieldProblem, steps: int, precision: str = "fp64") -> np.ndarray:
    """_scalar_rk4_output (scalar_crosscheck.jl:8-68) on the GPU engine."""
    if isinstance(steps, bool) or not isinstance(steps, (int, np.integer)) or steps <= 0:
        raise ValueError("steps must be a positive integer")
    if problem.modes != 1:
        raise ValueError("the fixed-step propagation cross-check supports scalar problems only")
    if problem.fiber.get("gain_parameters", 0.0) not in (0, 0.0):
        raise ValueError("the fixed-step propagation cross-check does not support gain")
    return propagate(problem, steps=int(steps), precision=precision).output_spectrum


def metrics(result: PropagationResult) -> dict:
    """_propagation_metrics, propagation.jl:204-237, from the device FP64
    evidence of every saved spectrum (fh_field_metrics)."""
    sim = result.problem.sim
    samples = result.plan.field_metrics(result.saved, sim)
    e0, p0 = samples[0]["energy_pj"], samples[0]["photon_number"]
    return dict(
        samples=result.problem.nt, modes=result.problem.modes, saved_positions=len(result.z_m),
        length_m=float(result.problem.fiber["L"]),
        input_energy_pj=e0, output_energy_pj=samples[-1]["energy_pj"],
        input_mode_energy_pj=samples[0]["mode_energy_pj"], output_mode_energy_pj=samples[-1]["mode_energy_pj"],
        relative_energy_change=samples[-1]["energy_pj"] / e0 - 1,
        input_peak_power_w=samples[0]["peak_power_w"], output_peak_power_w=samples[-1]["peak_power_w"],
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
    if photon_drift_limit is not None and not (math.isfinite(photon_drift_limit) and photon_drift_limit >= 0):
        raise ValueError("photon_drift_limit must be nonnegative and finite")
    if energy_drift_limit is not None and not (math.isfinite(energy_drift_limit) and energy_drift_limit >= 0):
        raise ValueError("energy_drift_limit must be nonnegative and finite")
    ev = metrics(result)
    finite = bool(np.all(np.isfinite(result.saved)))
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
