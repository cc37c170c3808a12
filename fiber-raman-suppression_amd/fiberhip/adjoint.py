"""The adjoint contract over libfiberhip (mirror of native_execution.jl /
adjoints.jl / spectral_shaper_model).

* control maps: ``FullGridPhase`` (one phase per frequency bin, identity
  decode/pullback, adjoints.jl:125-139) and ``PhaseBasis`` (phase = B c,
  pullback B^T g accumulated in FP64, adjoints.jl:215-228);
* objectives: ``raman_band_objective`` / ``spectral_band_objective`` (band
  energy fraction, physics_models.jl:851-863, physics_helpers.jl:222-232) and
  ``ObjectiveMap`` (caller cost + terminal seed, adjoints.jl:592-651);
* ``fiber_model(problem)`` -> ``AdjointModel`` whose forward and
  physical_gradient run on the GPU (spectral_shaper_model,
  physics_models.jl:1022-1081);
* ``run_adjoint_step`` and ``check_adjoint_gradient`` (native_execution.jl:
  315-349, 598-666).

The gradient is the exact gradient of the discrete fixed-step RK4IP map
(discrete adjoint), not the reference's continuous adjoint; the two agree to
the integrator's truncation error.
"""
from __future__ import annotations

import dataclasses
from typing import Callable, Optional

import numpy as np

from .engine import Plan
from .problem import FiberFieldProblem, frequency_band_mask


@dataclasses.dataclass
class FullGridPhase:
    problem: FiberFieldProblem

    def dimension(self) -> int:
        return self.problem.nt

    def decode(self, x) -> np.ndarray:
        x = np.asarray(x, dtype=np.float64)
        if x.shape != (self.problem.nt,) or not np.all(np.isfinite(x)):
            raise ValueError(f"phase must be {self.problem.nt} finite values")
        return x

    def pullback(self, g) -> np.ndarray:
        return np.asarray(g, dtype=np.float64)


@dataclasses.dataclass
class PhaseBasis:
    basis: np.ndarray

    def dimension(self) -> int:
        return self.basis.shape[1]

    def decode(self, c) -> np.ndarray:
        c = np.asarray(c, dtype=np.float64)
        if c.shape != (self.basis.shape[1],) or not np.all(np.isfinite(c)):
            raise ValueError(f"coefficients must be {self.basis.shape[1]} finite values")
        return self.basis @ c

    def pullback(self, g) -> np.ndarray:
        return self.basis.T.astype(np.float64) @ np.asarray(g, dtype=np.float64)


@dataclasses.dataclass
class ObjectiveMap:
    """cost(field) -> float and terminal_adjoint(field) -> dJ/dconj(field)."""
    name: str
    cost: Callable
    terminal_adjoint: Callable
    band_mask: Optional[np.ndarray] = None
    log_cost: bool = False


def spectral_band_objective(problem: FiberFieldProblem, band, log_cost: bool = False) -> ObjectiveMap:
    mask = frequency_band_mask(problem, band)
    return _band_objective("spectral_band", mask, log_cost)


def raman_band_objective(problem: FiberFieldProblem, log_cost: bool = False) -> ObjectiveMap:
    if problem.band_mask is None:
        raise ValueError("problem has no Raman band mask; pass band_mask or raman_threshold_thz")
    return _band_objective("raman_band", problem.band_mask, log_cost)


def _band_cost(u, mask, log_cost):
    """_field_spectral_band_cost + _maybe_log_cost (physics_helpers.jl:222-232, 320-326)."""
    total = np.sum(np.abs(u) ** 2)
    if not total > 0:
        raise ValueError("field must have nonzero spectral energy")
    cost = np.sum(np.abs(u[mask, :]) ** 2) / total
    seed = u * (mask[:, None] - cost) / total
    if log_cost:
        scale = 0.0 if cost <= 1e-15 else 10.0 / (np.log(10.0) * cost)
        return 10.0 * np.log10(max(cost, 1e-15)), scale * seed
    return float(cost), seed


def _band_objective(name, mask, log_cost):
    mask = np.asarray(mask, dtype=bool)
    return ObjectiveMap(name, cost=lambda u: _band_cost(u, mask, log_cost)[0],
                        terminal_adjoint=lambda u: _band_cost(u, mask, log_cost)[1], band_mask=mask,
                        log_cost=log_cost)


def _mode_energies(u, mask):
    u = np.asarray(u)
    mask = np.asarray(mask, dtype=bool)
    return np.sum(np.abs(u) ** 2, axis=0), np.sum(np.abs(u[mask, :]) ** 2, axis=0)


def mmf_cost_sum(u, band_mask):
    """mmf_cost_sum, mmf_cost.jl:49-62: J = sum_m E_band_m / sum_m E_m and dJ/dconj(u)."""
    et, eb = _mode_energies(u, band_mask)
    if not np.sum(et) > 0:
        raise ValueError("field must have nonzero spectral energy")
    J = float(np.sum(eb) / np.sum(et))
    return J, u * (np.asarray(band_mask, dtype=float)[:, None] - J) / np.sum(et)


def mmf_cost_fundamental(u, band_mask):
    """mmf_cost_fundamental, mmf_cost.jl:81-98: the band fraction of mode 1 only."""
    et, eb = _mode_energies(u, band_mask)
    if not et[0] > 0:
        raise ValueError("fundamental mode has zero energy")
    J = float(eb[0] / et[0])
    seed = np.zeros_like(np.asarray(u, dtype=np.complex128))
    seed[:, 0] = u[:, 0] * (np.asarray(band_mask, dtype=float) - J) / et[0]
    return J, seed


def mmf_cost_worst_mode(u, band_mask, tau: float = 50.0):
    """mmf_cost_worst_mode, mmf_cost.jl:124-166: log-sum-exp proxy of the worst
    mode's band fraction over the modes with energy, and its softmax-weighted seed."""
    if not (np.isfinite(tau) and tau > 0):
        raise ValueError("tau must be positive and finite")
    et, eb = _mode_energies(u, band_mask)
    active = np.flatnonzero(et > 0)
    if active.size == 0:
        raise ValueError("worst-mode cost requires a nonzero-energy mode")
    r = np.where(et > 0, eb / np.where(et > 0, et, 1.0), 0.0)
    rmax = float(np.max(r[active]))
    shifted = tau * (r[active] - rmax)
    denom = float(np.sum(np.exp(shifted)))
    J = rmax + (np.log(denom) - np.log(active.size)) / tau
    seed = np.zeros_like(np.asarray(u, dtype=np.complex128))
    m = np.asarray(band_mask, dtype=float)
    for w, k in zip(np.exp(shifted) / denom, active):
        seed[:, k] = w * u[:, k] * (m - r[k]) / et[k]
    return float(J), seed


def mmf_objective(problem: FiberFieldProblem, kind: str = "sum", band=None, tau: float = 50.0) -> ObjectiveMap:
    """Multimode Raman-band objective (mmf_cost.jl): kind "sum", "fundamental"
    or "worst_mode", over the problem's Raman band or `band` (THz)."""
    mask = problem.band_mask if band is None else frequency_band_mask(problem, band)
    if mask is None:
        raise ValueError("problem has no Raman band mask; pass band")
    mask = np.asarray(mask, dtype=bool)
    fns = {"sum": lambda u: mmf_cost_sum(u, mask), "fundamental": lambda u: mmf_cost_fundamental(u, mask),
           "worst_mode": lambda u: mmf_cost_worst_mode(u, mask, tau)}
    if kind not in fns:
        raise ValueError(f"kind must be one of {sorted(fns)}")
    f = fns[kind]
    obj = ObjectiveMap(f"mmf_{kind}", cost=lambda u: f(u)[0], terminal_adjoint=lambda u: f(u)[1], band_mask=mask)
    obj.device_objective = "band_sum" if kind == "sum" else kind
    obj.tau = tau
    return obj


def _field_matrix(value, nt: int, m: int, label: str):
    """_phase_matrix / _field_matrix, physics_models.jl:690-730: a scalar, an
    (nt,) vector repeated over modes or an (nt, m) matrix, and its shape tag."""
    a = np.asarray(value, dtype=np.float64)
    if a.ndim == 0:
        out, shape = np.full((nt, m), float(a)), "scalar"
    elif a.ndim == 1:
        if a.shape != (nt,):
            raise ValueError(f"{label} vector length {a.shape[0]} does not match Nt={nt}")
        out, shape = np.repeat(a[:, None], m, axis=1), "vector"
    elif a.ndim == 2:
        if a.shape != (nt, m):
            raise ValueError(f"{label} matrix size {a.shape} does not match ({nt}, {m})")
        out, shape = a.copy(), "matrix"
    else:
        raise ValueError(f"{label} must be a real scalar, vector, or matrix")
    if not np.all(np.isfinite(out)):
        raise ValueError(f"{label} values must be finite")
    return out, shape


def _gradient_for_shape(g, shape: str):
    """_physical_gradient_for_shape, physics_models.jl:732-737."""
    if shape == "scalar":
        return np.array([np.sum(g)])
    if shape == "vector":
        return np.sum(g, axis=1)
    return g


class AdjointModel:
    """GPU spectral-shaper model (spectral_shaper_model, physics_models.jl:1022-1081).

    The launch is alpha A cis(phi) uw0 with alpha = sqrt(energy) for a decoded
    control that is a phase array or a dict with "phase", "amplitude" and
    "energy" (_single_mode_fields, physics_models.jl:751-773).  forward() keeps
    the solve on the device as a tape (_cached_forward!, :799-819); the
    physical gradient sweeps backward over that tape, so a forward plus a
    gradient costs one forward sweep, and returns the phase, amplitude and
    energy gradients of physics_models.jl:1042-1047 for the keys the control
    has (the phase gradient alone for an array control).

    tape_segments: checkpoint segments whose stage tape the forward sweep keeps
    (fewer recomputations, more device memory; gradients are bit-identical).
    None keeps as many as fit in TAPE_BUDGET_BYTES (one gradient per call)."""

    TAPE_BUDGET_BYTES = 4 << 30

    def __init__(self, problem: FiberFieldProblem, steps: int = 2000, precision: str = "fp64",
                 checkpoint_every: int = 50, objective: Optional[ObjectiveMap] = None,
                 tape_segments: Optional[int] = None):
        g = np.asarray(problem.fiber["γ"])
        if not _gamma_symmetric(g):
            raise ValueError("fiber adjoint requires a fully permutation-symmetric gamma tensor; "
                             "use propagate(problem) for forward-only nonsymmetric models")
        self.problem = problem
        self.steps = steps
        self.precision = precision
        mask = objective.band_mask if objective is not None else problem.band_mask
        if tape_segments is None:
            tape_segments = self.default_tape_segments(problem, steps, precision, checkpoint_every)
        self.tape_segments = tape_segments
        self.plan = Plan.from_problem(problem, steps=steps, precision=precision, band_mask=mask,
                                      checkpoint_every=checkpoint_every, tape_segments=tape_segments,
                                      log_cost=bool(objective.log_cost) if objective is not None else False,
                                      objective=getattr(objective, "device_objective", "band_sum"),
                                      objective_tau=getattr(objective, "tau", 50.0))
        self._cache = None  # (signature, tape, fields)

    @classmethod
    def default_tape_segments(cls, problem: FiberFieldProblem, steps: int, precision: str,
                              checkpoint_every: int) -> int:
        """Kept segments whose regions (plus the shared recomputation region)
        fit in TAPE_BUDGET_BYTES: a stage of the tape holds u (complex) and
        h (real, single mode) or u of every mode and ceil(Q/2) complex h pairs
        (multimode kernels, which also serve one FP64 mode at 32768 samples),
        4 stages per step."""
        esz = 8 if precision == "fp64" else 4
        nt, m = problem.nt, problem.modes
        # the plan's rule (fh_plan_create): checkpoint_every = 0 means 50,
        # and the spacing never exceeds the step count
        every = min(int(checkpoint_every) if int(checkpoint_every) > 0 else 50, int(steps))
        q = m * (m + 1) // 2
        multimode = m > 1 or (precision == "fp64" and nt > 16384)
        per_stage = nt * (m + (q + 1) // 2) * 2 * esz if multimode else nt * 3 * esz
        per_segment = 4 * every * per_stage
        nseg = -(-int(steps) // every)
        return int(max(0, min(nseg, cls.TAPE_BUDGET_BYTES // per_segment - 1)))

    def fields(self, decoded) -> dict:
        """Phase, amplitude and energy of a decoded control (_single_mode_fields)."""
        nt, m = self.problem.nt, self.problem.modes
        get = (lambda k, d: decoded.get(k, d)) if isinstance(decoded, dict) else (
            lambda k, d: decoded if k == "phase" else d)
        phase, phase_shape = _field_matrix(get("phase", np.zeros((nt, m))), nt, m, "phase")
        amplitude, amplitude_shape = _field_matrix(get("amplitude", 1.0), nt, m, "amplitude")
        if not np.all(amplitude > 0):
            raise ValueError("fiber field amplitude control must be strictly positive")
        energy = float(get("energy", 1.0))
        if not (np.isfinite(energy) and energy > 0):
            raise ValueError("fiber field energy control must be positive and finite")
        return dict(phase=phase, phase_shape=phase_shape, amplitude=amplitude, amplitude_shape=amplitude_shape,
                    energy=energy, alpha=np.sqrt(energy))

    def shaped_input(self, decoded) -> np.ndarray:
        """_shaped_input, physics_models.jl:791-797: alpha A cis(phi) uw0."""
        f = decoded if isinstance(decoded, dict) and "alpha" in decoded else self.fields(decoded)
        shaped = f["alpha"] * f["amplitude"] * np.exp(1j * f["phase"]) * self.problem.uω0
        if not (np.all(np.isfinite(shaped)) and np.linalg.norm(shaped) > 0):
            raise ValueError("decoded controls must produce a finite, nonzero launch field")
        return shaped

    def _cached_forward(self, decoded):
        f = self.fields(decoded)
        sig = (f["phase"].tobytes(), f["amplitude"].tobytes(), f["energy"])
        if self._cache is None or self._cache[0] != sig:
            if self._cache is not None:
                self._cache[1].free()
                self._cache = None
            self._cache = (sig, self.plan.forward_tape(self.shaped_input(f)), f)
        return self._cache[1], self._cache[2]

    def forward(self, decoded) -> np.ndarray:
        return self._cached_forward(decoded)[0].output

    def physical_gradient(self, decoded, terminal_seed):
        tape, f = self._cached_forward(decoded)
        res = tape.adjoint(terminal_seed)
        phase_g = res["grad_phase"]
        if not isinstance(decoded, dict):
            return _gradient_for_shape(phase_g, f["phase_shape"])
        out = {}
        if "phase" in decoded:
            out["phase"] = _gradient_for_shape(phase_g, f["phase_shape"])
        if "amplitude" in decoded:
            out["amplitude"] = _gradient_for_shape(res["grad_amplitude"] / f["amplitude"], f["amplitude_shape"])
        if "energy" in decoded:
            out["energy"] = np.array([float(res["energy_dot"]) / f["energy"]])
        return out

    def device_objective_gradient(self, decoded) -> tuple[float, np.ndarray]:
        """Cost and phase gradient of the plan's objective computed on the
        device from the kept output (no host seed)."""
        tape, f = self._cached_forward(decoded)
        res = tape.adjoint(None)
        return float(res["cost"]), _gradient_for_shape(res["grad_phase"], f["phase_shape"])

    def close(self) -> None:
        if self._cache is not None:
            self._cache[1].free()
            self._cache = None
        self.plan.close()


def _gamma_symmetric(g) -> bool:
    """_adjoint_gamma_symmetric, sensitivity_disp_mmf.jl:103-112."""
    tol = 1e-12 * max(1.0, float(np.max(np.abs(g))))
    return bool(np.allclose(g, np.transpose(g, (1, 0, 2, 3)), rtol=1e-12, atol=tol)
                and np.allclose(g, np.transpose(g, (0, 2, 1, 3)), rtol=1e-12, atol=tol)
                and np.allclose(g, np.transpose(g, (0, 1, 3, 2)), rtol=1e-12, atol=tol))


def fiber_model(problem: FiberFieldProblem, steps: int = 2000, precision: str = "fp64",
                checkpoint_every: int = 50, objective: Optional[ObjectiveMap] = None,
                tape_segments: Optional[int] = None) -> AdjointModel:
    return AdjointModel(problem, steps, precision, checkpoint_every, objective, tape_segments)


@dataclasses.dataclass
class AdjointStepResult:
    cost: float
    final_state: np.ndarray
    terminal_seed: np.ndarray
    physical_gradient: np.ndarray
    control_gradient: np.ndarray


def run_adjoint_step(model: AdjointModel, control, objective: ObjectiveMap, coordinates) -> AdjointStepResult:
    """native_execution.jl:315-349: decode, forward, cost, seed, adjoint, pullback."""
    phase = control.decode(coordinates)
    final = model.forward(phase)
    cost = float(objective.cost(final))
    seed = objective.terminal_adjoint(final)
    g = model.physical_gradient(phase, seed)
    return AdjointStepResult(cost, final, seed, g, control.pullback(g))


def band_gradient(model: AdjointModel, control, coordinates) -> tuple[float, np.ndarray]:
    """One fused device call: forward, band cost and seed on the GPU, adjoint,
    chain rule; the plan's band objective replaces the host cost/seed."""
    if not model.plan.has_band:
        raise ValueError("the model's plan has no band objective")
    res = model.plan.gradient(model.shaped_input(control.decode(coordinates)))
    return float(res["cost"]), control.pullback(res["grad"])


@dataclasses.dataclass
class AdjointGradientCheckResult:
    pass_: bool
    coordinate_indices: np.ndarray
    adjoint_gradient: np.ndarray
    finite_difference_gradient: np.ndarray
    absolute_error: np.ndarray
    relative_error: np.ndarray


def check_adjoint_gradient(model: AdjointModel, control, objective: ObjectiveMap, coordinates, step=1e-6,
                           atol=1e-6, rtol=1e-4, coordinate_indices=None) -> AdjointGradientCheckResult:
    """native_execution.jl:598-666: centred differences of the cost at chosen
    coordinates (0-based indices here) against the adjoint gradient."""
    x = np.asarray(coordinates, dtype=np.float64)
    if not (np.isfinite(step) and step > 0):
        raise ValueError("gradient check step must be positive and finite")
    if not (np.isfinite(atol) and atol >= 0 and np.isfinite(rtol) and rtol >= 0):
        raise ValueError("gradient check tolerances must be non-negative and finite")
    n = control.dimension()
    idx = np.arange(n) if coordinate_indices is None else np.asarray(coordinate_indices, dtype=int)
    adj = run_adjoint_step(model, control, objective, x).control_gradient
    fd = np.empty(len(idx))
    for j, i in enumerate(idx):
        xp, xm = x.copy(), x.copy()
        xp[i] += step
        xm[i] -= step
        fd[j] = (objective.cost(model.forward(control.decode(xp)))
                 - objective.cost(model.forward(control.decode(xm)))) / (2 * step)
    sel = adj[idx]
    err = np.abs(sel - fd)
    rel = err / np.maximum(np.abs(fd), atol)
    return AdjointGradientCheckResult(bool(np.all(err <= atol + rtol * np.abs(fd))), idx, sel, fd, err, rel)
