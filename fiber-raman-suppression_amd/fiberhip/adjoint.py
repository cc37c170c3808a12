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


class AdjointModel:
    """GPU spectral-shaper model: launch alpha*A*cis(phi)*uw0 (here alpha = A = 1).

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
                                      log_cost=bool(objective.log_cost) if objective is not None else False)

    @classmethod
    def default_tape_segments(cls, problem: FiberFieldProblem, steps: int, precision: str,
                              checkpoint_every: int) -> int:
        """Kept segments whose regions (plus the shared recomputation region)
        fit in TAPE_BUDGET_BYTES: a stage of the tape holds u (complex) and
        h (real, single mode) or u of every mode and ceil(Q/2) complex h pairs
        (multimode), 4 stages per step."""
        esz = 8 if precision == "fp64" else 4
        nt, m = problem.nt, problem.modes
        # the plan's rule (fh_plan_create): checkpoint_every = 0 means 50,
        # and the spacing never exceeds the step count
        every = min(int(checkpoint_every) if int(checkpoint_every) > 0 else 50, int(steps))
        q = m * (m + 1) // 2
        per_stage = nt * 3 * esz if m == 1 else nt * (m + (q + 1) // 2) * 2 * esz
        per_segment = 4 * every * per_stage
        nseg = -(-int(steps) // every)
        return int(max(0, min(nseg, cls.TAPE_BUDGET_BYTES // per_segment - 1)))

    def shaped_input(self, phase) -> np.ndarray:
        """_shaped_input, physics_models.jl:791-797 (alpha = A = 1)."""
        return self.problem.uω0 * np.exp(1j * np.asarray(phase))[:, None]

    def forward(self, phase) -> np.ndarray:
        out, _ = self.plan.forward(self.shaped_input(phase))
        return out

    def physical_gradient(self, phase, terminal_seed) -> np.ndarray:
        res = self.plan.gradient(self.shaped_input(phase), seed=terminal_seed)
        return res["grad"]


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
