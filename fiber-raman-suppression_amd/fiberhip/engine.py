"""Contexts and plans over the libfiberhip C ABI."""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib
from ._lib import FH_PREC_FP32_BULK, FH_PREC_FP64, FH_SCHEME_RK4IP_FIXED, FH_SCHEME_TSIT5, check, lib

PRECISIONS = {"fp64": FH_PREC_FP64, "fp32": FH_PREC_FP32_BULK, "fp32_bulk": FH_PREC_FP32_BULK}
SCHEMES = {"rk4ip": FH_SCHEME_RK4IP_FIXED, "tsit5": FH_SCHEME_TSIT5}
# terminal objectives of gradients without a caller seed (mmf_cost.jl)
OBJECTIVES = {"band_sum": _lib.FH_OBJ_BAND_SUM, "fundamental": _lib.FH_OBJ_FUNDAMENTAL,
              "worst_mode": _lib.FH_OBJ_WORST_MODE}


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    """One HIP device and its stream (fh_context_create)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        check(lib().fh_context_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self) -> None:
        if self.handle:
            lib().fh_context_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


_DEFAULT: Optional[Context] = None


def default_context() -> Context:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


class DeviceArray:
    """A device buffer owned through the C ABI (fh_device_alloc)."""

    def __init__(self, context: Context, nbytes: int):
        self.context = context
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().fh_device_alloc(context.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, context: Context, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(context, a.nbytes)
        check(lib().fh_copy_to_device(context.handle, ctypes.c_void_p(d.ptr), a.ctypes.data_as(ctypes.c_void_p),
                                      a.nbytes))
        return d

    def to_host(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested view is larger than the buffer")
        check(lib().fh_copy_to_host(self.context.handle, out.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_void_p(self.ptr), out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            lib().fh_device_free(self.context.handle, ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):  # pragma: no cover
        try:
            self.free()
        except Exception:
            pass


def _column_major(a, dtype) -> np.ndarray:
    """(nt, m, ...) array -> contiguous column-major copy (t fastest)."""
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).T)


class Plan:
    """An immutable propagation plan (fh_plan_create) for one resolved fiber.

    dispersion (nt, m) rad/m, gamma (m, m, m, m), raman (nt,) complex hR(ω),
    shock (nt,) = fftshift(ωs/ω0), all in FFT order as the reference builds them.
    Gradients checkpoint every `checkpoint_every` steps; `tape_segments`
    checkpoint segments keep the forward sweep's stage tape instead of being
    recomputed (the last ones of a fixed-step solve, the first ones of an
    adaptive one): fewer transforms per gradient for more device memory, with
    bit-identical results.
    """

    def __init__(self, dispersion, gamma, raman, shock, one_m_fr: float, length_m: float, steps: Optional[int],
                 precision: str = "fp64", max_batch: int = 1, context: Optional[Context] = None,
                 band_mask=None, checkpoint_every: int = 50, log_cost: bool = False, scheme: Optional[str] = None,
                 rtol: float = 1e-8, atol: float = 1e-6, max_steps: int = 10000, tape_segments: int = 0,
                 objective: str = "band_sum", objective_tau: float = 50.0):
        desc = self._describe(dispersion, gamma, raman, shock, one_m_fr, length_m, steps, precision, max_batch,
                              band_mask, checkpoint_every, log_cost, scheme, rtol, atol, max_steps, tape_segments,
                              objective, objective_tau)
        self.context = context or default_context()
        h = ctypes.c_void_p()
        check(lib().fh_plan_create(self.context.handle, ctypes.byref(desc), ctypes.byref(h)))
        self.handle = h

    @classmethod
    def check(cls, *args, gradient: bool = False, **kwargs) -> int:
        """Device bytes a plan with these arguments holds (fh_plan_check; host
        only, no device touched), including its adjoint buffers with
        `gradient`; raises FiberHIPError as fh_plan_create would."""
        kwargs.pop("context", None)
        self = cls.__new__(cls)
        desc = self._describe(*args, **kwargs)
        n = ctypes.c_int64()
        check(lib().fh_plan_check(ctypes.byref(desc), int(bool(gradient)), ctypes.byref(n)))
        return n.value

    def _describe(self, dispersion, gamma, raman, shock, one_m_fr: float, length_m: float, steps: Optional[int],
                  precision: str = "fp64", max_batch: int = 1, band_mask=None, checkpoint_every: int = 50,
                  log_cost: bool = False, scheme: Optional[str] = None, rtol: float = 1e-8, atol: float = 1e-6,
                  max_steps: int = 10000, tape_segments: int = 0, objective: str = "band_sum",
                  objective_tau: float = 50.0) -> "_lib.ProblemDesc":
        """Resolve the arguments into an fh_problem_desc (arrays kept alive on self)."""
        if objective not in OBJECTIVES:
            raise ValueError(f"objective must be one of {sorted(OBJECTIVES)}")
        self.objective = objective
        D = np.asarray(dispersion, dtype=np.float64)
        if D.ndim == 1:
            D = D[:, None]
        self.nt, self.modes = D.shape
        self.precision = precision
        if scheme is None:  # a step count selects the fixed-step scheme, none the adaptive one
            scheme = "rk4ip" if steps is not None else "tsit5"
        if scheme not in SCHEMES:
            raise ValueError(f"scheme must be one of {sorted(SCHEMES)}")
        self.scheme = scheme
        self.steps = int(steps or 0)
        self.max_steps = int(max_steps)
        self.rtol, self.atol = float(rtol), float(atol)
        self.length_m = float(length_m)
        self.max_batch = int(max_batch)
        self._D = _column_major(D, np.float64)
        g = np.asarray(gamma, dtype=np.float64).reshape((self.modes,) * 4)
        self._g = np.ascontiguousarray(g.transpose(3, 2, 1, 0))
        self._hr = np.ascontiguousarray(np.asarray(raman, dtype=np.complex128).reshape(-1))
        self._sh = np.ascontiguousarray(np.asarray(shock, dtype=np.float64).reshape(-1))
        self._mask = None if band_mask is None else np.ascontiguousarray(np.asarray(band_mask, dtype=np.uint8))
        self.has_band = band_mask is not None
        desc = _lib.ProblemDesc(
            nt=self.nt, modes=self.modes, precision=PRECISIONS[precision], scheme=SCHEMES[scheme],
            steps=self.steps, max_batch=self.max_batch, length_m=self.length_m,
            dispersion=_dp(self._D), gamma=_dp(self._g), raman=_dp(self._hr.view(np.float64)),
            shock=_dp(self._sh), one_m_fr=float(one_m_fr), checkpoint_every=int(checkpoint_every),
            log_cost=int(bool(log_cost)),
            band_mask=None if self._mask is None else self._mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            rtol=self.rtol, atol=self.atol, max_steps=self.max_steps, tape_segments=int(tape_segments),
            objective=OBJECTIVES[objective], objective_tau=float(objective_tau))
        return desc

    @classmethod
    def from_problem(cls, problem, steps: Optional[int] = None, precision: str = "fp64", max_batch: int = 1,
                     context: Optional[Context] = None, band_mask=None, checkpoint_every: int = 50,
                     log_cost: bool = False, scheme: Optional[str] = None, rtol: Optional[float] = None,
                     atol: Optional[float] = None, max_steps: int = 10000, tape_segments: int = 0,
                     objective: str = "band_sum", objective_tau: float = 50.0) -> "Plan":
        """Plan for a FiberFieldProblem.  With `steps` the fixed-step RK4IP of
        _scalar_rk4_output; without, the adaptive Tsit5 of solve_disp_mmf with the
        fiber's reltol/abstol (defaults 1e-8 / 1e-6, simulate_disp_mmf.jl:186-190)."""
        f = problem.fiber
        from .problem import self_steepening

        rtol = f.get("reltol", 1e-8) if rtol is None else rtol
        atol = f.get("abstol", 1e-6) if atol is None else atol
        return cls(f["Dω"], f["γ"], f["hRω"], self_steepening(problem.sim), f["one_m_fR"], f["L"], steps,
                   precision=precision, max_batch=max_batch, context=context, band_mask=band_mask,
                   checkpoint_every=checkpoint_every, log_cost=log_cost, scheme=scheme, rtol=rtol, atol=atol,
                   max_steps=max_steps, tape_segments=tape_segments, objective=objective,
                   objective_tau=objective_tau)

    def solve_stats(self, batch: int) -> np.ndarray:
        """(batch, 3) accepted, rejected, rhs evaluations of the last call (fh_solve_stats)."""
        out = np.zeros((batch, 3), dtype=np.int64)
        check(lib().fh_solve_stats(self.handle, int(batch), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    def step_sizes(self, batch: int) -> list:
        """Accepted step sizes of each solve of the last adaptive call (fh_step_sizes)."""
        hs = np.zeros((batch, self.max_steps))
        check(lib().fh_step_sizes(self.handle, int(batch), _dp(hs)))
        n = self.solve_stats(batch)[:, 0]
        return [hs[b, : n[b]].copy() for b in range(batch)]

    @property
    def device_bytes(self) -> int:
        n = ctypes.c_int64()
        check(lib().fh_plan_device_bytes(self.handle, ctypes.byref(n)))
        return n.value

    def forward(self, u0) -> tuple[np.ndarray, dict]:
        """Batched forward solve.  u0: (nt, m) or (batch, nt, m) complex launch
        spectra; returns lab-frame outputs of the same shape and the stats."""
        u = np.asarray(u0, dtype=np.complex128)
        single = u.ndim == 2
        if single:
            u = u[None]
        if u.ndim != 3 or u.shape[1:] != (self.nt, self.modes):
            raise ValueError(f"launch has shape {u.shape}, expected (batch, {self.nt}, {self.modes})")
        batch = u.shape[0]
        buf = np.ascontiguousarray(np.transpose(u, (0, 2, 1)))  # [batch][mode][t]
        out = np.empty_like(buf)
        st = _lib.Stats()
        check(lib().fh_forward(self.handle, batch, _dp(buf.view(np.float64)), _dp(out.view(np.float64)),
                               ctypes.byref(st)))
        res = np.transpose(out, (0, 2, 1))
        stats = dict(accepted=st.accepted, rejected=st.rejected, rhs_evaluations=st.rhs_evals,
                     retcode=st.retcode)
        return (res[0] if single else res), stats

    def forward_saveat(self, u0, zsave) -> tuple[np.ndarray, dict, np.ndarray]:
        """Forward solve with dense output (fh_forward_saveat): lab-frame spectra
        at the sorted positions zsave, shape (nz, nt, m) per launch, beside the
        output and stats of forward()."""
        u = np.asarray(u0, dtype=np.complex128)
        single = u.ndim == 2
        if single:
            u = u[None]
        if u.ndim != 3 or u.shape[1:] != (self.nt, self.modes):
            raise ValueError(f"launch has shape {u.shape}, expected (batch, {self.nt}, {self.modes})")
        z = np.ascontiguousarray(zsave, dtype=np.float64).ravel()
        batch = u.shape[0]
        buf = np.ascontiguousarray(np.transpose(u, (0, 2, 1)))  # [batch][mode][t]
        out = np.empty_like(buf)
        saved = np.empty((batch, z.size, self.modes, self.nt), dtype=np.complex128)
        st = _lib.Stats()
        check(lib().fh_forward_saveat(self.handle, batch, _dp(buf.view(np.float64)), _dp(out.view(np.float64)),
                                      int(z.size), _dp(z), _dp(saved.view(np.float64)), ctypes.byref(st)))
        res = np.transpose(out, (0, 2, 1))
        sv = np.transpose(saved, (0, 1, 3, 2))
        stats = dict(accepted=st.accepted, rejected=st.rejected, rhs_evaluations=st.rhs_evals,
                     retcode=st.retcode)
        return (res[0] if single else res), stats, (sv[0] if single else sv)

    def field_metrics(self, spectra, sim: dict) -> list:
        """FP64 evidence of lab-frame spectra (fh_field_metrics; _field_metrics,
        propagation.jl:184-202).  spectra: (nt, m) or (count, nt, m)."""
        sp = np.asarray(spectra, dtype=np.complex128)
        if sp.ndim == 2:
            sp = sp[None]
        if sp.ndim != 3 or sp.shape[1:] != (self.nt, self.modes):
            raise ValueError(f"spectra have shape {sp.shape}, expected (count, {self.nt}, {self.modes})")
        buf = np.ascontiguousarray(np.transpose(sp, (0, 2, 1)))
        omega = np.ascontiguousarray(np.fft.fftshift(np.abs(np.asarray(sim["ωs"], dtype=np.float64))))
        out = np.zeros((sp.shape[0], self.modes + 5))
        check(lib().fh_field_metrics(self.handle, sp.shape[0], _dp(buf.view(np.float64)), _dp(omega),
                                     float(sim["Δt"]), _dp(out)))
        m = self.modes
        return [dict(energy_pj=float(r[m]), mode_energy_pj=r[:m].copy(), peak_power_w=float(r[m + 1]),
                     temporal_edge_fraction=float(r[m + 2]), spectral_edge_fraction=float(r[m + 3]),
                     photon_number=float(r[m + 4])) for r in out]

    def gradient(self, u0, seed=None) -> dict:
        """Discrete-adjoint gradient (fh_gradient).  u0: (nt, m) or (batch, nt, m)
        shaped launches; seed: matching Wirtinger terminal seeds, or None for
        the plan's band objective.  Returns uL, lam0, grad (per bin), cost."""
        u = np.asarray(u0, dtype=np.complex128)
        single = u.ndim == 2
        if single:
            u = u[None]
        if u.ndim != 3 or u.shape[1:] != (self.nt, self.modes):
            raise ValueError(f"launch has shape {u.shape}, expected (batch, {self.nt}, {self.modes})")
        batch = u.shape[0]
        buf = np.ascontiguousarray(np.transpose(u, (0, 2, 1)))
        sd = None
        if seed is not None:
            sv = np.asarray(seed, dtype=np.complex128)
            sv = sv[None] if sv.ndim == 2 else sv
            if sv.shape != u.shape:
                raise ValueError("seed must match the launch shape")
            sd = np.ascontiguousarray(np.transpose(sv, (0, 2, 1)))
        uL = np.empty_like(buf)
        lam0 = np.empty_like(buf)
        grad = np.empty((batch, self.nt))
        cost = np.empty(batch)
        st = _lib.Stats()
        check(lib().fh_gradient(self.handle, batch, _dp(buf.view(np.float64)),
                                None if sd is None else _dp(sd.view(np.float64)), _dp(uL.view(np.float64)),
                                _dp(lam0.view(np.float64)), _dp(grad), _dp(cost), ctypes.byref(st)))
        out = dict(uL=np.transpose(uL, (0, 2, 1)), lam0=np.transpose(lam0, (0, 2, 1)), grad=grad, cost=cost,
                   rhs_evaluations=st.rhs_evals)
        if single:
            out = {k: (v[0] if isinstance(v, np.ndarray) else v) for k, v in out.items()}
        return out

    def gradient_device(self, batch: int, d_u0: int, d_seed: int, d_uL: int, d_lam0: int, d_grad: int,
                        d_cost: int, stream: int = 0, noalloc: bool = False) -> None:
        """Enqueue a gradient on device-resident arrays (pointers as ints; d_seed
        0 for the plan's objective).  noalloc: fh_gradient_device_noalloc, which
        keeps its buffers in the library's fixed per-GPU scratch array."""
        fn = lib().fh_gradient_device_noalloc if noalloc else lib().fh_gradient_device
        check(fn(self.handle, int(batch), ctypes.c_void_p(d_u0), ctypes.c_void_p(d_seed) if d_seed else None,
                 ctypes.c_void_p(d_uL), ctypes.c_void_p(d_lam0), ctypes.c_void_p(d_grad), ctypes.c_void_p(d_cost),
                 ctypes.c_void_p(stream) if stream else None))

    def forward_tape(self, u0) -> "Tape":
        """Forward solve that keeps its tape on the device (fh_forward_tape);
        Tape.adjoint() then sweeps backward without a second forward solve."""
        u = np.asarray(u0, dtype=np.complex128)
        single = u.ndim == 2
        if single:
            u = u[None]
        if u.ndim != 3 or u.shape[1:] != (self.nt, self.modes):
            raise ValueError(f"launch has shape {u.shape}, expected (batch, {self.nt}, {self.modes})")
        batch = u.shape[0]
        buf = np.ascontiguousarray(np.transpose(u, (0, 2, 1)))
        uL = np.empty_like(buf)
        h = ctypes.c_void_p()
        st = _lib.Stats()
        check(lib().fh_forward_tape(self.handle, batch, _dp(buf.view(np.float64)), _dp(uL.view(np.float64)),
                                    ctypes.byref(h), ctypes.byref(st)))
        out = np.transpose(uL, (0, 2, 1))
        stats = dict(accepted=st.accepted, rejected=st.rejected, rhs_evaluations=st.rhs_evals, retcode=st.retcode)
        return Tape(self, h, batch, single, out[0] if single else out, stats)

    def time_gradient_device(self, batch: int, d_u0: int, d_uL: int, d_lam0: int, d_grad: int, d_cost: int,
                             launches: int) -> np.ndarray:
        ms = (ctypes.c_float * launches)()
        check(lib().fh_time_gradient_device(self.handle, int(batch), ctypes.c_void_p(d_u0), ctypes.c_void_p(d_uL),
                                            ctypes.c_void_p(d_lam0), ctypes.c_void_p(d_grad),
                                            ctypes.c_void_p(d_cost), int(launches), ms))
        return np.array(ms[:], dtype=np.float64)

    def forward_device(self, batch: int, d_u0: int, d_uL: int, stream: int = 0) -> None:
        """Enqueue a forward solve on device-resident fields (pointers as ints)."""
        check(lib().fh_forward_device(self.handle, int(batch), ctypes.c_void_p(d_u0), ctypes.c_void_p(d_uL),
                                      ctypes.c_void_p(stream) if stream else None))

    @property
    def field_dtype(self):
        return np.complex128 if self.precision == "fp64" else np.complex64

    def to_device(self, u) -> DeviceArray:
        """(batch, nt, m) complex -> device buffer in the plan's precision and layout."""
        u = np.asarray(u)
        if u.ndim == 2:
            u = u[None]
        buf = np.ascontiguousarray(np.transpose(u, (0, 2, 1))).astype(self.field_dtype)
        return DeviceArray.from_host(self.context, buf)

    def from_device(self, d: DeviceArray, batch: int) -> np.ndarray:
        a = d.to_host(self.field_dtype, (batch, self.modes, self.nt))
        return np.transpose(a, (0, 2, 1)).astype(np.complex128)

    def time_forward_device(self, batch: int, d_u0: int, d_uL: int, launches: int) -> np.ndarray:
        ms = (ctypes.c_float * launches)()
        check(lib().fh_time_forward_device(self.handle, int(batch), ctypes.c_void_p(d_u0), ctypes.c_void_p(d_uL),
                                           int(launches), ms))
        return np.array(ms[:], dtype=np.float64)

    def synchronize(self) -> None:
        check(lib().fh_synchronize(self.context.handle))

    def energy_device(self, batch: int, d_u: int) -> np.ndarray:
        out = np.zeros(batch * self.modes)
        check(lib().fh_energy_device(self.handle, int(batch), ctypes.c_void_p(d_u), _dp(out)))
        return out.reshape(batch, self.modes)

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().fh_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class Tape:
    """A forward solve kept on the device (fh_forward_tape) for later backward
    sweeps (fh_adjoint), the cached forward of the reference's adjoint
    contract (_cached_forward!, physics_models.jl:799-819)."""

    def __init__(self, plan: Plan, handle, batch: int, single: bool, output: np.ndarray, stats: dict):
        self.plan, self.handle, self.batch, self.single = plan, handle, batch, single
        self.output, self.stats = output, stats

    def adjoint(self, seed=None) -> dict:
        """Backward sweep with Wirtinger terminal seeds dJ/dconj(uL) (shape of
        the launch), or None for the plan's objective computed on the device.
        Returns lam0, grad_phase = 2 Re(conj(lam0) i u0), grad_amplitude =
        2 Re(conj(lam0) u0) (per mode), energy_dot = Re sum conj(lam0) u0 and
        cost (NaN with a seed)."""
        if not self.handle:
            raise ValueError("tape was freed")
        p, B = self.plan, self.batch
        sd = None
        if seed is not None:
            sv = np.asarray(seed, dtype=np.complex128)
            sv = sv[None] if sv.ndim == 2 else sv
            if sv.shape != (B, p.nt, p.modes):
                raise ValueError("seed must match the launch shape")
            sd = np.ascontiguousarray(np.transpose(sv, (0, 2, 1)))
        lam0 = np.empty((B, p.modes, p.nt), dtype=np.complex128)
        gph = np.empty((B, p.modes, p.nt))
        gam = np.empty((B, p.modes, p.nt))
        edot = np.empty(B)
        cost = np.empty(B)
        st = _lib.Stats()
        check(lib().fh_adjoint(self.handle, None if sd is None else _dp(sd.view(np.float64)),
                               _dp(lam0.view(np.float64)), _dp(gph), _dp(gam), _dp(edot), _dp(cost),
                               ctypes.byref(st)))
        out = dict(lam0=np.transpose(lam0, (0, 2, 1)), grad_phase=np.transpose(gph, (0, 2, 1)),
                   grad_amplitude=np.transpose(gam, (0, 2, 1)), energy_dot=edot, cost=cost)
        if self.single:
            out = {k: v[0] for k, v in out.items()}
        return out

    @property
    def device_bytes(self) -> int:
        n = ctypes.c_int64()
        check(lib().fh_tape_device_bytes(self.handle, ctypes.byref(n)))
        return n.value

    def free(self) -> None:
        if getattr(self, "handle", None):
            lib().fh_tape_free(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            if self.plan.handle:
                self.free()
        except Exception:
            pass
