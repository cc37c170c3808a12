"""Request server (fh_server_* of include/fiberhip.h): continuous batching of
independent forward and forward+adjoint solves of uneven grid size, mode count
and fiber length, admitted by device-memory footprint, with per-request
latency.  The reference runs its solves one at a time in the calling Julia
task; this is the serving loop around the same solver."""
from __future__ import annotations

import ctypes
import threading
from typing import Optional

import numpy as np

from . import _lib
from ._lib import FH_REQ_FORWARD, FH_REQ_GRADIENT, FiberHIPError, check, lib
from .engine import PRECISIONS, Context, default_context


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Server:
    """A libfiberhip request server on one context.

    memory_budget_bytes bounds the device memory of the server's plans; lanes
    is the number of launches in flight (the process needs lanes + 4 hardware
    queues, GPU_MAX_HW_QUEUES; on MI355X keep lanes at 12 or below: with more
    lanes, FP32 launches came back non-finite more often under load); max_batch the
    solves per launch.
    Single-mode requests use the adaptive Tsit5 scheme (rtol, atol), multimode
    requests the fixed-step RK4IP with ceil(L / multimode_step_m) steps (at
    most mm_grad_max_steps for multimode gradients).  Single-mode gradient plans
    keep the stage tape of their first tape_segments segments of 50 accepted
    steps (no recomputation there; more device memory per plan)."""

    def __init__(self, memory_budget_bytes: int = 16 << 30, lanes: int = 4, max_batch: int = 64,
                 precision: str = "fp32", rtol: float = 1e-6, atol: float = 1e-6, multimode_step_m: float = 1e-4,
                 max_steps: int = 10000, mm_grad_max_steps: int = 2000, context: Optional[Context] = None,
                 tape_segments: int = 0):
        self.context = context or default_context()
        self.precision = precision
        cfg = _lib.ServerConfig(memory_budget_bytes=int(memory_budget_bytes), lanes=int(lanes),
                                max_batch=int(max_batch), precision=PRECISIONS[precision], max_steps=int(max_steps),
                                rtol=float(rtol), atol=float(atol), multimode_step_m=float(multimode_step_m),
                                mm_grad_max_steps=int(mm_grad_max_steps), tape_segments=int(tape_segments))
        h = ctypes.c_void_p()
        check(lib().fh_server_create(self.context.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self._pending = {}
        self._lock = threading.Lock()

    def submit(self, problem, kind: str = "forward", u0=None, length_m: Optional[float] = None,
               band_mask=None) -> int:
        """Queue one solve of `problem` (a FiberFieldProblem); u0 defaults to the
        problem's launch, length_m to its fiber length.  Returns the request id."""
        from .problem import self_steepening

        f = problem.fiber
        D = np.asarray(f["Dω"], dtype=np.float64)
        D = D[:, None] if D.ndim == 1 else D
        nt, m = D.shape
        keep = dict(
            D=np.ascontiguousarray(D.T),
            g=np.ascontiguousarray(np.asarray(f["γ"], dtype=np.float64).reshape((m,) * 4).transpose(3, 2, 1, 0)),
            hr=np.ascontiguousarray(np.asarray(f["hRω"], dtype=np.complex128).reshape(-1)),
            sh=np.ascontiguousarray(np.asarray(self_steepening(problem.sim), dtype=np.float64).reshape(-1)),
            u0=np.ascontiguousarray(np.asarray(problem.uω0 if u0 is None else u0, dtype=np.complex128)
                                    .reshape(nt, m).T),
            uL=np.zeros((m, nt), dtype=np.complex128),
        )
        grad = kind == "gradient"
        if grad:
            if band_mask is None:
                raise ValueError("gradient requests need a band mask")
            keep["mask"] = np.ascontiguousarray(np.asarray(band_mask, dtype=np.uint8))
            keep["grad"] = np.zeros(nt)
            keep["cost"] = np.zeros(1)
        req = _lib.Request(
            kind=FH_REQ_GRADIENT if grad else FH_REQ_FORWARD, nt=nt, modes=m,
            length_m=float(f["L"] if length_m is None else length_m),
            dispersion=_dp(keep["D"]), gamma=_dp(keep["g"]), raman=_dp(keep["hr"].view(np.float64)),
            shock=_dp(keep["sh"]), one_m_fr=float(f["one_m_fR"]),
            band_mask=keep["mask"].ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)) if grad else None,
            u0=_dp(keep["u0"].view(np.float64)), uL=_dp(keep["uL"].view(np.float64)),
            grad=_dp(keep["grad"]) if grad else None, cost=_dp(keep["cost"]) if grad else None)
        rid = ctypes.c_int64()
        check(lib().fh_server_submit(self.handle, ctypes.byref(req), ctypes.byref(rid)))
        with self._lock:
            self._pending[rid.value] = keep
        return rid.value

    def poll(self, rid: int) -> int:
        st = ctypes.c_int32()
        check(lib().fh_server_poll(self.handle, int(rid), ctypes.byref(st)))
        return st.value

    def wait(self, rid: int) -> dict:
        """Block until request `rid` finished; returns uL (nt, m), grad and cost
        for gradients, the solve's counters, batch size and latencies (ms)."""
        res = _lib.RequestResult()
        rc = lib().fh_server_wait(self.handle, int(rid), ctypes.byref(res))
        with self._lock:
            keep = self._pending.pop(int(rid))
        out = dict(retcode=res.retcode, accepted=res.accepted, rejected=res.rejected, rhs_evaluations=res.rhs_evals,
                   batch=res.batch, attempts=res.attempts, latency_ms=(res.finish_ns - res.submit_ns) / 1e6,
                   submit_ns=res.submit_ns, start_ns=res.start_ns, finish_ns=res.finish_ns,
                   queue_ms=(res.start_ns - res.submit_ns) / 1e6 if res.start_ns else None,
                   uL=keep["uL"].T)
        if "grad" in keep:
            out["grad"], out["cost"] = keep["grad"], float(keep["cost"][0])
        if rc:
            try:
                check(rc)
            except FiberHIPError as exc:
                exc.result = out  # counters, batch and attempts of the failed request
                raise
        return out

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib().fh_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass
