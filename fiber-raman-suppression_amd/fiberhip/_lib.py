"""ctypes binding of libfiberhip (include/fiberhip.h).

The shared library is built in-tree next to this package
(``fiber-raman-suppression_amd/libfiberhip.so``).  There is no fallback: if
the library is missing, every entry point raises ``FiberHIPUnavailable``.
"""
from __future__ import annotations

import ctypes
import pathlib

PKG_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG_DIR.parent / "libfiberhip.so"
HEADER_PATH = PKG_DIR.parent.parent / "include" / "fiberhip.h"

FH_OK = 0
FH_EINVAL = 1
FH_EGRID = 2
FH_ENOTSYM = 3
FH_ENONFINITE = 4
FH_ENOMEM = 5
FH_EDEVICE = 6
FH_EUNSUPPORTED = 7
FH_EMAXSTEPS = 8

FH_PREC_FP64 = 0
FH_PREC_FP32_BULK = 1
FH_SCHEME_RK4IP_FIXED = 0
FH_SCHEME_TSIT5 = 1
FH_OBJ_BAND_SUM = 0
FH_OBJ_FUNDAMENTAL = 1
FH_OBJ_WORST_MODE = 2

_NAMES = {
    FH_EINVAL: "FH_EINVAL", FH_EGRID: "FH_EGRID", FH_ENOTSYM: "FH_ENOTSYM",
    FH_ENONFINITE: "FH_ENONFINITE", FH_ENOMEM: "FH_ENOMEM", FH_EDEVICE: "FH_EDEVICE",
    FH_EUNSUPPORTED: "FH_EUNSUPPORTED", FH_EMAXSTEPS: "FH_EMAXSTEPS",
}


class FiberHIPUnavailable(RuntimeError):
    """libfiberhip.so is not built or cannot be loaded."""


class FiberHIPError(RuntimeError):
    """A libfiberhip entry point returned an error code."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{_NAMES.get(code, code)}: {message}")
        self.code = code
        self.message = message


class ProblemDesc(ctypes.Structure):
    _fields_ = [
        ("nt", ctypes.c_int32),
        ("modes", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("scheme", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("max_batch", ctypes.c_int32),
        ("length_m", ctypes.c_double),
        ("dispersion", ctypes.POINTER(ctypes.c_double)),
        ("gamma", ctypes.POINTER(ctypes.c_double)),
        ("raman", ctypes.POINTER(ctypes.c_double)),
        ("shock", ctypes.POINTER(ctypes.c_double)),
        ("one_m_fr", ctypes.c_double),
        ("checkpoint_every", ctypes.c_int32),
        ("log_cost", ctypes.c_int32),
        ("band_mask", ctypes.POINTER(ctypes.c_uint8)),
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("max_steps", ctypes.c_int32),
        ("tape_segments", ctypes.c_int32),
        ("objective", ctypes.c_int32),
        ("objective_tau", ctypes.c_double),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("accepted", ctypes.c_int64),
        ("rejected", ctypes.c_int64),
        ("rhs_evals", ctypes.c_int64),
        ("retcode", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("memory_budget_bytes", ctypes.c_int64),
        ("lanes", ctypes.c_int32),
        ("max_batch", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("max_steps", ctypes.c_int32),
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("multimode_step_m", ctypes.c_double),
        ("mm_grad_max_steps", ctypes.c_int32),
        ("tape_segments", ctypes.c_int32),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("nt", ctypes.c_int32),
        ("modes", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("length_m", ctypes.c_double),
        ("dispersion", ctypes.POINTER(ctypes.c_double)),
        ("gamma", ctypes.POINTER(ctypes.c_double)),
        ("raman", ctypes.POINTER(ctypes.c_double)),
        ("shock", ctypes.POINTER(ctypes.c_double)),
        ("one_m_fr", ctypes.c_double),
        ("band_mask", ctypes.POINTER(ctypes.c_uint8)),
        ("u0", ctypes.POINTER(ctypes.c_double)),
        ("uL", ctypes.POINTER(ctypes.c_double)),
        ("grad", ctypes.POINTER(ctypes.c_double)),
        ("cost", ctypes.POINTER(ctypes.c_double)),
    ]


class RequestResult(ctypes.Structure):
    _fields_ = [
        ("submit_ns", ctypes.c_int64),
        ("start_ns", ctypes.c_int64),
        ("finish_ns", ctypes.c_int64),
        ("accepted", ctypes.c_int64),
        ("rejected", ctypes.c_int64),
        ("rhs_evals", ctypes.c_int64),
        ("retcode", ctypes.c_int32),
        ("batch", ctypes.c_int32),
        ("attempts", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class GainDesc(ctypes.Structure):
    _fields_ = [
        ("w12", ctypes.POINTER(ctypes.c_double)),
        ("w21", ctypes.POINTER(ctypes.c_double)),
        ("ga", ctypes.POINTER(ctypes.c_double)),
        ("ge", ctypes.POINTER(ctypes.c_double)),
        ("r12", ctypes.c_double),
        ("r21", ctypes.c_double),
        ("inv_tau", ctypes.c_double),
        ("gpa", ctypes.c_double),
        ("gpe", ctypes.c_double),
    ]


FH_REQ_FORWARD = 0
FH_REQ_GRADIENT = 1

_LIB = None

_VP = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_SIGNATURES = {
    "fh_context_create": ([ctypes.c_int, ctypes.POINTER(_VP)], ctypes.c_int),
    "fh_context_destroy": ([_VP], ctypes.c_int),
    "fh_plan_create": ([_VP, ctypes.POINTER(ProblemDesc), ctypes.POINTER(_VP)], ctypes.c_int),
    "fh_plan_destroy": ([_VP], ctypes.c_int),
    "fh_plan_device_bytes": ([_VP, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "fh_plan_check": ([ctypes.POINTER(ProblemDesc), ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "fh_forward_tape": ([_VP, ctypes.c_int32, _DP, _DP, ctypes.POINTER(_VP), ctypes.POINTER(Stats)], ctypes.c_int),
    "fh_adjoint": ([_VP, _DP, _DP, _DP, _DP, _DP, _DP, ctypes.POINTER(Stats)], ctypes.c_int),
    "fh_tape_device_bytes": ([_VP, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "fh_tape_free": ([_VP], ctypes.c_int),
    "fh_forward": ([_VP, ctypes.c_int32, _DP, _DP, ctypes.POINTER(Stats)], ctypes.c_int),
    "fh_forward_saveat": ([_VP, ctypes.c_int32, _DP, _DP, ctypes.c_int32, _DP, _DP, ctypes.POINTER(Stats)],
                          ctypes.c_int),
    "fh_forward_gain": ([_VP, ctypes.c_int32, ctypes.POINTER(GainDesc), _DP, _DP, _DP, _DP, ctypes.POINTER(Stats)],
                        ctypes.c_int),
    "fh_field_metrics": ([_VP, ctypes.c_int32, _DP, _DP, ctypes.c_double, _DP], ctypes.c_int),
    "fh_fd_modes": ([_VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double, _DP, ctypes.c_int32,
                     ctypes.c_double, ctypes.c_int32, _DP, _DP, ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "fh_overlap_tensor": ([_VP, ctypes.c_int64, ctypes.c_int32, _DP, ctypes.c_double, _DP], ctypes.c_int),
    "fh_forward_device": ([_VP, ctypes.c_int32, _VP, _VP, _VP], ctypes.c_int),
    "fh_energy_device": ([_VP, ctypes.c_int32, _VP, _DP], ctypes.c_int),
    "fh_gradient": ([_VP, ctypes.c_int32, _DP, _DP, _DP, _DP, _DP, _DP, ctypes.POINTER(Stats)], ctypes.c_int),
    "fh_gradient_device": ([_VP, ctypes.c_int32, _VP, _VP, _VP, _VP, _VP, _VP, _VP], ctypes.c_int),
    "fh_gradient_device_noalloc": ([_VP, ctypes.c_int32, _VP, _VP, _VP, _VP, _VP, _VP, _VP], ctypes.c_int),
    "fh_time_gradient_device": ([_VP, ctypes.c_int32, _VP, _VP, _VP, _VP, _VP, ctypes.c_int32,
                                 ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "fh_device_alloc": ([_VP, ctypes.c_int64, ctypes.POINTER(_VP)], ctypes.c_int),
    "fh_device_free": ([_VP, _VP], ctypes.c_int),
    "fh_copy_to_device": ([_VP, _VP, _VP, ctypes.c_int64], ctypes.c_int),
    "fh_copy_to_host": ([_VP, _VP, _VP, ctypes.c_int64], ctypes.c_int),
    "fh_synchronize": ([_VP], ctypes.c_int),
    "fh_device_copy": ([_VP, _VP, _VP, ctypes.c_int64], ctypes.c_int),
    "fh_time_forward_device": ([_VP, ctypes.c_int32, _VP, _VP, ctypes.c_int32, ctypes.POINTER(ctypes.c_float)],
                               ctypes.c_int),
    "fh_solve_stats": ([_VP, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "fh_step_sizes": ([_VP, ctypes.c_int32, _DP], ctypes.c_int),
    "fh_server_create": ([_VP, ctypes.POINTER(ServerConfig), ctypes.POINTER(_VP)], ctypes.c_int),
    "fh_server_submit": ([_VP, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
    "fh_server_poll": ([_VP, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)], ctypes.c_int),
    "fh_server_wait": ([_VP, ctypes.c_int64, ctypes.POINTER(RequestResult)], ctypes.c_int),
    "fh_server_destroy": ([_VP], ctypes.c_int),
    "fh_last_error": ([], ctypes.c_char_p),
}


def exported_names() -> list[str]:
    return list(_SIGNATURES)


def lib() -> ctypes.CDLL:
    """Load libfiberhip.so (raises FiberHIPUnavailable when it is not built)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise FiberHIPUnavailable(
                f"{LIB_PATH} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        try:
            handle = ctypes.CDLL(str(LIB_PATH))
        except OSError as exc:  # pragma: no cover - depends on the ROCm runtime
            raise FiberHIPUnavailable(f"cannot load {LIB_PATH}: {exc}") from exc
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = handle
    return _LIB


def check(code: int) -> None:
    if code != FH_OK:
        msg = lib().fh_last_error()
        raise FiberHIPError(code, msg.decode() if msg else "")
