"""The oracle: FP64 CPU restatements of FiberLab's GMMNLSE hot path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg, where it is the checker (or the timed
CPU baseline), never the thing measured or shipped.  The product package
(libfiberhip and its Python host mirror) never imports it.

Two independent restatements live here:

* ``fiberoracle.c`` (loaded below through ctypes): forward RHS, fixed-step
  global-frame RK4 (scalar_crosscheck.jl:8-68), its discrete adjoint, the
  reference's continuous adjoint RHS, and adaptive Tsit5 (solve_disp_mmf's
  integrator) with the discrete adjoint of its accepted-step map, with a
  self-written radix-2 FFT.
* ``np_ref.py``: numpy restatement of the setup helpers and the same
  RHS / RK4 / VJP, using numpy's FFT.

Pins (tests/test_oracle_pins.py): the reference's closed-form tests and its
published counterfactual numbers; see the test docstrings for file:line.
"""
from __future__ import annotations

import ctypes
import os
import pathlib

import numpy as np

from . import np_ref  # noqa: F401  (re-exported)

_HERE = pathlib.Path(__file__).resolve().parent
_LIB = None


class _Fiber(ctypes.Structure):
    _fields_ = [
        ("nt", ctypes.c_int),
        ("m", ctypes.c_int),
        ("disp", ctypes.POINTER(ctypes.c_double)),
        ("gamma", ctypes.POINTER(ctypes.c_double)),
        ("hr", ctypes.POINTER(ctypes.c_double)),
        ("shock", ctypes.POINTER(ctypes.c_double)),
        ("one_m_fr", ctypes.c_double),
    ]


def build() -> pathlib.Path:
    """Compile liboracle.so in place with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-B", "-C", str(_HERE), "liboracle.so"], check=True)
    return _HERE / "liboracle.so"


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build()
        _LIB = ctypes.CDLL(str(path))
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(_Fiber)
        _LIB.orc_fft.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        _LIB.orc_rhs.argtypes = [fp, ctypes.c_double, dp, dp]
        _LIB.orc_rhs_vjp.argtypes = [fp, ctypes.c_double, dp, dp, dp]
        _LIB.orc_rk4_forward.argtypes = [fp, ctypes.c_double, ctypes.c_int, dp, dp, dp]
        _LIB.orc_rk4_adjoint.argtypes = [fp, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, dp, dp]
        _LIB.orc_adjoint_rhs.argtypes = [fp, ctypes.c_double, dp, dp, dp]
        i64p = ctypes.POINTER(ctypes.c_int64)
        _LIB.orc_tsit5_tableau.argtypes = [dp, dp, dp]
        _LIB.orc_tsit5_forward.argtypes = [fp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           dp, dp, dp, i64p]
        _LIB.orc_tsit5_forward_save.argtypes = [fp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_int, dp, dp, dp, i64p, ctypes.c_int, dp, dp]
        _LIB.orc_tsit5_dense_weights.argtypes = [ctypes.c_double, dp]
        _LIB.orc_tsit5_replay.argtypes = [fp, ctypes.c_double, dp, ctypes.c_int, dp, dp]
        _LIB.orc_tsit5_adjoint.argtypes = [fp, ctypes.c_double, dp, ctypes.c_int, ctypes.c_int, dp, dp, dp]
        for name in ("orc_fft", "orc_rhs", "orc_rhs_vjp", "orc_rk4_forward", "orc_rk4_adjoint",
                     "orc_adjoint_rhs", "orc_tsit5_tableau", "orc_tsit5_forward", "orc_tsit5_replay",
                     "orc_tsit5_adjoint", "orc_tsit5_forward_save", "orc_tsit5_dense_weights"):
            getattr(_LIB, name).restype = ctypes.c_int
    return _LIB


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _field(a) -> np.ndarray:
    """(nt, m) complex -> C-contiguous (m, nt) complex128, i.e. the column-major
    (t fastest) layout of the (nt, m) array."""
    a = np.asarray(a, dtype=np.complex128)
    if a.ndim == 1:
        a = a[:, None]
    return np.ascontiguousarray(a.T)


class Fiber:
    """Holds the resolved operators of one problem in the C layout."""

    def __init__(self, D, gamma, hR, shock, one_m_fR):
        D = np.asarray(D, dtype=np.float64)
        if D.ndim == 1:
            D = D[:, None]
        self.nt, self.m = D.shape
        self.D = D
        self._Dc = np.ascontiguousarray(D.T)
        self.gamma = np.asarray(gamma, dtype=np.float64).reshape((self.m,) * 4)
        self._gc = np.ascontiguousarray(self.gamma.transpose(3, 2, 1, 0))  # column-major gamma[i,j,k,l]
        self.hR = np.ascontiguousarray(np.asarray(hR, dtype=np.complex128).reshape(-1))
        self.shock = np.ascontiguousarray(np.asarray(shock, dtype=np.float64).reshape(-1))
        self.one_m_fR = float(one_m_fR)
        self._s = _Fiber(self.nt, self.m, _dp(self._Dc), _dp(self._gc), _dp(self.hR.view(np.float64)),
                         _dp(self.shock), self.one_m_fR)

    @property
    def ref(self):
        return ctypes.byref(self._s)


def _check(rc: int, what: str):
    if rc == -2:
        raise FloatingPointError(f"{what}: non-finite result")
    if rc != 0:
        raise RuntimeError(f"{what} failed (rc={rc})")


def fft(x, inverse=False) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    _check(lib().orc_fft(_dp(a.view(np.float64)), a.size, int(inverse)), "orc_fft")
    return a


def rhs(fib: Fiber, z: float, x) -> np.ndarray:
    x = _field(x)
    out = np.zeros_like(x)
    _check(lib().orc_rhs(fib.ref, z, _dp(x.view(np.float64)), _dp(out.view(np.float64))), "orc_rhs")
    return out.T


def rhs_vjp(fib: Fiber, z: float, x, gy) -> np.ndarray:
    x, gy = _field(x), _field(gy)
    out = np.zeros_like(x)
    _check(lib().orc_rhs_vjp(fib.ref, z, _dp(x.view(np.float64)), _dp(gy.view(np.float64)),
                             _dp(out.view(np.float64))), "orc_rhs_vjp")
    return out.T


def rk4_forward(fib: Fiber, length: float, steps: int, u0, trajectory=False):
    """Fixed-step global-frame RK4 (scalar_crosscheck.jl:8-68, any M) -> lab spectrum."""
    u0 = _field(u0)
    out = np.zeros_like(u0)
    traj = np.zeros((steps, fib.m, fib.nt), dtype=np.complex128) if trajectory else None
    tp = _dp(traj.view(np.float64)) if trajectory else None
    _check(lib().orc_rk4_forward(fib.ref, length, steps, _dp(u0.view(np.float64)),
                                 _dp(out.view(np.float64)), tp), "orc_rk4_forward")
    if trajectory:
        return out.T, np.transpose(traj, (2, 1, 0))
    return out.T


def rk4_adjoint(fib: Fiber, length: float, steps: int, u0, lam_out, every: int = 50) -> np.ndarray:
    """Discrete adjoint of rk4_forward: Wirtinger seed at the lab output -> seed at u0."""
    u0, lo = _field(u0), _field(lam_out)
    out = np.zeros_like(u0)
    _check(lib().orc_rk4_adjoint(fib.ref, length, steps, every, _dp(u0.view(np.float64)),
                                 _dp(lo.view(np.float64)), _dp(out.view(np.float64))), "orc_rk4_adjoint")
    return out.T


def adjoint_rhs(fib: Fiber, z: float, lam, ufwd) -> np.ndarray:
    """adjoint_disp_mmf! restated (sensitivity_disp_mmf.jl:32-79)."""
    lam, ufwd = _field(lam), _field(ufwd)
    out = np.zeros_like(lam)
    _check(lib().orc_adjoint_rhs(fib.ref, z, _dp(lam.view(np.float64)), _dp(ufwd.view(np.float64)),
                                 _dp(out.view(np.float64))), "orc_adjoint_rhs")
    return out.T


def continuous_adjoint(fib: Fiber, length: float, steps: int, u0, lam_out) -> np.ndarray:
    """Backward RK4 of the reference's continuous adjoint ODE (solve_adjoint_disp_mmf,
    sensitivity_disp_mmf.jl:309-332) with the forward state taken from a forward
    RK4 run at half the step (so every stage position is a stored state)."""
    u0 = np.asarray(u0, dtype=np.complex128).reshape(fib.nt, fib.m)
    _, traj = rk4_forward(fib, length, 2 * steps, u0, trajectory=True)
    states = [u0] + [traj[:, :, i] for i in range(2 * steps)]
    h = length / steps
    lam = np.exp(-1j * fib.D * length) * np.asarray(lam_out, dtype=np.complex128).reshape(fib.nt, fib.m)
    for n in range(steps, 0, -1):
        z = n * h
        f_hi, f_mid, f_lo = states[2 * n], states[2 * n - 1], states[2 * n - 2]
        k1 = adjoint_rhs(fib, z, lam, f_hi)
        k2 = adjoint_rhs(fib, z - h / 2, lam - h / 2 * k1, f_mid)
        k3 = adjoint_rhs(fib, z - h / 2, lam - h / 2 * k2, f_mid)
        k4 = adjoint_rhs(fib, z - h, lam - h * k3, f_lo)
        lam = lam - h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return lam


def tsit5_tableau():
    """(c[7], A[7][6], btilde[7]) of the Tsitouras 5(4) pair used by the oracle."""
    c, a, bt = np.zeros(7), np.zeros((7, 6)), np.zeros(7)
    _check(lib().orc_tsit5_tableau(_dp(c), _dp(a), _dp(bt)), "orc_tsit5_tableau")
    return c, a, bt


def tsit5_dense_weights(theta: float) -> np.ndarray:
    """b_i(theta), i = 1..7, of the Tsit5 continuous extension."""
    b = np.zeros(7)
    _check(lib().orc_tsit5_dense_weights(ctypes.c_double(theta), _dp(b)), "orc_tsit5_dense_weights")
    return b


class Tsit5Result:
    def __init__(self, out, hs, stats, saved=None):
        self.out, self.hs, self.stats, self.saved = out, hs, stats, saved

    @property
    def accepted(self) -> int:
        return int(self.stats[0])

    @property
    def rejected(self) -> int:
        return int(self.stats[1])

    @property
    def rhs_evals(self) -> int:
        return int(self.stats[2])


def tsit5_forward(fib: Fiber, length: float, u0, rtol: float = 1e-6, atol: float = 1e-6,
                  max_steps: int = 100000, saveat=None) -> Tsit5Result:
    """Adaptive Tsit5 (solve_disp_mmf, simulate_disp_mmf.jl:183-207) -> lab spectrum,
    accepted step sizes and {accepted, rejected, rhs evaluations}.  With
    `saveat` (sorted positions in [0, L]) also the lab spectra there,
    shape (nz, nt, m), from the continuous extension."""
    u0 = _field(u0)
    out = np.zeros_like(u0)
    hs = np.zeros(max_steps)
    st = np.zeros(3, dtype=np.int64)
    z = np.ascontiguousarray([] if saveat is None else saveat, dtype=np.float64)
    saved = np.zeros((z.size,) + u0.shape, dtype=np.complex128)
    rc = lib().orc_tsit5_forward_save(fib.ref, length, rtol, atol, max_steps, _dp(u0.view(np.float64)),
                                      _dp(out.view(np.float64)), _dp(hs),
                                      st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(z.size), _dp(z),
                                      _dp(saved.view(np.float64)))
    if rc == -3:
        raise RuntimeError("orc_tsit5_forward: max_steps reached")
    _check(rc, "orc_tsit5_forward")
    return Tsit5Result(out.T, hs[: st[0]].copy(), st,
                       None if saveat is None else np.transpose(saved, (0, 2, 1)).copy())


def tsit5_replay(fib: Fiber, length: float, hs, u0) -> np.ndarray:
    """The accepted-step map of tsit5_forward for a given step sequence."""
    u0 = _field(u0)
    out = np.zeros_like(u0)
    hs = np.ascontiguousarray(hs, dtype=np.float64)
    _check(lib().orc_tsit5_replay(fib.ref, length, _dp(hs), hs.size, _dp(u0.view(np.float64)),
                                  _dp(out.view(np.float64))), "orc_tsit5_replay")
    return out.T


def tsit5_adjoint(fib: Fiber, length: float, hs, u0, lam_out, every: int = 50) -> np.ndarray:
    """Discrete adjoint of tsit5_replay: Wirtinger seed at the lab output -> seed at u0."""
    u0, lo = _field(u0), _field(lam_out)
    out = np.zeros_like(u0)
    hs = np.ascontiguousarray(hs, dtype=np.float64)
    _check(lib().orc_tsit5_adjoint(fib.ref, length, _dp(hs), hs.size, every, _dp(u0.view(np.float64)),
                                   _dp(lo.view(np.float64)), _dp(out.view(np.float64))), "orc_tsit5_adjoint")
    return out.T


class _Gain(ctypes.Structure):
    _fields_ = [("w12", ctypes.POINTER(ctypes.c_double)), ("w21", ctypes.POINTER(ctypes.c_double)),
                ("ga", ctypes.POINTER(ctypes.c_double)), ("ge", ctypes.POINTER(ctypes.c_double)),
                ("r12", ctypes.c_double), ("r21", ctypes.c_double), ("inv_tau", ctypes.c_double),
                ("gpa", ctypes.c_double), ("gpe", ctypes.c_double)]


class Gain:
    """Gain coefficients (per-bin arrays in FFT order, scalars) of the pump +
    signal model in the C layout; `c` is any object with the fields w12, w21,
    ga, ge, r12, r21, inv_tau, gpa, gpe."""

    def __init__(self, c):
        self._a = [np.ascontiguousarray(getattr(c, k), dtype=np.float64) for k in ("w12", "w21", "ga", "ge")]
        self._s = _Gain(*[_dp(a) for a in self._a], float(c.r12), float(c.r21), float(c.inv_tau), float(c.gpa),
                        float(c.gpe))

    @property
    def ref(self):
        return ctypes.byref(self._s)


def gain_rhs(fib: Fiber, gain: Gain, z: float, pump: float, x):
    """One evaluation of disp_gain_smf! (simulate_disp_gain_mmf.jl:21-63) -> (dPp, dx)."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.complex128).reshape(-1))
    dx = np.zeros_like(x)
    dp = np.zeros(2)
    _check(lib().orc_gain_rhs(fib.ref, gain.ref, ctypes.c_double(z), ctypes.c_double(pump),
                              _dp(x.view(np.float64)), _dp(dp), _dp(dx.view(np.float64))), "orc_gain_rhs")
    return complex(dp[0], dp[1]), dx[:, None]


def gain_forward(fib: Fiber, gain: Gain, length: float, u0, pump0: float, rtol: float = 1e-5, atol: float = 1e-6,
                 max_steps: int = 100000):
    """solve_disp_gain_smf (simulate_disp_gain_mmf.jl:200-239): Tsit5 on [Pp; x]
    -> (lab output spectrum (nt, 1), Pp(L), accepted step sizes, stats)."""
    u0 = _field(u0)
    out = np.zeros_like(u0)
    pl = np.zeros(1)
    hs = np.zeros(max_steps)
    st = np.zeros(3, dtype=np.int64)
    rc = lib().orc_gain_forward(fib.ref, gain.ref, ctypes.c_double(length), ctypes.c_double(rtol),
                                ctypes.c_double(atol), int(max_steps), ctypes.c_double(pump0),
                                _dp(u0.view(np.float64)), _dp(out.view(np.float64)), _dp(pl), _dp(hs),
                                st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    if rc == -3:
        raise RuntimeError("orc_gain_forward: max_steps reached")
    _check(rc, "orc_gain_forward")
    return out.T, float(pl[0]), hs[: st[0]].copy(), st


__all__ = ["Fiber", "build", "lib", "Gain", "gain_rhs", "gain_forward", "fft", "rhs", "rhs_vjp", "rk4_forward", "rk4_adjoint",
           "adjoint_rhs", "continuous_adjoint", "tsit5_tableau", "tsit5_dense_weights", "tsit5_forward", "tsit5_replay",
           "tsit5_adjoint", "np_ref"]
_ = os  # keep import for callers that patch paths
