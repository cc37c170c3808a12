"""Multimode cross-mode coupling on the GPU against closed forms.

A CW launch at one frequency bin with gamma = a on (i,i,i,i) and b on every
arrangement of (i,i,j,j) has an exact solution (tests/conftest.py:
cw_multimode_case): every mode turns at the rate (a + 3(M-1)b) P when all
modes are in phase, or (a + (2M-3)b) P when half of them are a quarter period
behind, times the shock factor of the bin, on top of the dispersion phase.
It ties the i != j and k != l terms of the Kerr/Raman contraction of
disp_mmf! (simulate_disp_mmf.jl:38-53) to mathematics rather than to the
oracle alone.  Bounds: FP64 1e-10 relative (normwise field), FP32 1e-5.  The
FP32 cases at M = 6 and 10 with at most 512 threads per solve run the Kerr
contraction on the matrix cores; the 16384-sample M = 10 case runs the
vector-unit contraction.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import cw_multimode_case, rel

pytestmark = pytest.mark.gpu


def _window(nt):
    return max(3.0, 1.2 * nt / (2 * 193.41449))


def _plan(case, precision, steps=None, scheme=None, rtol=1e-12, atol=1e-12):
    import fiberhip

    D, g, hR, S, omf, u0, expected, L = case
    return fiberhip.Plan(D, g, hR, S, omf, L, steps, precision=precision, scheme=scheme, rtol=rtol, atol=atol,
                         max_steps=20000)


@pytest.mark.parametrize("m,pattern", [(3, "phase"), (4, "quad"), (6, "phase"), (6, "quad"), (10, "phase"),
                                       (10, "quad")])
def test_cw_closed_form_fp64(gpu_context, m, pattern):
    case = cw_multimode_case(m, pattern)
    u0, expected = case[5], case[6]
    print(f"case: cw closed form fp64 M={m} {pattern} nt=256 rk4 1000 steps and tsit5 1e-12", flush=True)
    out, _ = _plan(case, "fp64", steps=1000).forward(u0)
    e_rk4 = rel(out, expected)
    out, st = _plan(case, "fp64", scheme="tsit5").forward(u0)
    e_ts = rel(out, expected)
    print(f"M={m} {pattern}: rk4 {e_rk4:.3e}  tsit5 {e_ts:.3e} ({st['accepted']} steps)", flush=True)
    assert e_rk4 <= 1e-10
    assert e_ts <= 1e-10


@pytest.mark.parametrize("m,pattern,nt", [(2, "quad", 256), (6, "phase", 256), (6, "quad", 4096),
                                          (10, "phase", 256), (10, "quad", 1024), (10, "phase", 16384)])
def test_cw_closed_form_fp32(gpu_context, m, pattern, nt):
    case = cw_multimode_case(m, pattern, nt=nt, window_ps=_window(nt))
    u0, expected = case[5], case[6]
    print(f"case: cw closed form fp32 M={m} {pattern} nt={nt} rk4 1000 steps and tsit5 1e-8", flush=True)
    out, _ = _plan(case, "fp32", steps=1000).forward(u0)
    e_rk4 = rel(out, expected)
    out, st = _plan(case, "fp32", scheme="tsit5", rtol=1e-8, atol=1e-8).forward(u0)
    e_ts = rel(out, expected)
    print(f"M={m} {pattern} nt={nt}: rk4 {e_rk4:.3e}  tsit5 {e_ts:.3e} ({st['accepted']} steps)", flush=True)
    assert e_rk4 <= 1e-5
    assert e_ts <= 1e-5
