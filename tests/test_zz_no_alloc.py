"""The allocation-free dot product (asam_dev_dot_noalloc) against asam_dev_dot through the
C-ABI: the same bits on two calls in a row, the second over fewer elements, with its
scratch shared between the calls; sizes beyond its maximum and null pointers rejected."""
import ctypes as C
import math

import numpy as np
import pytest

from asam_amd import workload as W

pytestmark = pytest.mark.gpu

E_INVALID = -1


def hip():
    h = C.CDLL("libamdhip64.so")
    h.hipDeviceSynchronize.restype = C.c_int
    return h


def test_dot_noalloc_matches_context_dot(sound_mod):
    p = W.config(4, np1=2, scale=0.5)
    s = sound_mod.SoundSolver.from_problem(p, setup=False)
    L = s.L
    f = L.asam_dev_dot_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
    h = hip()
    out = s.alloc(8)
    for n in (s.n, s.n // 3 + 7):  # the second call is shorter than the first
        x = W.normal(s.n, 70 + n)
        y = W.uniform01(s.n, 71 + n) - 0.5
        xb, yb = s.vec(x), s.vec(y)
        ref = s.dot(xb, yb, n)
        rc = f(None, n, xb.ptr, yb.ptr, out.ptr)
        assert rc == 0
        assert h.hipDeviceSynchronize() == 0
        got = float(out.download()[0])
        exact = math.fsum(x[:n] * y[:n])
        bound = n * 2.0 ** -53 * float(np.abs(x[:n] * y[:n]).sum())
        print(f"n={n}: noalloc {got:.17e} context dot {ref:.17e} fsum {exact:.17e}")
        assert got == ref  # same partials, same fixed summation order
        assert abs(got - exact) <= bound
    xb = s.vec(np.ones(s.n))
    assert f(None, (1 << 32) + 1, xb.ptr, xb.ptr, out.ptr) == E_INVALID  # host-side check
    assert f(None, 8, None, xb.ptr, out.ptr) == E_INVALID
    assert f(None, 8, xb.ptr, xb.ptr, None) == E_INVALID
    assert f(None, 0, xb.ptr, xb.ptr, out.ptr) == 0
    assert h.hipDeviceSynchronize() == 0
    assert float(out.download()[0]) == 0.0
