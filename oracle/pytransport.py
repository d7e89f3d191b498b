"""ctypes binding of the CPU transport oracle (transport.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke`` for its reference answer
and by bench.py's cpu_baseline leg.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import pyoracle

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class SpDiag(C.Structure):
    _fields_ = [("n", C.c_int), ("numdiag", C.c_int), ("diag", C.c_int), ("ptr", _ip),
                ("val", _dp)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        L = pyoracle.lib()
        L.orc_tr_spavec.argtypes = [C.POINTER(SpDiag), C.POINTER(SpDiag), C.c_double, _dp, _dp]
        L.orc_tr_gauss_seidel.argtypes = [C.POINTER(SpDiag), C.POINTER(SpDiag), C.c_double,
                                          C.c_int, _dp, _dp]
        L.orc_tr_lu_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_tr_gefa.argtypes = [C.c_long, C.c_int, C.c_double, _dp]
        L.orc_tr_create.restype = C.c_void_p
        L.orc_tr_create.argtypes = [C.c_int] * 5 + [_dp] * 4 + [_ip, _ip, _dp]
        L.orc_tr_free.argtypes = [C.c_void_p]
        L.orc_tr_m.restype = C.c_long
        L.orc_tr_m.argtypes = [C.c_void_p]
        L.orc_tr_matvec.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.orc_tr_psolve.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_tr_dot.restype = C.c_double
        L.orc_tr_dot.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_tr_bicgstab.restype = C.c_int
        L.orc_tr_bicgstab.argtypes = [C.c_void_p, _dp, _dp, _ip, _dp]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def offsets7(nx, ny, nz):
    X, XY = nx + 2, (nx + 2) * (ny + 2)
    return np.array([-XY, -X, -1, 0, 1, X, XY], dtype=np.int32)


def offsets2(nx, ny, nz):
    return np.array([0, (nx + 2) * (ny + 2)], dtype=np.int32)


def _sp(val, ptr, diag):
    val = np.ascontiguousarray(val, dtype=np.float64)
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    s = SpDiag(val.shape[1], val.shape[0], diag, ptr.ctypes.data_as(_ip), _p(val))
    return s, (val, ptr)


def spavec(val, ptr, diag, x, bval=None, bptr=None, bdiag=1, alpha=0.0):
    """SpAVec_SpDiag (or SpABVec_SpDiag with alpha*B); val (numdiag, n)."""
    A, keep = _sp(val, ptr, diag)
    y = np.empty(val.shape[1])
    x = np.ascontiguousarray(x, dtype=np.float64)
    if bval is None:
        lib().orc_tr_spavec(C.byref(A), None, 0.0, _p(y), _p(x))
    else:
        B, keep2 = _sp(bval, bptr, bdiag)
        lib().orc_tr_spavec(C.byref(A), C.byref(B), alpha, _p(y), _p(x))
    return y


def gauss_seidel(val, ptr, diag, b, iterg=5, cval=None, cptr=None, cdiag=1, alpha=0.0):
    """GaussSeidel_SpDiag (or _SpDiagAB with alpha*C)."""
    A, keep = _sp(val, ptr, diag)
    x = np.empty(val.shape[1])
    b = np.ascontiguousarray(b, dtype=np.float64)
    if cval is None:
        lib().orc_tr_gauss_seidel(C.byref(A), None, 0.0, iterg, _p(x), _p(b))
    else:
        Cm, keep2 = _sp(cval, cptr, cdiag)
        lib().orc_tr_gauss_seidel(C.byref(A), C.byref(Cm), alpha, iterg, _p(x), _p(b))
    return x


def lu_solve(nx, ny, nz, nc, lu, v):
    v = np.array(v, dtype=np.float64, order="C")
    lib().orc_tr_lu_solve(nx, ny, nz, nc, _p(np.ascontiguousarray(lu)), _p(v))
    return v


def gefa(J, nc, fac):
    lu = np.array(J, dtype=np.float64, order="C")
    lib().orc_tr_gefa(lu.shape[1], nc, fac, _p(lu))
    return lu


class TransportOracle:
    """FP64 restatement of the transport stage solve of one TransportProblem."""

    def __init__(self, tp):
        self.tp = tp
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
                      for a in (tp.tmom, tp.tpot, tp.fall, tp.lu)]
        self._ops = [np.ascontiguousarray(tp.op_mv, dtype=np.int32),
                     np.ascontiguousarray(tp.op_ps, dtype=np.int32)]
        self._alpha = np.ascontiguousarray(tp.alpha, dtype=np.float64)
        p = [(_p(a) if a is not None else None) for a in self._keep]
        self.h = lib().orc_tr_create(tp.nx, tp.ny, tp.nz, tp.ncomp, tp.iterg, *p,
                                     self._ops[0].ctypes.data_as(_ip),
                                     self._ops[1].ctypes.data_as(_ip), _p(self._alpha))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tr_free(self.h)
            self.h = None

    def _vec(self, x):
        return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(self.tp.ncomp, -1))

    def matvec(self, x, init=False):
        x = self._vec(x)
        y = np.empty_like(x)
        lib().orc_tr_matvec(self.h, _p(y), _p(x), int(bool(init)))
        return y

    def psolve(self, x):
        x = self._vec(x)
        y = np.empty_like(x)
        lib().orc_tr_psolve(self.h, _p(y), _p(x))
        return y

    def dot(self, x, y):
        return lib().orc_tr_dot(self.h, _p(self._vec(x)), _p(self._vec(y)))

    def bicgstab(self, b, maxiter=30, tol=1e-6, x0=None):
        """Returns (x, MaxIter, Tol, status) with the reference's in/out conventions."""
        b = self._vec(b)
        x = np.zeros_like(b) if x0 is None else self._vec(x0).copy()
        mi, t = C.c_int(maxiter), C.c_double(tol)
        st = lib().orc_tr_bicgstab(self.h, _p(x), _p(b), C.byref(mi), C.byref(t))
        return x, mi.value, t.value, st
