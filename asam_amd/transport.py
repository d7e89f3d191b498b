"""Python host mirror of ASAM's transport stage solve over the gfx950 C-ABI.

Mirrors the reference's interface for the second implicit system of a Rosenbrock stage:
``BiCGStabChemSp(A, x, b, MaxIter, Tol)`` / ``BiCGStabMetSp`` (reference
SRC/MODULE/SpMatrix4CellPar_Mod.f90:140-348) with the operator ``MatVecChem`` /
``MatVecMet`` (:56-84) and the preconditioner ``PSolve`` (:86-102) over the transport
Jacobian (JacTMom, JacTPot, JacFall as SpDiag matrices, the per-cell LU factors).
All arithmetic runs in libasam_sound.so on the GPU; there is no CPU fallback: a missing
library or device raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import sound

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

MOM, POT, FALL = 0, 1, 2
JACTMOM, JACTPOT, JACFALL = 0, 1, 2


class TrGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("ncomp", C.c_int),
                ("iterg", C.c_int), ("reserved", C.c_int * 3)]


class TrInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("status", C.c_int), ("rnorm0", C.c_double),
                ("rnorm", C.c_double), ("psolves", C.c_int), ("matvecs", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class TrTuning(C.Structure):
    _fields_ = [("gs_rows", C.c_int), ("gs_chunk", C.c_int), ("gs_spin_limit", C.c_int),
                ("reserved", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    L = sound.load()
    i = C.c_int
    sig = {
        "asam_tr_create": (i, [C.POINTER(_vp), C.POINTER(TrGrid)]),
        "asam_tr_destroy": (i, [_vp]),
        "asam_tr_last_error": (C.c_char_p, [_vp]),
        "asam_tr_set_matrix": (i, [_vp, i, i, i, _ip, _dp]),
        "asam_tr_set_lu": (i, [_vp, _dp]),
        "asam_tr_set_components": (i, [_vp, _ip, _ip, _dp]),
        "asam_tr_matvec": (i, [_vp, _dp, _dp, i]),
        "asam_tr_psolve": (i, [_vp, _dp, _dp]),
        "asam_tr_solve": (i, [_vp, _dp, _dp, _ip, _dp, C.POINTER(TrInfo)]),
        "asam_tr_matvec_device": (i, [_vp, _vp, _vp, i]),
        "asam_tr_psolve_device": (i, [_vp, _vp, _vp]),
        "asam_tr_solve_device": (i, [_vp, _vp, _vp, _ip, _dp, C.POINTER(TrInfo)]),
        "asam_tr_get_tuning": (i, [_vp, C.POINTER(TrTuning)]),
        "asam_tr_set_tuning": (i, [_vp, C.POINTER(TrTuning)]),
        "asam_tr_time_kernel": (i, [_vp, i, i, _dp]),
        "asam_tr_test_hold_boundary": (i, [_vp, i]),
        "asam_tr_dev_alloc": (i, [_vp, C.POINTER(_vp), C.c_size_t]),
        "asam_tr_dev_free": (i, [_vp, _vp]),
        "asam_tr_dev_upload": (i, [_vp, _vp, _vp, C.c_size_t]),
        "asam_tr_dev_download": (i, [_vp, _vp, _vp, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def offsets7(nx, ny, nz):
    X, XY = nx + 2, (nx + 2) * (ny + 2)
    return np.array([-XY, -X, -1, 0, 1, X, XY], dtype=np.int32)


def offsets2(nx, ny, nz):
    return np.array([0, (nx + 2) * (ny + 2)], dtype=np.int32)


class DeviceVector:
    """A device buffer of ncomp*m doubles on the solver's device."""

    def __init__(self, solver, n):
        self.s, self.n = solver, n
        p = _vp()
        solver._check(load().asam_tr_dev_alloc(solver.h, C.byref(p), 8 * n))
        self.ptr = p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.n
        self.s._check(load().asam_tr_dev_upload(self.s.h, self.ptr, a.ctypes.data, 8 * self.n))
        return self

    def download(self):
        out = np.empty(self.n)
        self.s._check(load().asam_tr_dev_download(self.s.h, out.ctypes.data, self.ptr,
                                                  8 * self.n))
        return out

    def free(self):
        if self.ptr:
            load().asam_tr_dev_free(self.s.h, self.ptr)
            self.ptr = None


class TransportSolver:
    """The transport stage system of one block on the GPU (JacSpMatrix4_T of one block)."""

    def __init__(self, nx, ny, nz, ncomp, iterg=5):
        L = load()
        self.nx, self.ny, self.nz, self.ncomp, self.iterg = nx, ny, nz, ncomp, iterg
        self.m = (nx + 2) * (ny + 2) * (nz + 2)
        h = _vp()
        rc = L.asam_tr_create(C.byref(h), C.byref(TrGrid(nx, ny, nz, ncomp, iterg)))
        if rc != 0:
            raise sound.AsamError(f"asam_tr_create failed ({rc})")
        self.h = h

    @classmethod
    def from_problem(cls, tp):
        """Set up from a TransportProblem (asam_amd.workload)."""
        s = cls(tp.nx, tp.ny, tp.nz, tp.ncomp, tp.iterg)
        s.set_matrix(JACTPOT, tp.tpot)
        s.set_matrix(JACTMOM, tp.tmom)
        if tp.fall is not None:
            s.set_matrix(JACFALL, tp.fall)
        s.set_lu(tp.lu)
        s.set_components(tp.op_mv, tp.op_ps, tp.alpha)
        return s

    def __del__(self):
        if getattr(self, "h", None):
            load().asam_tr_destroy(self.h)
            self.h = None

    def _check(self, rc):
        if rc != 0:
            msg = load().asam_tr_last_error(self.h)
            raise sound.AsamError(f"error {rc}: {msg.decode() if msg else ''}")

    def set_matrix(self, which, val, diagptr=None, diag=None):
        """Val of JacTMom / JacTPot (7, m) or JacFall (2, m) -- the Fortran Val(m, NumDiag)."""
        val = np.ascontiguousarray(val, dtype=np.float64)
        nd = val.shape[0]
        if diagptr is None:
            diagptr = (offsets2 if nd == 2 else offsets7)(self.nx, self.ny, self.nz)
        if diag is None:
            diag = 1 if nd == 2 else 4
        dp = np.ascontiguousarray(diagptr, dtype=np.int32)
        self._check(load().asam_tr_set_matrix(self.h, which, nd, diag, dp.ctypes.data_as(_ip),
                                              _p(val)))

    def set_lu(self, lu):
        if lu is None:
            self._check(load().asam_tr_set_lu(self.h, None))
        else:
            self._check(load().asam_tr_set_lu(self.h, _p(np.ascontiguousarray(lu, np.float64))))

    def set_components(self, op_mv=None, op_ps=None, alpha=None):
        a = [None if v is None else np.ascontiguousarray(v, dtype=np.int32)
             for v in (op_mv, op_ps)]
        al = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
        self._check(load().asam_tr_set_components(
            self.h, *(None if v is None else v.ctypes.data_as(_ip) for v in a),
            None if al is None else _p(al)))

    def _vec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        if x.size != self.ncomp * self.m:
            raise ValueError("vector size")
        return x

    def MatVec(self, x, init=False):
        """MatVecChem / MatVecMet_Vec4CellSp (init: the MatVecChem form of the initial
        residual)."""
        x = self._vec(x)
        y = np.empty_like(x)
        self._check(load().asam_tr_matvec(self.h, _p(x), _p(y), int(bool(init))))
        return y.reshape(self.ncomp, self.m)

    def PSolve(self, x):
        """PSolve_Vec4CellSp: 1 + IterG forward Gauss-Seidel sweeps per component."""
        x = self._vec(x)
        y = np.empty_like(x)
        self._check(load().asam_tr_psolve(self.h, _p(x), _p(y)))
        return y.reshape(self.ncomp, self.m)

    def BiCGStab(self, x, b, MaxIter=30, Tol=1e-6):
        """BiCGStabChem/Met_Vec4CellSp(A, x, b, MaxIter, Tol): returns (x, MaxIter, Tol,
        info) with the reference's in/out conventions."""
        x = self._vec(x).copy()
        b = self._vec(b)
        mi, tol, info = C.c_int(MaxIter), C.c_double(Tol), TrInfo()
        self._check(load().asam_tr_solve(self.h, _p(x), _p(b), C.byref(mi), C.byref(tol),
                                         C.byref(info)))
        return x.reshape(self.ncomp, self.m), mi.value, tol.value, info

    BiCGStabChemSp = BiCGStab
    BiCGStabMetSp = BiCGStab

    def device_vector(self):
        return DeviceVector(self, self.ncomp * self.m)

    def matvec_device(self, x: DeviceVector, y: DeviceVector, init=False):
        self._check(load().asam_tr_matvec_device(self.h, x.ptr, y.ptr, int(bool(init))))

    def psolve_device(self, x: DeviceVector, y: DeviceVector):
        self._check(load().asam_tr_psolve_device(self.h, x.ptr, y.ptr))

    def solve_device(self, x: DeviceVector, b: DeviceVector, MaxIter=30, Tol=1e-6):
        mi, tol, info = C.c_int(MaxIter), C.c_double(Tol), TrInfo()
        self._check(load().asam_tr_solve_device(self.h, x.ptr, b.ptr, C.byref(mi),
                                                C.byref(tol), C.byref(info)))
        return mi.value, tol.value, info

    def get_tuning(self):
        t = TrTuning()
        self._check(load().asam_tr_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, **kw):
        t = self.get_tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._check(load().asam_tr_set_tuning(self.h, C.byref(t)))

    def test_hold_boundary(self, hold=True):
        """Test hook: no block publishes its boundary row, so later blocks wait until the
        gs_spin_limit and the sweep fails (asam_tr_test_hold_boundary)."""
        self._check(load().asam_tr_test_hold_boundary(self.h, int(bool(hold))))

    def time_kernel(self, which, reps=10):
        """Device ms of one Gauss-Seidel sweep (0), one MatVec (1), one PSolve (2)."""
        ms = C.c_double()
        self._check(load().asam_tr_time_kernel(self.h, which, reps, C.byref(ms)))
        return ms.value
