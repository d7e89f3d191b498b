"""Python host mirror of ASAM's SolveSound interface over the gfx950 C-ABI.

Mirrors the reference's solver-call interface for the stage solve:
``SolveSound(x, b, MaxIter, Tol)`` (reference SRC/SOLU/Iter_Mod.f90:9-30), with
the operator/preconditioner plug points ``MatVec1`` (SRC/MATRIX/MatVec_Mod.f90:8)
and ``PSolve1`` (SRC/MATRIX/PreCond_Mod.f90:1555) and the set-up ``JacAccGrav``
(SRC/MULTI/JacAccGrav_Mod.f90:304).  All arithmetic runs in libasam_sound.so on
the GPU; there is no CPU fallback: a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libasam_sound.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "asam_sound.h")

METHODS = {"BICGSTAB": 0, "FGMRES": 1, "GMRES": 2, "FGMRESV": 3}
PRECISIONS = {"FP64": 0, "MIXED": 1}

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("np1", C.c_int),
                ("grad_full", C.c_int), ("beta0_dt", C.c_double), ("grav", C.c_double),
                ("fac_anela", C.c_double)]


class Control(C.Structure):
    _fields_ = [("method", C.c_int), ("step_gmres", C.c_int), ("precision", C.c_int),
                ("inner_tol", C.c_double), ("max_inner", C.c_int), ("max_outer", C.c_int)]


class Info(C.Structure):
    _fields_ = [("iterations", C.c_int), ("restarts", C.c_int), ("status", C.c_int),
                ("reserved", C.c_int), ("rel_residual", C.c_double),
                ("abs_residual", C.c_double), ("ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class Tuning(C.Structure):
    _fields_ = [("apply_variant", C.c_int), ("apply_wg", C.c_int), ("line_window", C.c_int),
                ("line_lds", C.c_int), ("fused_tail", C.c_int), ("apply_xcd", C.c_int),
                ("apply_bx", C.c_int), ("line_tiled", C.c_int), ("apply_band_rows", C.c_int),
                ("apply_rows", C.c_int), ("line_nt", C.c_int),
                ("apply_depth", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class AsamError(RuntimeError):
    pass


_lib = None


def exported_names_from_header(path: str = None):
    """Function names declared in include/asam_sound.h, asam_ensemble.h and
    asam_transport.h (or in the one header named)."""
    import re
    inc = os.path.dirname(HEADER)
    paths = [path] if path else [HEADER, os.path.join(inc, "asam_ensemble.h"),
                                 os.path.join(inc, "asam_transport.h")]
    names = set()
    for pth in paths:
        txt = open(pth).read()
        names |= set(re.findall(r"^\s*(?:int|long|const char \*)\s*\**\s*(asam_\w+)\s*\(",
                                txt, flags=re.M))
    return sorted(names)


def load(path: str = LIB_PATH):
    """Load libasam_sound.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: run __graft_entry__.build() first")
    L = C.CDLL(path)
    i, ip, dp, pp = C.c_int, C.POINTER(C.c_int), _dp, C.POINTER(_vp)
    sig = {
        "asam_sound_create": (i, [C.POINTER(_vp), C.POINTER(Grid)]),
        "asam_sound_destroy": (i, [_vp]),
        "asam_sound_default_control": (i, [C.POINTER(Control)]),
        "asam_sound_set_control": (i, [_vp, C.POINTER(Control)]),
        "asam_sound_last_error": (C.c_char_p, [_vp]),
        "asam_sound_setup": (i, [_vp] + [dp] * 10),
        "asam_sound_setup_weights": (i, [_vp] + [dp] * 11),
        "asam_sound_setup_moist": (i, [_vp] + [dp] * 13),
        "asam_sound_solve": (i, [_vp, dp, dp, ip, dp, C.POINTER(Info)]),
        "asam_sound_solve_device": (i, [_vp, _vp, _vp, ip, dp, C.POINTER(Info)]),
        "asam_sound_solve_T": (i, [_vp, dp, dp, ip, dp, C.POINTER(Info)]),
        "asam_sound_solve_T_device": (i, [_vp, _vp, _vp, ip, dp, C.POINTER(Info)]),
        "asam_sound_apply": (i, [_vp, dp, dp]),
        "asam_sound_apply_T": (i, [_vp, dp, dp]),
        "asam_sound_psolve": (i, [_vp, dp, dp]),
        "asam_sound_psolve_T": (i, [_vp, dp, dp]),
        "asam_sound_apply_device": (i, [_vp, _vp, _vp, i]),
        "asam_sound_psolve_device": (i, [_vp, _vp, _vp, i]),
        "asam_sound_theta_vjp_device": (i, [_vp, _vp, _vp, _vp]),
        "asam_dev_alloc": (i, [_vp, pp, C.c_size_t]),
        "asam_dev_free": (i, [_vp, _vp]),
        "asam_dev_upload": (i, [_vp, _vp, _vp, C.c_size_t]),
        "asam_dev_download": (i, [_vp, _vp, _vp, C.c_size_t]),
        "asam_dev_memset": (i, [_vp, _vp, i, C.c_size_t]),
        "asam_dev_axpby": (i, [_vp, C.c_size_t, C.c_double, _vp, C.c_double, _vp]),
        "asam_dev_dot": (i, [_vp, C.c_size_t, _vp, _vp, dp]),
        "asam_dev_mass": (i, [_vp, _vp, _vp]),
        "asam_sound_time_kernel": (i, [_vp, i, i, i, dp]),
        "asam_sound_last_counts": (i, [_vp, ip]),
        "asam_sound_profile_enable": (i, [_vp, i]),
        "asam_sound_get_tuning": (i, [_vp, C.POINTER(Tuning)]),
        "asam_sound_time_copy": (i, [_vp, C.c_size_t, i, dp]),
        "asam_sound_set_tuning": (i, [_vp, C.POINTER(Tuning)]),
        "asam_sound_profile_read": (i, [_vp, dp, dp, ip]),
        "asam_sound_autotune": (i, [_vp, i, C.POINTER(Tuning), dp]),
        "asam_sound_rhs": (i, [_vp] + [dp] * 6),
        "asam_sound_rhs_device": (i, [_vp] * 7),
        "asam_sound_update": (i, [_vp] + [dp] * 6),
        "asam_sound_update_device": (i, [_vp] * 7),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class DeviceBuffer:
    """A device allocation owned by a solver context."""

    def __init__(self, solver: "SoundSolver", nbytes: int):
        self.solver = solver
        self.nbytes = nbytes
        ptr = _vp()
        solver._chk(solver.L.asam_dev_alloc(solver.h, C.byref(ptr), nbytes))
        self.ptr = ptr

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        self.solver._chk(self.solver.L.asam_dev_upload(self.solver.h, self.ptr,
                                                       a.ctypes.data_as(_vp), a.nbytes))
        return self

    def download(self, dtype=np.float64) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        self.solver._chk(self.solver.L.asam_dev_download(self.solver.h, out.ctypes.data_as(_vp),
                                                         self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr:
            self.solver.L.asam_dev_free(self.solver.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class SoundSolver:
    """One stage-system context on the current GPU (the JacAccGrav/SolveSound pair)."""

    def __init__(self, nx, ny, nz, np1=1, grad_full=1, beta0_dt=0.871733, grav=0.0,
                 fac_anela=1.0):
        self.L = load()
        self.grid = Grid(nx, ny, nz, np1, grad_full, beta0_dt, grav, fac_anela)
        self.h = _vp()
        rc = self.L.asam_sound_create(C.byref(self.h), C.byref(self.grid))
        if rc != 0:
            msg = self.L.asam_sound_last_error(self.h).decode() if self.h else ""
            if self.h:
                self.L.asam_sound_destroy(self.h)
                self.h = None
            raise AsamError(f"asam_sound_create failed ({rc}): {msg}")
        self.n = nx * ny * nz * np1
        self.ctl = Control()
        self.L.asam_sound_default_control(C.byref(self.ctl))

    @classmethod
    def from_problem(cls, prob, setup=True):
        g = prob.grid
        s = cls(g.nx, g.ny, g.nz, prob.np1, prob.grad_full, prob.beta0_dt, prob.grav)
        if setup:
            s.setup(prob.FU, prob.FV, prob.FW, prob.VolC, prob.dxv, prob.dyv, prob.dzv,
                    prob.rho, prob.rhotheta, prob.T)
        return s

    def _chk(self, rc):
        if rc != 0:
            raise AsamError(f"libasam_sound error {rc}: "
                            f"{self.L.asam_sound_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.asam_sound_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- set-up
    def setup(self, FU, FV, FW, VolC, dx, dy, dz, rho, rhotheta, T):
        arrs = [_c(a) for a in (FU, FV, FW, VolC, dx, dy, dz, rho, rhotheta, T)]
        self._chk(self.L.asam_sound_setup(self.h, *[_p(a) for a in arrs]))

    def setup_moist(self, FU, FV, FW, VolC, dx, dy, dz, rho, rhotheta, T, rhov=None, rhol=None,
                    rhor=None):
        """JacAccGrav with moisture: vapour, cloud water and rain densities (None: none)."""
        arrs = [_c(a) for a in (FU, FV, FW, VolC, dx, dy, dz, rho, rhotheta, T)]
        q = [None if a is None else _c(a) for a in (rhov, rhol, rhor)]
        self._chk(self.L.asam_sound_setup_moist(self.h, *[_p(a) for a in arrs],
                                                *[_p(a) if a is not None else None for a in q]))

    def setup_weights(self, FU, FV, FW, VolC, dx, dy, dz, DTU, DTV, DTW, DUT):
        arrs = [_c(a) for a in (FU, FV, FW, VolC, dx, dy, dz, DTU, DTV, DTW, DUT)]
        self._chk(self.L.asam_sound_setup_weights(self.h, *[_p(a) for a in arrs]))

    def set_control(self, method="BICGSTAB", step_gmres=10, precision="FP64", inner_tol=1e-5,
                    max_inner=200, max_outer=12):
        self.ctl = Control(METHODS[method], step_gmres, PRECISIONS[precision], inner_tol,
                           max_inner, max_outer)
        self._chk(self.L.asam_sound_set_control(self.h, C.byref(self.ctl)))

    def get_tuning(self) -> dict:
        t = Tuning()
        self._chk(self.L.asam_sound_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = Tuning()
        self._chk(self.L.asam_sound_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._chk(self.L.asam_sound_set_tuning(self.h, C.byref(t)))

    # ---------------------------------------------------------------- solve
    def SolveSound(self, x, b, MaxIter, Tol):
        """Reference semantics: solves MatVec1(x) = -A x = b in place of x.

        Returns (x, MaxIter, Tol, info): MaxIter/Tol as the reference leaves them
        (BiCGStab: +-Iter / 2*Iter and the final residual norm)."""
        x = _c(x).copy()
        b = _c(b)
        mi = C.c_int(int(MaxIter))
        tl = C.c_double(float(Tol))
        info = Info()
        self._chk(self.L.asam_sound_solve(self.h, _p(x), _p(b), C.byref(mi), C.byref(tl),
                                          C.byref(info)))
        return x, mi.value, tl.value, info

    def SolveSound_T(self, x, b, MaxIter, Tol):
        x = _c(x).copy()
        b = _c(b)
        mi = C.c_int(int(MaxIter))
        tl = C.c_double(float(Tol))
        info = Info()
        self._chk(self.L.asam_sound_solve_T(self.h, _p(x), _p(b), C.byref(mi), C.byref(tl),
                                            C.byref(info)))
        return x, mi.value, tl.value, info

    def solve_device(self, xbuf: DeviceBuffer, bbuf: DeviceBuffer, MaxIter, Tol,
                     transpose=False):
        mi = C.c_int(int(MaxIter))
        tl = C.c_double(float(Tol))
        info = Info()
        f = self.L.asam_sound_solve_T_device if transpose else self.L.asam_sound_solve_device
        self._chk(f(self.h, xbuf.ptr, bbuf.ptr, C.byref(mi), C.byref(tl), C.byref(info)))
        return mi.value, tl.value, info

    # ---------------------------------------------------------------- pieces
    def _unary(self, f, x):
        x = _c(x)
        y = np.empty_like(x)
        self._chk(f(self.h, _p(y), _p(x)))
        return y

    def MatVec1(self, x):
        return self._unary(self.L.asam_sound_apply, x)

    def MatVec1_T(self, x):
        return self._unary(self.L.asam_sound_apply_T, x)

    def PSolve1(self, x):
        return self._unary(self.L.asam_sound_psolve, x)

    def PSolve1_T(self, x):
        return self._unary(self.L.asam_sound_psolve_T, x)

    def alloc(self, nbytes) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def vec(self, a=None) -> DeviceBuffer:
        buf = DeviceBuffer(self, self.n * 8)
        if a is None:
            self._chk(self.L.asam_dev_memset(self.h, buf.ptr, 0, self.n * 8))
        else:
            buf.upload(_c(a))
        return buf

    def autotune(self, reps: int = 5):
        """Pick the fastest launch shapes for the current precision by timing whole BiCGStab
        iterations; returns (tuning dict, [ms per iteration, 0])."""
        t = Tuning()
        ms = (C.c_double * 2)()
        self._chk(self.L.asam_sound_autotune(self.h, reps, C.byref(t), ms))
        return t.as_dict(), list(ms)

    # ------------------------------------------------------ either side of the solve
    def Rhs(self, U, V, W, incr_rt, incr_r=None):
        """b of the stage system from predicted face velocities U (nz,ny,nx+1), V
        (nz,ny+1,nx), W (nz+1,ny,nx) and the stage increments of rho*theta and rho
        (reference Rhs, MATRIX/Rhs_Mod.f90:110-157)."""
        b = np.empty(self.n)
        a = [_c(v) for v in (U, V, W, incr_rt)]
        ir = _c(incr_r) if incr_r is not None else None
        self._chk(self.L.asam_sound_rhs(self.h, _p(b), *[_p(v) for v in a],
                                        _p(ir) if ir is not None else None))
        return b

    def UpdateVelocity(self, x, U, V, W, incr_rt, incr_r=None):
        """Post-solve update (reference UpdateVelocity, MATRIX/Rhs_Mod.f90:589-952);
        returns updated copies (x, U, V, W, incr_rt, incr_r)."""
        x, U, V, W, it = (np.array(v, dtype=np.float64) for v in (x, U, V, W, incr_rt))
        ir = None if incr_r is None else np.array(incr_r, dtype=np.float64)
        self._chk(self.L.asam_sound_update(self.h, _p(x), _p(U), _p(V), _p(W), _p(it),
                                           _p(ir) if ir is not None else None))
        return x, U, V, W, it, ir

    def rhs_device(self, b, U, V, W, it, ir=None):
        self._chk(self.L.asam_sound_rhs_device(self.h, b.ptr, U.ptr, V.ptr, W.ptr, it.ptr,
                                               ir.ptr if ir is not None else None))

    def update_device(self, x, U, V, W, it, ir=None):
        self._chk(self.L.asam_sound_update_device(self.h, x.ptr, U.ptr, V.ptr, W.ptr, it.ptr,
                                                  ir.ptr if ir is not None else None))

    def zero(self, buf: DeviceBuffer):
        self._chk(self.L.asam_dev_memset(self.h, buf.ptr, 0, buf.nbytes))

    def apply_device(self, y: DeviceBuffer, x: DeviceBuffer, transpose=False):
        self._chk(self.L.asam_sound_apply_device(self.h, y.ptr, x.ptr, int(transpose)))

    def psolve_device(self, y: DeviceBuffer, x: DeviceBuffer, transpose=False):
        self._chk(self.L.asam_sound_psolve_device(self.h, y.ptr, x.ptr, int(transpose)))

    def axpby(self, a, x: DeviceBuffer, b, y: DeviceBuffer, n=None):
        self._chk(self.L.asam_dev_axpby(self.h, n or self.n, a, x.ptr, b, y.ptr))

    def dot(self, x: DeviceBuffer, y: DeviceBuffer, n=None) -> float:
        out = C.c_double(0.0)
        self._chk(self.L.asam_dev_dot(self.h, n or self.n, x.ptr, y.ptr, C.byref(out)))
        return out.value

    def mass(self, y: DeviceBuffer, x: DeviceBuffer):
        self._chk(self.L.asam_dev_mass(self.h, y.ptr, x.ptr))

    def theta_vjp(self, grad: DeviceBuffer, mu: DeviceBuffer, x: DeviceBuffer):
        self._chk(self.L.asam_sound_theta_vjp_device(self.h, grad.ptr, mu.ptr, x.ptr))

    def time_kernel(self, which: int, fp32: bool, reps: int = 20) -> float:
        ms = C.c_double(0.0)
        self._chk(self.L.asam_sound_time_kernel(self.h, which, int(fp32), reps, C.byref(ms)))
        return ms.value

    PROFILE_CLASSES = ("apply_fp64", "apply_fp32", "line_fp64", "line_fp32", "blas_fp64",
                       "blas_fp32", "reduce", "other")

    def profile(self, on: bool = True):
        self._chk(self.L.asam_sound_profile_enable(self.h, int(on)))

    def profile_read(self):
        """Per kernel class: (ms summed, algorithmic bytes summed, launches) since last read."""
        ms, by, ln = (C.c_double * 8)(), (C.c_double * 8)(), (C.c_int * 8)()
        self._chk(self.L.asam_sound_profile_read(self.h, ms, by, ln))
        return {k: {"ms": ms[q], "bytes": by[q], "launches": ln[q]}
                for q, k in enumerate(self.PROFILE_CLASSES)}

    def time_copy(self, nbytes: int, reps: int = 20) -> float:
        ms = C.c_double(0.0)
        self._chk(self.L.asam_sound_time_copy(self.h, nbytes, reps, C.byref(ms)))
        return ms.value

    def last_counts(self):
        arr = (C.c_int * 8)()
        self._chk(self.L.asam_sound_last_counts(self.h, arr))
        return list(arr)
