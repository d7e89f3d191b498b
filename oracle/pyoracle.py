"""ctypes binding of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke`` for its
reference answer, and by bench.py's cpu_baseline leg.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liborc.so")

_dp = C.POINTER(C.c_double)


class OrcGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("np1", C.c_int),
                ("grad_full", C.c_int), ("beta0_dt", C.c_double), ("grav", C.c_double),
                ("fac_anela", C.c_double)]


class MGControl(C.Structure):
    _fields_ = [("ref_level", C.c_int), ("gauss_iter2", C.c_int), ("gauss_iter3", C.c_int),
                ("mult_iter2", C.c_int), ("mult_iter3", C.c_int)]


def build(quiet: bool = True) -> str:
    subprocess.run(["make", "-C", HERE], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_assemble.restype = C.c_void_p
        L.orc_assemble.argtypes = [C.POINTER(OrcGrid)] + [_dp] * 10
        L.orc_assemble_weights.restype = C.c_void_p
        L.orc_assemble_weights.argtypes = [C.POINTER(OrcGrid)] + [_dp] * 11
        L.orc_weights.argtypes = [C.POINTER(OrcGrid)] + [_dp] * 8
        L.orc_weights_moist.argtypes = [C.POINTER(OrcGrid)] + [_dp] * 11
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_n.restype = C.c_long
        L.orc_n.argtypes = [C.c_void_p]
        L.orc_export.argtypes = [C.c_void_p] + [_dp] * 7
        for f in ("orc_apply", "orc_matvec1", "orc_matvec1_T", "orc_psolve1", "orc_psolve1_T",
                  "orc_line_solve"):
            getattr(L, f).argtypes = [C.c_void_p, _dp, _dp]
        L.orc_bicgstab.restype = C.c_int
        L.orc_bicgstab.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), C.c_int]
        for f in ("orc_fgmres", "orc_fgmresv", "orc_gmres"):
            getattr(L, f).restype = C.c_int
            getattr(L, f).argtypes = [C.c_void_p, _dp, _dp, C.c_int, C.c_double, C.c_int,
                                      C.c_int, _dp]
        L.orc_rhs.argtypes = [C.POINTER(OrcGrid)] + [_dp] * 13
        L.orc_update.argtypes = [C.POINTER(OrcGrid)] + [_dp] * 16
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_mg_default_control.argtypes = [C.POINTER(MGControl)]
        L.orc_mg_build.restype = C.c_void_p
        L.orc_mg_build.argtypes = [C.POINTER(OrcGrid), C.POINTER(MGControl)] + [_dp] * 11
        L.orc_mg_free.argtypes = [C.c_void_p]
        L.orc_mg_pressurepre.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_mg_levels.restype = C.c_int
        L.orc_mg_levels.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int)]
        for f in ("orc_mg_apply_level", "orc_mg_restrict", "orc_mg_prolong"):
            getattr(L, f).restype = C.c_int
            getattr(L, f).argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.orc_set_pressurepre.argtypes = [C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """FP64 reference-restating operator, preconditioner and Krylov solvers."""

    def __init__(self, prob, weights=None):
        L = lib()
        g = prob.grid
        self.prob = prob
        self.g = OrcGrid(g.nx, g.ny, g.nz, prob.np1, prob.grad_full, prob.beta0_dt,
                         prob.grav, 1.0)
        self._keep = [_c(prob.FU), _c(prob.FV), _c(prob.FW), _c(prob.VolC), _c(prob.dxv),
                      _c(prob.dyv), _c(prob.dzv)]
        self._w = None if weights is None else [_c(a) for a in weights]
        if weights is None:
            st = [_c(prob.rho), _c(prob.rhotheta), _c(prob.T)]
            self._h = L.orc_assemble(C.byref(self.g), *[_p(a) for a in self._keep + st])
        else:
            w = [_c(a) for a in weights]
            self._h = L.orc_assemble_weights(C.byref(self.g), *[_p(a) for a in self._keep + w])
        self.n = int(L.orc_n(self._h))
        self.N = g.ncell

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().orc_free(self._h)
                self._h = None
        except Exception:
            pass

    def _face_weights(self):
        if self._w is None:
            self._w = [_c(a) for a in self.weights()]
        return self._w

    def rhs(self, U, V, W, incr_rt, incr_r=None):
        """b of the stage system from predicted face velocities and the stage increments
        (Rhs, Rhs_Mod.f90:110-157)."""
        p = self.prob
        DTU, DTV, DTW, _ = self._face_weights()
        b = np.empty(self.n)
        ir = _c(incr_r if incr_r is not None else np.zeros(self.N))
        lib().orc_rhs(C.byref(self.g), _p(self._keep[0]), _p(self._keep[1]), _p(self._keep[2]),
                      _p(self._keep[3]), _p(DTU), _p(DTV), _p(DTW), _p(_c(U)), _p(_c(V)),
                      _p(_c(W)), _p(_c(incr_rt)), _p(ir), _p(b))
        return b

    def update(self, x, U, V, W, incr_rt, incr_r=None):
        """UpdateVelocity (Rhs_Mod.f90:589-952): returns updated copies
        (x, U, V, W, incr_rt, incr_r)."""
        DTU, DTV, DTW, DUT = self._face_weights()
        x, U, V, W, it = (np.array(a, dtype=np.float64) for a in (x, U, V, W, incr_rt))
        ir = None if incr_r is None else np.array(incr_r, dtype=np.float64)
        k = self._keep
        lib().orc_update(C.byref(self.g), _p(k[0]), _p(k[1]), _p(k[2]), _p(k[3]), _p(k[4]),
                         _p(k[5]), _p(k[6]), _p(DTU), _p(DTV), _p(DTW), _p(_c(DUT)), _p(x),
                         _p(U), _p(V), _p(W), _p(it), _p(ir) if ir is not None else None)
        return x, U, V, W, it, ir

    def weights(self):
        p, g = self.prob, self.prob.grid
        DTU = np.empty((g.nz, g.ny, g.nx + 1))
        DTV = np.empty((g.nz, g.ny + 1, g.nx))
        DTW = np.empty((g.nz + 1, g.ny, g.nx))
        DUT = np.empty((g.nz, g.ny, g.nx))
        lib().orc_weights(C.byref(self.g), _p(_c(p.VolC)), _p(_c(p.rho)), _p(_c(p.rhotheta)),
                          _p(_c(p.T)), _p(DTU), _p(DTV), _p(DTW), _p(DUT))
        return DTU, DTV, DTW, DUT

    def weights_moist(self, rhov=None, rhol=None, rhor=None):
        """Face theta and DUT of moist air (WeightDiagCompute 'Density') for this problem's
        state with vapour, cloud water and rain densities (None: none)."""
        p, g = self.prob, self.prob.grid
        DTU = np.empty((g.nz, g.ny, g.nx + 1))
        DTV = np.empty((g.nz, g.ny + 1, g.nx))
        DTW = np.empty((g.nz + 1, g.ny, g.nx))
        DUT = np.empty((g.nz, g.ny, g.nx))
        q = [None if a is None else _c(a) for a in (rhov, rhol, rhor)]
        lib().orc_weights_moist(C.byref(self.g), _p(_c(p.VolC)), _p(_c(p.rho)),
                                _p(_c(p.rhotheta)), _p(_c(p.T)),
                                *[_p(a) if a is not None else None for a in q],
                                _p(DTU), _p(DTV), _p(DTW), _p(DUT))
        return DTU, DTV, DTW, DUT

    def export(self):
        N = self.N
        w = 1 if self.prob.np1 == 1 else 4
        wh = 1 if self.prob.np1 == 1 else 2
        d, uz, lz = (np.empty(N * w) for _ in range(3))
        ux, lx, uy, ly = (np.empty(N * wh) for _ in range(4))
        lib().orc_export(self._h, _p(d), _p(ux), _p(lx), _p(uy), _p(ly), _p(uz), _p(lz))
        return dict(d=d, ux=ux, lx=lx, uy=uy, ly=ly, uz=uz, lz=lz)

    def _unary(self, fname, x):
        x = _c(x)
        y = np.empty_like(x)
        getattr(lib(), fname)(self._h, _p(y), _p(x))
        return y

    def apply(self, x):
        return self._unary("orc_apply", x)

    def matvec1(self, x):
        return self._unary("orc_matvec1", x)

    def matvec1_T(self, x):
        return self._unary("orc_matvec1_T", x)

    def psolve1(self, x):
        return self._unary("orc_psolve1", x)

    def psolve1_T(self, x):
        return self._unary("orc_psolve1_T", x)

    def line_solve(self, x):
        return self._unary("orc_line_solve", x)

    def bicgstab(self, b, maxiter=200, tol=1e-8, transpose=False, x0=None):
        x = np.zeros(self.n) if x0 is None else _c(x0).copy()
        mi = C.c_int(maxiter)
        tl = C.c_double(tol)
        st = lib().orc_bicgstab(self._h, _p(x), _p(_c(b)), C.byref(mi), C.byref(tl),
                                int(transpose))
        return x, mi.value, tl.value, st

    def fgmres(self, b, maxiter=10, tol=1e-8, m=10, transpose=False, x0=None):
        x = np.zeros(self.n) if x0 is None else _c(x0).copy()
        rr = C.c_double(0.0)
        it = lib().orc_fgmres(self._h, _p(x), _p(_c(b)), maxiter, tol, m, int(transpose),
                              C.byref(rr))
        return x, it, rr.value

    def fgmresv(self, b, maxiter=10, tol=1e-8, m=10, transpose=False, x0=None):
        x = np.zeros(self.n) if x0 is None else _c(x0).copy()
        rr = C.c_double(0.0)
        it = lib().orc_fgmresv(self._h, _p(x), _p(_c(b)), maxiter, tol, m, int(transpose),
                               C.byref(rr))
        return x, it, rr.value

    def gmres(self, b, maxiter=10, tol=1e-8, m=10, transpose=False, x0=None):
        x = np.zeros(self.n) if x0 is None else _c(x0).copy()
        rr = C.c_double(0.0)
        it = lib().orc_gmres(self._h, _p(x), _p(_c(b)), maxiter, tol, m, int(transpose),
                             C.byref(rr))
        return x, it, rr.value


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def to_csr(orc: Oracle):
    """Explicit sparse matrix of A (not -A) assembled from the exported coefficients."""
    import scipy.sparse as sp
    p = orc.prob
    g = p.grid
    nx, ny, nz = g.nx, g.ny, g.nz
    N = g.ncell
    c = orc.export()
    rows, cols, vals = [], [], []
    idx = np.arange(N)
    i = idx % nx
    j = (idx // nx) % ny
    k = idx // (nx * ny)
    if p.np1 == 1:
        rows.append(idx); cols.append(idx); vals.append(c["d"])
        for (u, l, s, m) in (("ux", "lx", 1, i + 1 < nx), ("uy", "ly", nx, j + 1 < ny),
                             ("uz", "lz", nx * ny, k + 1 < nz)):
            a = idx[m]
            rows.append(a); cols.append(a + s); vals.append(c[u][m])
            rows.append(a + s); cols.append(a); vals.append(c[l][m])
        A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(N, N))
        return A
    d = c["d"].reshape(N, 4)  # [11,21,12,22]
    for (r, q, col) in ((0, 0, 0), (1, 1, 0), (0, 2, 1), (1, 3, 1)):
        rows.append(2 * idx + r); cols.append(2 * idx + col); vals.append(d[:, q])
    uz = c["uz"].reshape(N, 4)
    lz = c["lz"].reshape(N, 4)
    m = k + 1 < nz
    a = idx[m]
    s = nx * ny
    for (r, q, col) in ((0, 0, 0), (1, 1, 0), (0, 2, 1), (1, 3, 1)):
        rows.append(2 * a + r); cols.append(2 * (a + s) + col); vals.append(uz[m, q])
        rows.append(2 * (a + s) + r); cols.append(2 * a + col); vals.append(lz[m, q])
    for (u, l, s, m) in (("ux", "lx", 1, i + 1 < nx), ("uy", "ly", nx, j + 1 < ny)):
        uu = c[u].reshape(N, 2)
        ll = c[l].reshape(N, 2)
        a = idx[m]
        for r in (0, 1):
            rows.append(2 * a + r); cols.append(2 * (a + s)); vals.append(uu[m, r])
            rows.append(2 * (a + s) + r); cols.append(2 * a); vals.append(ll[m, r])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(2 * N, 2 * N))
    return A


class PressurePre:
    """The reference's own preconditioner (PressurePre -> VCycleCG: semicoarsened multigrid
    with symmetric y-plane / x-line Gauss-Seidel smoothing and z-line Thomas solves) built
    on an Oracle's geometry and weights.  ``attach`` makes it the PSolve1 of that oracle's
    forward Krylov solves."""

    def __init__(self, orc: Oracle, ref_level=4, gauss_iter2=1, gauss_iter3=1, mult_iter2=1,
                 mult_iter3=1):
        L = lib()
        self.orc = orc
        self.ctl = MGControl(ref_level, gauss_iter2, gauss_iter3, mult_iter2, mult_iter3)
        DTU, DTV, DTW, DUT = (_c(a) for a in orc._face_weights())
        self._w = (DTU, DTV, DTW, DUT)
        self._h = L.orc_mg_build(C.byref(orc.g), C.byref(self.ctl),
                                 *[_p(a) for a in orc._keep], _p(DTU), _p(DTV), _p(DTW),
                                 _p(DUT))
        if not self._h:
            raise ValueError("invalid multigrid control")

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self.detach()
                lib().orc_mg_free(self._h)
                self._h = None
        except Exception:
            pass

    def __call__(self, r):
        r = _c(r)
        x = np.empty_like(r)
        lib().orc_mg_pressurepre(self._h, _p(x), _p(r))
        return x

    def levels(self, level3):
        dims = (C.c_int * 3)()
        nq = lib().orc_mg_levels(self._h, level3, dims)
        return tuple(dims), nq

    def _lvl(self, fname, level3, x, nout):
        y = np.empty(nout)
        rc = getattr(lib(), fname)(self._h, level3, _p(y), _p(_c(x)))
        if rc:
            raise ValueError(f"{fname}: no level {level3}")
        return y

    def size(self, level3):
        (nx, ny, nz), _ = self.levels(level3)
        return nx * ny * nz * self.orc.prob.np1

    def apply_level(self, level3, x):
        return self._lvl("orc_mg_apply_level", level3, x, self.size(level3))

    def restrict(self, level3, r):
        return self._lvl("orc_mg_restrict", level3, r, self.size(level3 + 1))

    def prolong(self, level3, xc):
        return self._lvl("orc_mg_prolong", level3, xc, self.size(level3))

    def attach(self):
        lib().orc_set_pressurepre(self.orc._h, self._h)

    def detach(self):
        if getattr(self.orc, "_h", None):
            lib().orc_set_pressurepre(self.orc._h, None)
