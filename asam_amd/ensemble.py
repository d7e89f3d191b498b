"""Python host side of the batched (ensemble) stage solver over the gfx950 C-ABI
(include/asam_ensemble.h).

Each member is one SolveSound stage system (reference SRC/SOLU/Iter_Mod.f90:9-30, set up
as JacAccGrav, SRC/MULTI/JacAccGrav_Mod.f90:304); all members are solved together by
FP32-inner / FP64-refined BiCGStab in one HIP-graph replay.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import sound as S

_vp = C.c_void_p
_dp = C.POINTER(C.c_double)


class EnsTuning(C.Structure):
    _fields_ = [("apply_bx", C.c_int), ("apply_zsplit", C.c_int), ("apply_pf", C.c_int),
                ("line_window", C.c_int), ("line_lds", C.c_int), ("line_pf", C.c_int),
                ("blas_chunk", C.c_int), ("blas_pf", C.c_int), ("use_graph", C.c_int),
                ("apply_rows", C.c_int), ("reserved", C.c_int * 2)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class EnsStats(C.Structure):
    _fields_ = [("graph_launches", C.c_int), ("graph_nodes", C.c_int),
                ("eager_launches", C.c_int), ("outer_steps", C.c_int), ("ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_sigs_done = False


def load():
    global _sigs_done
    L = S.load()
    if _sigs_done:
        return L
    i, ip, lp = C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_long)
    sig = {
        "asam_ens_create": (i, [C.POINTER(_vp), i, C.POINTER(S.Grid)]),
        "asam_ens_destroy": (i, [_vp]),
        "asam_ens_last_error": (C.c_char_p, [_vp]),
        "asam_ens_setup": (i, [_vp, i] + [_dp] * 10),
        "asam_ens_length": (C.c_long, [_vp]),
        "asam_ens_member_range": (i, [_vp, i, lp, lp]),
        "asam_ens_set_control": (i, [_vp, C.POINTER(S.Control)]),
        "asam_ens_solve": (i, [_vp, _dp, _dp, C.c_double, C.POINTER(S.Info)]),
        "asam_ens_solve_device": (i, [_vp, _vp, _vp, C.c_double, C.POINTER(S.Info)]),
        "asam_ens_get_tuning": (i, [_vp, C.POINTER(EnsTuning)]),
        "asam_ens_set_tuning": (i, [_vp, C.POINTER(EnsTuning)]),
        "asam_ens_autotune": (i, [_vp, i, C.POINTER(EnsTuning), _dp]),
        "asam_ens_get_stats": (i, [_vp, C.POINTER(EnsStats)]),
        "asam_ens_dev_alloc": (i, [_vp, C.POINTER(_vp), C.c_size_t]),
        "asam_ens_dev_free": (i, [_vp, _vp]),
        "asam_ens_dev_upload": (i, [_vp, _vp, _vp, C.c_size_t]),
        "asam_ens_dev_download": (i, [_vp, _vp, _vp, C.c_size_t]),
        "asam_ens_dev_memset": (i, [_vp, _vp, i, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _sigs_done = True
    return L


class EnsBuffer:
    """A device allocation of an ensemble context (e.g. an ensemble vector)."""

    def __init__(self, ens: "Ensemble", nbytes: int):
        self.ens = ens
        self.nbytes = nbytes
        p = _vp()
        ens._chk(ens.L.asam_ens_dev_alloc(ens.h, C.byref(p), nbytes))
        self.ptr = p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        self.ens._chk(self.ens.L.asam_ens_dev_upload(self.ens.h, self.ptr, a.ctypes.data_as(_vp),
                                                     a.nbytes))
        return self

    def download(self, dtype=np.float64):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        self.ens._chk(self.ens.L.asam_ens_dev_download(self.ens.h, out.ctypes.data_as(_vp),
                                                       self.ptr, self.nbytes))
        return out

    def zero(self):
        self.ens._chk(self.ens.L.asam_ens_dev_memset(self.ens.h, self.ptr, 0, self.nbytes))
        return self

    def free(self):
        if self.ptr and getattr(self.ens, "h", None):
            self.ens.L.asam_ens_dev_free(self.ens.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Ensemble:
    """Uneven ensemble of stage systems on the current GPU."""

    def __init__(self, problems, setup=True):
        self.L = load()
        problems = list(problems)
        grids = (S.Grid * len(problems))()
        for m, p in enumerate(problems):
            g = p.grid
            grids[m] = S.Grid(g.nx, g.ny, g.nz, p.np1, p.grad_full, p.beta0_dt, p.grav, 1.0)
        self.h = _vp()
        rc = self.L.asam_ens_create(C.byref(self.h), len(problems), grids)
        if rc != 0:
            msg = self.L.asam_ens_last_error(self.h).decode() if self.h else ""
            self.close()
            raise S.AsamError(f"asam_ens_create failed ({rc}) {msg}")
        self.nmem = len(problems)
        self.n = int(self.L.asam_ens_length(self.h))
        self.ranges = []
        for m in range(self.nmem):
            off, ln = C.c_long(), C.c_long()
            self._chk(self.L.asam_ens_member_range(self.h, m, C.byref(off), C.byref(ln)))
            self.ranges.append((off.value, ln.value))
        if setup:
            for m, p in enumerate(problems):
                self.setup_member(m, p)

    def setup_member(self, m, p):
        """Operator set-up of member m from a SoundProblem (JacAccGrav inputs)."""
        a = [S._c(v) for v in (p.FU, p.FV, p.FW, p.VolC, p.dxv, p.dyv, p.dzv, p.rho,
                               p.rhotheta, p.T)]
        self._chk(self.L.asam_ens_setup(self.h, m, *[S._p(v) for v in a]))

    def _chk(self, rc):
        if rc != 0:
            raise S.AsamError(f"libasam_sound ensemble error {rc}: "
                              f"{self.L.asam_ens_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.asam_ens_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ vectors
    def concat(self, vecs):
        return np.concatenate([np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
                               for v in vecs])

    def split(self, x):
        return [x[o:o + n] for o, n in self.ranges]

    def buffer(self, data=None):
        b = EnsBuffer(self, self.n * 8)
        if data is not None:
            b.upload(np.ascontiguousarray(data, dtype=np.float64))
        return b

    # ------------------------------------------------------------------ controls
    def set_control(self, inner_tol=1e-5, max_inner=200, max_outer=12):
        c = S.Control(S.METHODS["BICGSTAB"], 10, S.PRECISIONS["MIXED"], inner_tol, max_inner,
                      max_outer)
        self._chk(self.L.asam_ens_set_control(self.h, C.byref(c)))

    def get_tuning(self) -> EnsTuning:
        t = EnsTuning()
        self._chk(self.L.asam_ens_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, **kw):
        t = self.get_tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        self._chk(self.L.asam_ens_set_tuning(self.h, C.byref(t)))

    def autotune(self, reps=3):
        t = EnsTuning()
        ms = (C.c_double * 3)()
        self._chk(self.L.asam_ens_autotune(self.h, reps, C.byref(t), ms))
        return t.as_dict(), list(ms)

    def stats(self) -> dict:
        s = EnsStats()
        self._chk(self.L.asam_ens_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    # ------------------------------------------------------------------ solves
    def solve(self, b, tol=1e-8, x0=None):
        """Host arrays: b (concatenated or a list of member vectors) -> x, infos."""
        if isinstance(b, (list, tuple)):
            b = self.concat(b)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        infos = (S.Info * self.nmem)()
        self._chk(self.L.asam_ens_solve(self.h, S._p(x), S._p(b), tol, infos))
        return x, [infos[m] for m in range(self.nmem)]

    def solve_device(self, xbuf: EnsBuffer, bbuf: EnsBuffer, tol=1e-8):
        infos = (S.Info * self.nmem)()
        self._chk(self.L.asam_ens_solve_device(self.h, xbuf.ptr, bbuf.ptr, tol, infos))
        return [infos[m] for m in range(self.nmem)]
