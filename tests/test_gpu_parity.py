"""GPU parity: the HIP library against the FP64 CPU oracle on the same seeded inputs.

Bars (from the stage-solve contract): plain FP64 solve within 1e-10 relative of the
oracle's solution and iteration count within +-1; operator pieces to FP64 round-off;
mixed precision reaches the 1e-8 relative residual with its error vs the FP64
oracle solve stated (printed) for every case.
"""
import numpy as np
import pytest

from asam_amd import workload as W

pytestmark = pytest.mark.gpu

OP_TOL = 1e-12     # operator / preconditioner pieces (FP64 round-off)
SOL_TOL = 1e-10    # FP64 solve vs oracle solve


def small(np1, grad_full=1, nx=70, ny=5, nz=7, H=600.0, seed=3):
    return W.make_problem(nx, ny, nz, 1000.0, 1000.0, 250.0, H=H, a=2500.0,
                          ridge_xc=0.5 * nx * 1000.0, np1=np1, grad_full=grad_full, seed=seed)


def live(p, v):
    """Zero v on dead (fully buried, V=0) cells: Krylov vectors vanish there because
    b does (the reference's 1e-40 dead-cell diagonal, TriDiagT_Mod.f90:101, would
    otherwise scale them by 1e40; the GPU masks such cells)."""
    m = np.repeat((p.VolC.reshape(-1) > 0.0), p.np1)
    return np.where(m, v, 0.0)


def rel(a, b):
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


CASES = [
    dict(np1=1, grad_full=1), dict(np1=1, grad_full=0),
    dict(np1=2, grad_full=1), dict(np1=2, grad_full=0),
    dict(np1=1, grad_full=1, nx=1, ny=1, nz=9), dict(np1=2, grad_full=1, nx=3, ny=1, nz=1),
    dict(np1=2, grad_full=1, nx=130, ny=3, nz=2),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
def test_operator_pieces(case, oracle_mod, sound_mod):
    p = small(**case)
    orc = oracle_mod.Oracle(p)
    s = sound_mod.SoundSolver.from_problem(p)
    x = live(p, W.normal(p.n, 11))
    for name, f_gpu, f_orc in (("MatVec1", s.MatVec1, orc.matvec1),
                               ("MatVec1_T", s.MatVec1_T, orc.matvec1_T),
                               ("PSolve1", s.PSolve1, orc.psolve1),
                               ("PSolve1_T", s.PSolve1_T, orc.psolve1_T)):
        g, o = f_gpu(x), f_orc(x)
        e = rel(g, o)
        print(f"{case} {name}: rel err {e:.3e}")
        assert e < OP_TOL, (name, e)


@pytest.mark.parametrize("np1", [1, 2])
def test_setup_weights_matches_state_setup(np1, oracle_mod, sound_mod):
    p = small(np1)
    orc = oracle_mod.Oracle(p)
    DTU, DTV, DTW, DUT = orc.weights()
    s1 = sound_mod.SoundSolver.from_problem(p)
    g = p.grid
    s2 = sound_mod.SoundSolver(g.nx, g.ny, g.nz, np1, 1, p.beta0_dt, p.grav)
    s2.setup_weights(p.FU, p.FV, p.FW, p.VolC, p.dxv, p.dyv, p.dzv, DTU, DTV, DTW, DUT)
    x = live(p, W.normal(p.n, 5))
    assert rel(s1.MatVec1(x), s2.MatVec1(x)) < 1e-14


SOLVE_CASES = [
    ("cfg1-T", lambda: W.config(1, np1=1, scale=0.5)),
    ("cfg1-TB", lambda: W.config(1, np1=2, scale=0.5)),
    ("cfg4-T", lambda: W.config(4, np1=1, scale=0.5)),
    ("cfg4-TB", lambda: W.config(4, np1=2, scale=0.5)),
    ("ragged-TB", lambda: small(2, nx=70, ny=9, nz=11)),
]


@pytest.mark.parametrize("name,mk", SOLVE_CASES, ids=[c[0] for c in SOLVE_CASES])
def test_bicgstab_fp64_parity(name, mk, oracle_mod, sound_mod):
    p = mk()
    orc = oracle_mod.Oracle(p)
    xo, mo, to, so = orc.bicgstab(p.b, 300, 1e-8)
    s = sound_mod.SoundSolver.from_problem(p)
    x, mg, tg, info = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
    e = rel(x, xo)
    print(f"{name}: oracle MaxIter={mo} Tol={to:.3e}; gpu MaxIter={mg} Tol={tg:.3e} "
          f"rel diff {e:.3e}, {info.ms:.2f} ms")
    assert so == 0 and info.status == 0
    assert abs(abs(mg) - abs(mo)) <= 1
    assert e < SOL_TOL
    r = p.b - orc.matvec1(x)
    assert np.linalg.norm(r) <= 1.0001e-8 * np.linalg.norm(p.b)


@pytest.mark.parametrize("method", ["FGMRES", "FGMRESV", "GMRES"])
@pytest.mark.parametrize("np1", [1, 2])
def test_gmres_fp64_parity(method, np1, oracle_mod, sound_mod):
    p = W.config(4, np1=np1, scale=0.5)
    orc = oracle_mod.Oracle(p)
    f = {"FGMRES": orc.fgmres, "FGMRESV": orc.fgmresv, "GMRES": orc.gmres}[method]
    xo, ito, rro = f(p.b, 20, 1e-8, 10)
    s = sound_mod.SoundSolver.from_problem(p)
    s.set_control(method=method, step_gmres=10)
    x, mi, tl, info = s.SolveSound(np.zeros(p.n), p.b, 20, 1e-8)
    e = rel(x, xo)
    print(f"{method} np1={np1}: oracle steps {ito} rel {rro:.2e}; gpu steps {info.iterations} "
          f"cycles {info.restarts} rel {info.rel_residual:.2e}; diff {e:.3e}")
    assert abs(info.iterations - ito) <= 1
    assert e < SOL_TOL
    assert mi == 20  # FGMRES/GMRES leave MaxIter unchanged
    if method == "FGMRESV":  # the adjoint solve runs the same iteration with -A^T, -M^-T
        xo, ito, _ = f(p.b, 20, 1e-8, 10, transpose=True)
        x, mi, tl, info = s.SolveSound_T(np.zeros(p.n), p.b, 20, 1e-8)
        print(f"{method} np1={np1} transposed: oracle steps {ito}, gpu {info.iterations}, "
              f"diff {rel(x, xo):.3e}")
        assert abs(info.iterations - ito) <= 1 and rel(x, xo) < SOL_TOL


@pytest.mark.parametrize("m", [1, 3, 32])
@pytest.mark.parametrize("method", ["FGMRES", "FGMRESV"])
def test_gmres_restart_lengths(method, m, oracle_mod, sound_mod):
    """Restart lengths from 1 (a single projection per cycle) to the largest the library
    takes.  m = 32 runs one whole cycle with Tol = 0: every projection of 32 Arnoldi steps
    (for classical Gram-Schmidt, 8 groups of 4 at the last step), to round-off."""
    p = W.config(4, np1=2, scale=0.5)
    orc = oracle_mod.Oracle(p)
    f = orc.fgmresv if method == "FGMRESV" else orc.fgmres
    cycles, tol = {1: (60, 1e-8), 3: (20, 1e-8), 32: (1, 0.0)}[m]
    xo, ito, rro = f(p.b, cycles, tol, m)
    s = sound_mod.SoundSolver.from_problem(p)
    s.set_control(method=method, step_gmres=m)
    x, mi, tl, info = s.SolveSound(np.zeros(p.n), p.b, cycles, tol)
    e = rel(x, xo)
    if m == 32:
        assert ito == 32 and info.iterations == 32
    print(f"{method} m={m}: oracle steps {ito} rel {rro:.2e}; gpu steps {info.iterations} "
          f"cycles {info.restarts} rel {info.rel_residual:.2e}; diff {e:.3e}")
    assert abs(info.iterations - ito) <= 1
    assert e < SOL_TOL


@pytest.mark.parametrize("np1", [1, 2])
def test_transposed_solve_parity(np1, oracle_mod, sound_mod):
    p = W.config(4, np1=np1, scale=0.5)
    orc = oracle_mod.Oracle(p)
    xo, mo, to, so = orc.bicgstab(p.b, 300, 1e-8, transpose=True)
    s = sound_mod.SoundSolver.from_problem(p)
    x, mg, tg, info = s.SolveSound_T(np.zeros(p.n), p.b, 300, 1e-8)
    e = rel(x, xo)
    print(f"adjoint np1={np1}: oracle {mo}, gpu {mg}, diff {e:.3e}")
    assert abs(abs(mg) - abs(mo)) <= 1
    assert e < SOL_TOL


@pytest.mark.parametrize("method", ["BICGSTAB", "FGMRES", "FGMRESV"])
@pytest.mark.parametrize("np1", [1, 2])
def test_mixed_precision_reaches_1e8(np1, method, oracle_mod, sound_mod):
    """FP32 inner solves (BiCGStab, or one FGMRES / FGMRESV cycle per refinement step)
    with FP64 refinement reach the 1e-8 relative residual of the FP64 system (recomputed
    by the oracle's operator), forward and transposed; the error against the oracle's
    FP64 solve is stated."""
    p = W.config(4, np1=np1, scale=0.5)
    orc = oracle_mod.Oracle(p)
    s = sound_mod.SoundSolver.from_problem(p)
    s.set_control(method=method, precision="MIXED", step_gmres=10)
    for tr in (False, True):
        xo, mo, to, so = orc.bicgstab(p.b, 300, 1e-12, transpose=tr)
        solve = s.SolveSound_T if tr else s.SolveSound
        x, mi, tl, info = solve(np.zeros(p.n), p.b, 300, 1e-8)
        r = p.b - (orc.matvec1_T(x) if tr else orc.matvec1(x))
        rr = np.linalg.norm(r) / np.linalg.norm(p.b)
        err = rel(x, xo)
        print(f"mixed {method} np1={np1} {'transposed' if tr else 'forward'}: inner its "
              f"{info.iterations}, outer {info.restarts}, true rel residual {rr:.3e}, "
              f"error vs FP64 oracle {err:.3e}")
        assert info.status == 0
        assert rr <= 1.0001e-8
        assert err < 1e-6


@pytest.mark.parametrize("method", ["BICGSTAB", "FGMRES", "FGMRESV"])
def test_mixed_refinement_cap(method, oracle_mod, sound_mod):
    """With one refinement step allowed the mixed solve stops there and says it did not
    converge; the step it took still reduced the FP64 residual, and the returned x is that
    step's iterate."""
    p = W.config(4, np1=2, scale=0.5)
    orc = oracle_mod.Oracle(p)
    s = sound_mod.SoundSolver.from_problem(p)
    s.set_control(method=method, precision="MIXED", step_gmres=10, inner_tol=1e-3,
                  max_outer=1)
    x, mi, tl, info = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-12)
    rr = np.linalg.norm(p.b - orc.matvec1(x)) / np.linalg.norm(p.b)
    print(f"{method} capped at one refinement step: status {info.status}, steps "
          f"{info.restarts}, device rel residual {info.rel_residual:.3e}, true {rr:.3e}")
    assert info.status == 1 and info.restarts == 1
    assert rr < 1e-2 and abs(rr - info.rel_residual) <= 1e-6 * rr + 1e-15


@pytest.mark.parametrize("method,precision", [("BICGSTAB", "FP64"), ("BICGSTAB", "MIXED"),
                                              ("FGMRES", "FP64"), ("FGMRESV", "FP64"),
                                              ("GMRES", "FP64")])
@pytest.mark.parametrize("np1", [1, 2])
def test_zero_guess_residual_form(np1, method, precision, oracle_mod, sound_mod):
    """The initial residual of an all-zero guess comes from the x = 0 form (b outside the
    dead cells); with profiling on, the full apply computes it.  Both give the same bits,
    and a nonzero guess takes the full apply in either mode."""
    p = W.config(4, np1=np1, scale=0.5)
    dead = int(np.sum(p.VolC == 0.0))
    x1 = live(p, 1e-3 * W.normal(p.n, 5))
    maxiter = 300 if method == "BICGSTAB" else 20
    for x0 in (np.zeros(p.n), x1):
        out = []
        for prof in (False, True):
            s = sound_mod.SoundSolver.from_problem(p)
            s.set_control(method=method, precision=precision)
            s.profile(prof)
            out.append(s.SolveSound(x0, p.b, maxiter, 1e-8))
            s.close()
        (xa, ma, ta, ia), (xb, mb, tb, ib) = out
        print(f"np1={np1} {method} {precision} guess {'zero' if not x0.any() else 'nonzero'} "
              f"({dead} dead cells): iterations {ia.iterations}/{ib.iterations}, "
              f"bitwise {np.array_equal(xa, xb)}")
        assert np.array_equal(xa, xb) and ma == mb and ta == tb
        assert ia.iterations == ib.iterations and ia.status == ib.status == 0


@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("np1", [1, 2])
def test_line_windows_agree(np1, transpose, oracle_mod, sound_mod):
    """The FP64 solve (forward and transposed) matches the oracle for every load window of
    the line solve and every staging of its forward sweep: all levels in LDS, none (the
    output vector), and a 16 KiB budget that stages the lowest levels in LDS and the rest
    in the output vector.  The variants agree with each other to round-off (a wrong
    preconditioner still converges, but to a solution that differs at the tolerance
    level, about 1e-9 here)."""
    p = W.config(1, np1=np1, scale=0.5)
    xo, mo, _, _ = oracle_mod.Oracle(p).bicgstab(p.b, 300, 1e-8, transpose=transpose)
    s = sound_mod.SoundSolver.from_problem(p)
    solve = s.SolveSound_T if transpose else s.SolveSound
    ref = None
    for win in (1, 4, 8):
        for lds in (0, 1, 16):
            s.set_tuning(line_window=win, line_lds=lds)
            x, mg, tg, info = solve(np.zeros(p.n), p.b, 300, 1e-8)
            if ref is None:
                ref = x.copy()
            d, e = rel(x, ref), rel(x, xo)
            print(f"np1={np1} transpose={transpose} window {win} lds {lds}: vs window 1 "
                  f"{d:.2e}, vs oracle {e:.2e}, MaxIter {mg} (oracle {mo})")
            assert info.status == 0
            assert d < 1e-13
            assert e < SOL_TOL and abs(abs(mg) - abs(mo)) <= 1


@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("np1", [1, 2])
def test_line_cache_policy_same_bits(np1, transpose, oracle_mod, sound_mod):
    """The line solve's load cache policy (line_nt: plain, nontemporal, by footprint) changes
    no bit of the FP64 or the mixed solve, forward or transposed; the FP64 solve matches the
    oracle."""
    p = W.config(1, np1=np1, scale=0.5)
    xo, mo, _, _ = oracle_mod.Oracle(p).bicgstab(p.b, 300, 1e-8, transpose=transpose)
    s = sound_mod.SoundSolver.from_problem(p)
    solve = s.SolveSound_T if transpose else s.SolveSound
    for prec in ("FP64", "MIXED"):
        s.set_control(precision=prec)
        ref = None
        for nt in (0, 1, 2):
            s.set_tuning(line_nt=nt)
            x, mg, tg, info = solve(np.zeros(p.n), p.b, 300, 1e-8)
            print(f"np1={np1} transpose={transpose} {prec} line_nt {nt}: vs oracle "
                  f"{rel(x, xo):.2e}, status {info.status}")
            assert info.status == 0
            if ref is None:
                ref = x.copy()
                if prec == "FP64":
                    assert rel(x, xo) < SOL_TOL and abs(abs(mg) - abs(mo)) <= 1
            assert np.array_equal(x, ref)


def test_zero_rhs_breakdown_semantics(oracle_mod, sound_mod):
    p = small(1)
    s = sound_mod.SoundSolver.from_problem(p)
    b = np.zeros(p.n)
    x, mi, tl, info = s.SolveSound(np.zeros(p.n), b, 50, 1e-8)
    orc = oracle_mod.Oracle(p)
    xo, mo, to, so = orc.bicgstab(b, 50, 1e-8)
    assert np.all(x == 0) and mi == mo == 1 and tl == to == 0.0


def test_maxiter_exhaustion_semantics(oracle_mod, sound_mod):
    p = W.config(4, np1=1, scale=0.5)
    orc = oracle_mod.Oracle(p)
    xo, mo, to, so = orc.bicgstab(p.b, 3, 1e-14)
    s = sound_mod.SoundSolver.from_problem(p)
    x, mg, tg, info = s.SolveSound(np.zeros(p.n), p.b, 3, 1e-14)
    assert mo == 6 and mg == 6 and info.status == 1
    assert rel(x, xo) < 1e-12
    assert abs(tg - to) <= 1e-9 * abs(to)


def test_deterministic(sound_mod):
    p = W.config(4, np1=2, scale=0.5)
    s = sound_mod.SoundSolver.from_problem(p)
    x1, *_ = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
    x2, *_ = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
    assert np.array_equal(x1, x2)


@pytest.mark.parametrize("np1", [1, 2])
def test_config2_full_size_parity(np1, oracle_mod, sound_mod):
    """256x256x64 cut-cell system at its real size: FP64 and mixed vs the oracle."""
    p = W.config(2, np1=np1)
    orc = oracle_mod.Oracle(p)
    xo, mo, to, so = orc.bicgstab(p.b, 300, 1e-8)
    s = sound_mod.SoundSolver.from_problem(p)
    x, mg, tg, info = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
    e = rel(x, xo)
    print(f"cfg2 np1={np1} FP64: oracle {mo}, gpu {mg}, diff {e:.3e}, {info.ms:.2f} ms")
    assert abs(abs(mg) - abs(mo)) <= 1
    assert e < SOL_TOL
    s.set_control(precision="MIXED")
    xm, _, _, im = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
    rr = np.linalg.norm(p.b - orc.matvec1(xm)) / np.linalg.norm(p.b)
    em = rel(xm, xo)
    print(f"cfg2 np1={np1} mixed: inner {im.iterations} outer {im.restarts} true rel res "
          f"{rr:.3e} err vs FP64 oracle {em:.3e}, {im.ms:.2f} ms")
    assert rr <= 1.0001e-8


@pytest.mark.parametrize("method", ["FGMRES", "FGMRESV"])
def test_config2_full_size_fgmres(method, oracle_mod, sound_mod):
    """The reference's default method and its classical Gram-Schmidt variant on the
    256x256x64 MultiTriTB system at its real size: FP64 vs the oracle (steps +-1, 1e-10),
    and the mixed solve to the 1e-8 residual of the FP64 system."""
    p = W.config(2, np1=2)
    orc = oracle_mod.Oracle(p)
    f = orc.fgmresv if method == "FGMRESV" else orc.fgmres
    xo, ito, rro = f(p.b, 20, 1e-8, 10)
    s = sound_mod.SoundSolver.from_problem(p)
    s.set_control(method=method, step_gmres=10)
    x, mi, tl, info = s.SolveSound(np.zeros(p.n), p.b, 20, 1e-8)
    e = rel(x, xo)
    print(f"cfg2 {method} FP64: oracle steps {ito}, gpu {info.iterations}, diff {e:.3e}, "
          f"{info.ms:.2f} ms")
    assert abs(info.iterations - ito) <= 1 and e < SOL_TOL
    s.set_control(method=method, step_gmres=10, precision="MIXED")
    xm, _, _, im = s.SolveSound(np.zeros(p.n), p.b, 20, 1e-8)
    rr = np.linalg.norm(p.b - orc.matvec1(xm)) / np.linalg.norm(p.b)
    print(f"cfg2 {method} mixed: steps {im.iterations} refinements {im.restarts} true rel res "
          f"{rr:.3e} err vs FP64 oracle {rel(xm, xo):.3e}, {im.ms:.2f} ms")
    assert im.status == 0 and rr <= 1.0001e-8


def test_device_dot_matches_fsum(sound_mod):
    """Two-stage reduction (DPP wavefront sums, LDS across waves, fixed-order finalize)
    against a correctly rounded sum, within the FP64 bound n u sum|x_i y_i|, on lengths
    that leave partial wavefronts and workgroups; repeated calls agree bit for bit."""
    import math
    p = W.config(4, np1=1, scale=0.25)
    s = sound_mod.SoundSolver.from_problem(p)
    for n in (1, 63, 64, 65, 1000, 12345, (1 << 20) + 3):
        n = min(n, s.n)
        x = W.normal(s.n, 60 + n)
        y = W.uniform01(s.n, 61 + n) - 0.5
        xb, yb = s.vec(x), s.vec(y)
        d1, d2 = s.dot(xb, yb, n), s.dot(xb, yb, n)
        exact = math.fsum(x[:n] * y[:n])
        bound = n * 2.0 ** -53 * float(np.abs(x[:n] * y[:n]).sum())
        print(f"n={n}: dot {d1:.17e} fsum {exact:.17e} err {abs(d1 - exact):.2e} bound {bound:.2e}")
        assert d1 == d2
        assert abs(d1 - exact) <= bound


@pytest.mark.parametrize("precision", ["FP64", "MIXED"])
def test_autotuned_shapes_keep_parity(precision, oracle_mod, sound_mod):
    """The autotuner only changes launch shapes: the FP64 solve after it still matches the
    oracle (1e-10, iterations +-1) and the mixed solve still reaches 1e-8."""
    p = W.config(4, np1=2, scale=0.5)
    orc = oracle_mod.Oracle(p)
    s = sound_mod.SoundSolver.from_problem(p)
    s.set_control(precision=precision)
    t, ms = s.autotune(2)
    print(f"{precision}: autotuned {t}, ms per iteration {ms[0]:.4f}")
    assert ms[0] > 0
    x, mi, tl, info = s.SolveSound(np.zeros(p.n), p.b, 200, 1e-8)
    if precision == "FP64":
        xo, mo, to, so = orc.bicgstab(p.b, 200, 1e-8)
        assert rel(x, xo) < SOL_TOL and abs(abs(mi) - abs(mo)) <= 1
    else:
        rr = np.linalg.norm(p.b - orc.matvec1(x)) / np.linalg.norm(p.b)
        assert rr <= 1.0001e-8


@pytest.mark.parametrize("np1", [1, 2])
def test_moist_setup_matches_oracle(np1, oracle_mod, sound_mod):
    """asam_sound_setup_moist (vapour, cloud water, rain) gives the oracle's moist operator:
    MatVec1 to round-off, the FP64 solve to 1e-10 with iterations +-1; without moisture it
    is the dry set-up bit for bit."""
    p = small(np1)
    g = p.grid
    shape = (g.nz, g.ny, g.nx)
    rv = 0.015 * p.rho * W.uniform01(g.ncell, 81).reshape(shape)
    rl = 0.003 * p.rho * W.uniform01(g.ncell, 82).reshape(shape)
    rr = 0.002 * p.rho * W.uniform01(g.ncell, 83).reshape(shape)
    orc = oracle_mod.Oracle(p, weights=oracle_mod.Oracle(p).weights_moist(rv, rl, rr))
    s = sound_mod.SoundSolver(g.nx, g.ny, g.nz, np1, p.grad_full, p.beta0_dt, p.grav)
    s.setup_moist(p.FU, p.FV, p.FW, p.VolC, p.dxv, p.dyv, p.dzv, p.rho, p.rhotheta, p.T,
                  rv, rl, rr)
    x = live(p, W.normal(p.n, 84))
    e = rel(s.MatVec1(x), orc.matvec1(x))
    xo, mo, _, _ = orc.bicgstab(p.b, 300, 1e-8)
    xg, mg, _, info = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
    es = rel(xg, xo)
    print(f"np1={np1} moist: MatVec1 {e:.2e}, solve {es:.2e}, MaxIter {mg} (oracle {mo})")
    assert e < OP_TOL and es < SOL_TOL and abs(abs(mg) - abs(mo)) <= 1
    s.setup_moist(p.FU, p.FV, p.FW, p.VolC, p.dxv, p.dyv, p.dzv, p.rho, p.rhotheta, p.T)
    dry = sound_mod.SoundSolver.from_problem(p)
    assert np.array_equal(s.MatVec1(x), dry.MatVec1(x))


@pytest.mark.parametrize("np1", [1, 2])
def test_apply_variants_keep_parity(np1, oracle_mod, sound_mod):
    """Every apply variant (z-march, one cell per thread, band order with one band per XCD
    or 8-row bands, z-march of 2 or 4 rows per thread with row blocks overhanging ny) gives
    the oracle's MatVec1 to round-off and the FP64 solve to 1e-10, with the residual form of
    the refinement step and the all-zero-guess form included (mixed solve to 1e-8)."""
    p = small(np1, nx=70, ny=37, nz=11)
    orc = oracle_mod.Oracle(p)
    xo, mo, _, _ = orc.bicgstab(p.b, 300, 1e-8)
    x = live(p, W.normal(p.n, 85))
    s = sound_mod.SoundSolver.from_problem(p)
    for var, bx, rows, wg, rr in ((0, 256, 0, 4096, 4), (1, 64, 0, 4096, 4),
                                  (2, 256, 0, 512, 4), (2, 64, 8, 1024, 4),
                                  (2, 128, 16, 256, 4), (3, 256, 0, 4096, 2),
                                  (3, 64, 0, 4096, 4), (3, 128, 0, 256, 4),
                                  (3, 32, 0, 2048, 2), (3, 32, 0, 4096, 4)):
        s.set_tuning(apply_variant=var, apply_bx=bx, apply_band_rows=rows, apply_wg=wg,
                     apply_rows=rr)
        e = rel(s.MatVec1(x), orc.matvec1(x))
        et = rel(s.MatVec1_T(x), orc.matvec1_T(x))
        s.set_control(precision="FP64")
        xg, mg, _, info = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
        s.set_control(precision="MIXED")
        xm, _, _, im = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
        rr = np.linalg.norm(p.b - orc.matvec1(xm)) / np.linalg.norm(p.b)
        print(f"np1={np1} variant {var} bx {bx} rows {rows}/{rr} wg {wg}: MatVec1 {e:.2e} "
              f"transposed {et:.2e}, FP64 solve {rel(xg, xo):.2e} MaxIter {mg} (oracle {mo}), "
              f"mixed true residual {rr:.2e}")
        assert e < OP_TOL and et < OP_TOL
        assert rel(xg, xo) < SOL_TOL and abs(abs(mg) - abs(mo)) <= 1
        assert im.status == 0 and rr <= 1.0001e-8


@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("dims", [(70, 37, 11), (9, 1, 6), (5, 3, 4), (66, 5, 3)])
def test_deep_apply_same_bits(dims, transpose, oracle_mod, sound_mod):
    """The row-blocked apply with its loads two levels ahead (apply_depth = 2, used by the
    FP32 applies of the mixed solve with 2x2 blocks) does march_rows' arithmetic: the mixed
    solve gives the same bits at every width, XCD order and z-tiling (one tile of all
    levels, and tiles of a few levels), forward and transposed, on grids down to one row
    and fewer levels than the look-ahead."""
    p = small(2, nx=dims[0], ny=dims[1], nz=dims[2])
    orc = oracle_mod.Oracle(p)
    s = sound_mod.SoundSolver.from_problem(p)
    s.set_control(precision="MIXED")
    solve = s.SolveSound_T if transpose else s.SolveSound
    for bx in (32, 64, 128, 256):
        for wg, xcd in ((256, 1), (4096, 0)):
            out = []
            for depth in (1, 2):
                s.set_tuning(apply_variant=3, apply_rows=2, apply_bx=bx, apply_wg=wg,
                             apply_xcd=xcd, apply_depth=depth)
                out.append(solve(np.zeros(p.n), p.b, 300, 1e-8))
            (xa, ma, ta, ia), (xb, mb, tb, ib) = out
            op = orc.matvec1_T if transpose else orc.matvec1
            rr = np.linalg.norm(p.b - op(xb)) / np.linalg.norm(p.b)
            print(f"{dims} transpose={transpose} bx {bx} wg {wg} xcd {xcd}: max|dx| "
                  f"{np.abs(xa - xb).max():.1e}, inner {ia.iterations}/{ib.iterations}, "
                  f"true residual {rr:.2e}")
            assert ib.status == 0 and rr <= 1.0001e-8
            assert np.array_equal(xa, xb)
            assert (ia.iterations, ia.restarts, ia.abs_residual) == \
                (ib.iterations, ib.restarts, ib.abs_residual)


@pytest.mark.parametrize("np1", [1, 2])
@pytest.mark.parametrize("dims", [(9, 1, 6), (5, 3, 4), (66, 5, 3)])
def test_row_blocked_apply_thin_grids(np1, dims, oracle_mod, sound_mod):
    """The row-blocked z-march on grids thinner than its row blocks (one row, three rows,
    five rows against blocks of 2 and 4, widths 32 to 256): operator, transposed operator
    and the residual form against the oracle."""
    p = small(np1, nx=dims[0], ny=dims[1], nz=dims[2])
    orc = oracle_mod.Oracle(p)
    x = live(p, W.normal(p.n, 86))
    s = sound_mod.SoundSolver.from_problem(p)
    for bx, rr in ((32, 2), (32, 4), (64, 2), (64, 4), (256, 4)):
        s.set_tuning(apply_variant=3, apply_bx=bx, apply_rows=rr)
        e = rel(s.MatVec1(x), orc.matvec1(x))
        et = rel(s.MatVec1_T(x), orc.matvec1_T(x))
        print(f"np1={np1} {dims} bx {bx} rows {rr}: MatVec1 {e:.2e} transposed {et:.2e}")
        assert e < OP_TOL and et < OP_TOL
        s.set_control(precision="FP64")
        xg, mg, _, info = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
        xo, mo, _, _ = orc.bicgstab(p.b, 300, 1e-8)
        assert rel(xg, xo) < SOL_TOL and abs(abs(mg) - abs(mo)) <= 1


@pytest.mark.parametrize("transpose", [False, True])
def test_line_large_lds_budget(transpose, oracle_mod, sound_mod):
    """96-level FP64 2x2 columns staged entirely in LDS (96 KiB per workgroup, budgets 100
    and 128 KiB) and partly past 64 KiB (76 KiB: 76 levels in LDS, 20 in the output vector)
    against the default 60 KiB budget (60 levels in LDS, 36 in the output vector) and the
    oracle: FP64 solve to 1e-10 with the oracle's iteration count (+-1), forward and
    transposed."""
    p = W.make_problem(64, 8, 96, 1000.0, 1000.0, 150.0, H=600.0, a=2500.0,
                       ridge_xc=32000.0, np1=2, grad_full=1, seed=5)
    orc = oracle_mod.Oracle(p)
    xo, mo, _, _ = orc.bicgstab(p.b, 300, 1e-8, transpose=transpose)
    s = sound_mod.SoundSolver.from_problem(p)
    s.set_control(precision="FP64")
    solve = s.SolveSound_T if transpose else s.SolveSound
    ref = None
    for lds in (1, 76, 100, 128):
        s.set_tuning(line_lds=lds)
        x, mg, _, info = solve(np.zeros(p.n), p.b, 300, 1e-8)
        if ref is None:
            ref = x.copy()
        print(f"transpose={transpose} line_lds {lds}: vs oracle {rel(x, xo):.2e}, vs 60 KiB "
              f"{rel(x, ref):.2e}, MaxIter {mg} (oracle {mo})")
        assert info.status == 0 and rel(x, xo) < SOL_TOL and abs(abs(mg) - abs(mo)) <= 1
        assert rel(x, ref) < 1e-13
