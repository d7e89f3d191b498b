"""Transport stage solve on the GPU against the oracle (oracle/transport.c):
  - the operator (MatVecChem / MatVecMet) and the Gauss-Seidel preconditioner (PSolve)
    equal the oracle bit for bit (both round every product and sum separately, in the
    reference's order) on ragged grids, with one or two z-slots per lane,
    JacTMom / JacFall components, and every launch shape of the pipelined sweep (one row
    per workgroup hands every row over between workgroups);
  - BiCGStab (Chem and Met forms) matches the oracle's solution within 1e-10 relative,
    the iteration count within one, and the MaxIter/Tol return conventions;
  - device entry points equal the host ones; couplings across the linear-index wrap are
    rejected."""
import numpy as np
import pytest

from asam_amd import workload as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def trm(sound_mod):
    from asam_amd import transport
    transport.load()
    return transport


@pytest.fixture(scope="module")
def orc(oracle_mod):
    from oracle import pytransport
    return pytransport


def problem(nx, ny, nz, seed=3, **kw):
    p = W.make_problem(nx, ny, nz, 500.0, 500.0, 250.0, H=700.0, a=1200.0,
                       ridge_xc=0.5 * nx * 500.0, np1=1, seed=seed)
    return W.transport_problem(p, **kw)


SHAPES = [(5, 3, 2), (24, 17, 9), (6, 5, 70), (40, 36, 30)]


def report(name, a, b):
    d = np.abs(a - b)
    rel = float(d.max() / max(np.abs(b).max(), 1e-300))
    print(f"{name}: bitwise={np.array_equal(a, b)} max rel diff {rel:.2e}")
    return rel


@pytest.mark.parametrize("shape", SHAPES)
def test_matvec_equals_oracle(shape, trm, orc):
    tp = problem(*shape, ncomp=3, fall_comp=1, mom_comps=(2,), met=True)
    s = trm.TransportSolver.from_problem(tp)
    o = orc.TransportOracle(tp)
    x = W.normal(tp.ncomp * tp.m, 41).reshape(tp.ncomp, tp.m)
    for init in (False, True):
        y = s.MatVec(x, init=init)
        yo = o.matvec(x, init=init)
        print(f"shape {shape} init {init}")
        report("matvec", y, yo)
        assert np.array_equal(y, yo)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("tune", [dict(gs_rows=8, gs_chunk=4), dict(gs_rows=1, gs_chunk=2),
                                  dict(gs_rows=3, gs_chunk=4), dict(gs_rows=5, gs_chunk=2)])
def test_psolve_equals_oracle(shape, tune, trm, orc):
    tp = problem(*shape, ncomp=3, fall_comp=1, mom_comps=(2,))
    s = trm.TransportSolver.from_problem(tp)
    s.set_tuning(**tune)
    o = orc.TransportOracle(tp)
    x = W.normal(tp.ncomp * tp.m, 43).reshape(tp.ncomp, tp.m)
    print(f"shape {shape} tuning {s.get_tuning().as_dict()}", flush=True)
    y = s.PSolve(x)
    yo = o.psolve(x)
    report("psolve", y, yo)
    assert np.array_equal(y, yo)
    # repeated calls give the same bits (no state leaks between sweeps or calls)
    assert np.array_equal(s.PSolve(x), y)


@pytest.mark.parametrize("iterg", [0, 2])
def test_psolve_iterg(iterg, trm, orc):
    tp = problem(12, 9, 7, ncomp=2, iterg=iterg)
    s = trm.TransportSolver.from_problem(tp)
    x = W.normal(tp.ncomp * tp.m, 45).reshape(tp.ncomp, tp.m)
    y, yo = s.PSolve(x), orc.TransportOracle(tp).psolve(x)
    report(f"psolve iterg={iterg}", y, yo)
    assert np.array_equal(y, yo)


@pytest.mark.parametrize("met", [False, True])
@pytest.mark.parametrize("shape", [(24, 17, 9), (40, 36, 30), (6, 5, 70)])
def test_bicgstab_matches_oracle(shape, met, trm, orc):
    tp = problem(*shape, ncomp=3, fall_comp=1, mom_comps=(2,), met=met, wind=25.0, noise=4.0)
    s = trm.TransportSolver.from_problem(tp)
    o = orc.TransportOracle(tp)
    for tol in (1e-6, 1e-12):
        x, mi, t, info = s.BiCGStab(np.zeros_like(tp.b), tp.b, 60, tol)
        xo, mio, to, sto = o.bicgstab(tp.b, 60, tol)
        print(f"shape {shape} met {met} tol {tol}: GPU MaxIter {mi} Tol {t:.3e} "
              f"({info.iterations} its, {info.psolves} psolves); oracle {mio} {to:.3e}")
        assert info.status == 0 and sto == 0
        assert abs(abs(mi) - abs(mio)) <= 1 and (mi < 0) == (mio < 0)
        inner = np.concatenate([tp.interior(x).ravel()])
        innero = np.concatenate([tp.interior(xo).ravel()])
        err = np.linalg.norm(inner - innero) / np.linalg.norm(innero)
        print(f"  error vs oracle {err:.2e}")
        assert err <= 1e-10
        if mi == mio:
            assert abs(t - to) <= 1e-6 * to + 1e-300


@pytest.mark.parametrize("tune", [dict(gs_rows=8, gs_chunk=2), dict(gs_rows=2, gs_chunk=4)])
def test_limits_nz126_16_components(tune, trm, orc):
    """The largest column (nz = 126: Z = 128, two full wavefronts per row) and the most
    components (16, every operator kind, Met form): operator and preconditioner bitwise,
    the solve within 1e-10 of the oracle."""
    p = W.make_problem(7, 4, 126, 500.0, 500.0, 100.0, H=700.0, a=1200.0,
                       ridge_xc=0.5 * 7 * 500.0, np1=1, seed=9)
    tp = W.transport_problem(p, ncomp=16, fall_comp=5, mom_comps=(2, 11), met=True, wind=25.0,
                             noise=4.0)
    s = trm.TransportSolver.from_problem(tp)
    s.set_tuning(**tune)
    o = orc.TransportOracle(tp)
    x = W.normal(tp.ncomp * tp.m, 51).reshape(tp.ncomp, tp.m)
    print(f"tuning {s.get_tuning().as_dict()}", flush=True)
    assert np.array_equal(s.MatVec(x), o.matvec(x))
    assert np.array_equal(s.PSolve(x), o.psolve(x))
    xg, mi, t, info = s.BiCGStab(np.zeros_like(tp.b), tp.b, 60, 1e-12)
    xo, mio, to, sto = o.bicgstab(tp.b, 60, 1e-12)
    err = np.linalg.norm(tp.interior(xg) - tp.interior(xo)) / np.linalg.norm(tp.interior(xo))
    print(f"MaxIter {mi}/{mio}, error vs oracle {err:.2e}")
    assert info.status == 0 and sto == 0 and abs(abs(mi) - abs(mio)) <= 1 and err <= 1e-10


def test_return_conventions(trm, orc):
    tp = problem(16, 12, 10, ncomp=2, wind=25.0, noise=4.0)
    s = trm.TransportSolver.from_problem(tp)
    _, mi, t, info = s.BiCGStab(np.zeros_like(tp.b), tp.b, 1, 1e-14)
    assert mi == 2 and t > 0 and info.status == 1     # not converged: 2*Iter
    z, mi, t, info = s.BiCGStab(np.zeros_like(tp.b), np.zeros_like(tp.b), 5, 1e-6)
    assert mi == 1 and t == 0.0 and info.status == 2 and not np.any(z)  # rho == 0


def test_device_entry_points(trm, orc):
    tp = problem(20, 14, 11, ncomp=2, fall_comp=0)
    s = trm.TransportSolver.from_problem(tp)
    x = W.normal(tp.ncomp * tp.m, 47)
    dx, dy, db = s.device_vector().upload(x), s.device_vector(), s.device_vector().upload(tp.b)
    s.matvec_device(dx, dy)
    assert np.array_equal(dy.download(), s.MatVec(x).ravel())
    s.psolve_device(dx, dy)
    assert np.array_equal(dy.download(), s.PSolve(x).ravel())
    dx.upload(np.zeros(tp.ncomp * tp.m))
    mi, t, info = s.solve_device(dx, db, 30, 1e-8)
    xh, mih, th, _ = s.BiCGStab(np.zeros_like(tp.b), tp.b, 30, 1e-8)
    assert mi == mih and np.array_equal(dx.download(), xh.ravel())
    for v in (dx, dy, db):
        v.free()


def test_bounded_wait_reports_error(trm, orc):
    """A block whose boundary row never arrives stops waiting after gs_spin_limit polls:
    the call returns ASAM_E_HIP instead of hanging, every kernel finishes, and the context
    sweeps correctly again once the hold is lifted."""
    from asam_amd import sound
    tp = problem(20, 30, 9, ncomp=2)
    s = trm.TransportSolver.from_problem(tp)
    s.set_tuning(gs_rows=2, gs_chunk=2, gs_spin_limit=64)
    x = W.normal(tp.ncomp * tp.m, 49).reshape(tp.ncomp, tp.m)
    s.test_hold_boundary(True)
    with pytest.raises(sound.AsamError, match="waited past its limit"):
        s.PSolve(x)
    s.test_hold_boundary(False)
    s.set_tuning(gs_spin_limit=0)
    assert np.array_equal(s.PSolve(x), orc.TransportOracle(tp).psolve(x))


def test_wrap_coupling_rejected(trm):
    from asam_amd import sound
    tp = problem(6, 4, 3, ncomp=1)
    s = trm.TransportSolver(tp.nx, tp.ny, tp.nz, 1)
    bad = tp.tpot.copy()
    X = tp.nx + 2
    bad[2, X - 1] = -0.1    # Val(i,3) of the last cell of an x-row: couples to the next row
    with pytest.raises(sound.AsamError, match="consecutive"):
        s.set_matrix(trm.JACTPOT, bad)
    with pytest.raises(sound.AsamError, match="DiagPtr"):
        s.set_matrix(trm.JACTPOT, tp.tpot, diagptr=np.array([-1, -2, -3, 0, 1, 2, 3]))
    with pytest.raises(sound.AsamError, match="JacTPot"):
        s.PSolve(np.zeros(tp.m))
