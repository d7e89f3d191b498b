"""The named workloads at their full sizes on the GPU, against the FP64 CPU oracle.

- 64x64x40 box without terrain (T and TB): plain FP64 BiCGStab within 1e-10 of the
  oracle's solution with iterations +-1; the mixed solve reaches the 1e-8 relative
  residual of the FP64 system (recomputed by the oracle's operator).
- 512x512x96 cut-cell ridge (TB, 50.3 M unknowns): the same FP64 bar, and the mixed
  solve to 1e-8 with its error against the oracle's FP64 solve bounded.
- 64 uneven members (nx, ny ~ U{64..256}, nz ~ U{40..64}, TB) solved together in one
  HIP-graph replay: every member's residual recomputed by the oracle's operator, and its
  error against the oracle's FP64 solve stated and bounded.
"""
import numpy as np
import pytest

from asam_amd import workload as W

pytestmark = pytest.mark.gpu

SOL_TOL = 1e-10


def rel(a, b):
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


@pytest.mark.parametrize("np1", [1, 2])
def test_config1_full_size(np1, oracle_mod, sound_mod):
    p = W.config(1, np1=np1)
    assert (p.grid.nx, p.grid.ny, p.grid.nz) == (64, 64, 40)
    orc = oracle_mod.Oracle(p)
    xo, mo, to, so = orc.bicgstab(p.b, 300, 1e-8)
    s = sound_mod.SoundSolver.from_problem(p)
    x, mg, tg, info = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
    e = rel(x, xo)
    print(f"cfg1 64x64x40 np1={np1} FP64: oracle MaxIter {mo}, gpu {mg}, diff {e:.3e}, "
          f"{info.ms:.2f} ms")
    assert so == 0 and info.status == 0
    assert abs(abs(mg) - abs(mo)) <= 1 and (mg < 0) == (mo < 0)
    assert e < SOL_TOL
    s.set_control(precision="MIXED")
    xm, _, _, im = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
    rr = np.linalg.norm(p.b - orc.matvec1(xm)) / np.linalg.norm(p.b)
    em = rel(xm, xo)
    print(f"cfg1 np1={np1} mixed: inner {im.iterations} refinements {im.restarts}, true rel "
          f"residual {rr:.3e}, error vs FP64 oracle {em:.3e}")
    assert im.status == 0 and rr <= 1.0001e-8
    assert em < 1e-7


def test_config3_full_size(oracle_mod, sound_mod):
    """512x512x96, ridge 2 km: FP64 and mixed against the oracle.  The oracle's solve of
    this system takes seconds on the host threads."""
    p = W.config(3, np1=2)
    assert (p.grid.nx, p.grid.ny, p.grid.nz) == (512, 512, 96)
    s = sound_mod.SoundSolver.from_problem(p)
    x, mg, tg, info = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
    s.set_control(precision="MIXED")
    xm, _, _, im = s.SolveSound(np.zeros(p.n), p.b, 300, 1e-8)
    s.close()
    orc = oracle_mod.Oracle(p)
    xo, mo, to, so = orc.bicgstab(p.b, 300, 1e-8)
    e = rel(x, xo)
    print(f"cfg3 512x512x96 FP64: oracle MaxIter {mo}, gpu {mg}, diff {e:.3e}, "
          f"{info.ms:.2f} ms", flush=True)
    assert so == 0 and info.status == 0
    assert abs(abs(mg) - abs(mo)) <= 1
    assert e < SOL_TOL
    bn = np.linalg.norm(p.b)
    rr = np.linalg.norm(p.b - orc.matvec1(xm)) / bn
    em = rel(xm, xo)
    print(f"cfg3 mixed: inner {im.iterations} refinements {im.restarts}, true rel residual "
          f"{rr:.3e}, error vs FP64 oracle {em:.3e}, {im.ms:.2f} ms", flush=True)
    assert im.status == 0 and rr <= 1.0001e-8
    assert em < 1e-7


@pytest.fixture(scope="module")
def config5(sound_mod):
    """The 64-member ensemble, solved once to learn its schedule and then by one graph
    replay; returns the members, their solutions and infos, and the replay's stats."""
    from asam_amd import ensemble
    ps = W.ensemble_members(64, seed=5, np1=2)
    E = ensemble.Ensemble(ps)
    b = E.concat([p.b for p in ps])
    E.solve(b, 1e-8)
    x, infos = E.solve(b, 1e-8)
    st = E.stats()
    xs = [v.copy() for v in E.split(x)]
    E.close()
    print(f"config 5: {len(ps)} members, {sum(p.grid.ncell for p in ps) / 1e6:.1f} M cells, "
          f"replay stats {st}", flush=True)
    return ps, xs, infos, st


@pytest.mark.parametrize("part", range(4))
def test_config5_ensemble_full_size(part, config5, oracle_mod):
    ps, xs, infos, st = config5
    assert len(ps) == 64
    assert st["graph_launches"] == 1 and st["graph_nodes"] > 0
    for m in range(16 * part, 16 * part + 16):
        p, xm, inf = ps[m], xs[m], infos[m]
        g = p.grid
        assert 64 <= g.nx <= 256 and 64 <= g.ny <= 256 and 40 <= g.nz <= 64
        orc = oracle_mod.Oracle(p)
        rr = float(np.linalg.norm(p.b - orc.matvec1(xm)) / np.linalg.norm(p.b))
        xo, _, _, _ = orc.bicgstab(p.b, 500, 1e-12)
        err = rel(xm, xo)
        print(f"member {m} {g.nx}x{g.ny}x{g.nz}: {inf.iterations} inner its, {inf.restarts} "
              f"refinements, rel residual {rr:.2e} (device {inf.rel_residual:.2e}), error vs "
              f"FP64 oracle {err:.2e}", flush=True)
        assert inf.status == 0
        assert rr <= 1.0001e-8
        assert abs(rr - inf.rel_residual) <= 1e-3 * rr + 1e-15
        assert err < 1e-7
