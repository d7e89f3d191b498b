"""Input generator checks: exact cut-cell geometry against quadrature, the counter-based
RNG's determinism and moments."""
import numpy as np
import pytest
from scipy.integrate import quad

from asam_amd import workload as W


def test_agnesi_cut_areas_match_quadrature():
    g = W.Grid(12, 2, 8, 500.0, 500.0, 200.0)
    H, a, xc = 1000.0, 2500.0, 3000.0
    FU, FV, FW, V = W.cut_geometry(g, H, a, xc, vf_min=0.0)
    h = lambda x: H / (1.0 + ((x - xc) / a) ** 2)  # noqa: E731
    for i in range(g.nx):
        for k in range(g.nz):
            zlo, zhi = k * g.dz, (k + 1) * g.dz
            f = lambda x: zhi - min(max(h(x), zlo), zhi)  # noqa: E731
            area, _ = quad(f, i * g.dx, (i + 1) * g.dx, limit=200, epsabs=1e-9)
            vf = area / (g.dx * g.dz)
            expect = 0.0 if vf < 1e-9 else area * g.dy
            assert abs(V[k, 0, i] - expect) <= 1e-7 * g.dx * g.dy * g.dz, (i, k)
            assert abs(FV[k, 1, i] - (0.0 if vf < 1e-9 else area)) <= 1e-7 * g.dx * g.dz
    # x-faces: open height at the face position
    for i in range(g.nx + 1):
        for k in range(g.nz):
            zlo, zhi = k * g.dz, (k + 1) * g.dz
            ex = (zhi - min(max(h(i * g.dx), zlo), zhi)) * g.dy
            assert abs(FU[k, 0, i] - (ex if ex >= 1e-9 * g.dz * g.dy else 0.0)) < 1e-6
    # z-faces: open length {h < z}
    for i in range(g.nx):
        for k in range(g.nz + 1):
            z = k * g.dz
            xs = np.linspace(i * g.dx, (i + 1) * g.dx, 200001)
            meas = np.mean(h(xs) < z) * g.dx
            assert abs(FW[k, 0, i] / g.dy - meas) < 0.05, (i, k)


def test_volume_fractions_and_clamp():
    p = W.config(2, np1=1, scale=0.25)
    vf = p.VolC / (500.0 * 500.0 * 200.0)
    assert vf.min() >= 0.0 and vf.max() <= 1.0 + 1e-12
    pos = vf[vf > 0]
    assert pos.min() >= 1e-3 - 1e-15
    cut = ((vf > 0) & (vf < 1 - 1e-12)).mean()
    assert 0.005 < cut < 0.05
    # dead cells carry b = 0
    assert np.all(p.b[(p.VolC.reshape(-1) == 0)] == 0)


def test_rng_is_counter_based_and_standard_normal():
    a = W.normal(100000, 42)
    b = W.normal(100000, 42)
    c = W.normal(100000, 43)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert abs(a.mean()) < 0.02 and abs(a.std() - 1.0) < 0.02
    # prefix property of the counter-based stream
    assert np.array_equal(W.uniform01(10, 7), W.uniform01(20, 7)[:10])


def test_splitmix64_known_values():
    # SplitMix64 (Vigna) from state 0: the first outputs of the reference generator
    z = W.splitmix64(np.arange(3, dtype=np.uint64), 0)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_hydrostatic_background():
    p = W.config(1, np1=1, scale=0.25)
    # far from the bubble theta = 300 K; p = rho Rd T
    th = p.rhotheta / p.rho
    assert abs(th[-1, 0, 0] - 300.0) < 1e-6
    prs = p.rho * W.RD * p.T
    z = (np.arange(p.grid.nz) + 0.5) * p.grid.dz
    pi = 1 - W.GRAV * z / (W.CPD * 300.0)
    assert np.allclose(prs[:, 0, 0], W.P0 * pi ** (W.CPD / W.RD), rtol=1e-12)
