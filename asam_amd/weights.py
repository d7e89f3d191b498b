"""Cut-cell geometry files (`.Weight2`) of ASAM: reading one Cartesian block into the
arrays the stage solve is set up from, and writing them back in the same layout.

The layout is the one the reference's grid generator writes (GRID/OutputWeightBlk_Mod.f90,
WriteWeightBlk :501-760) and its model reads (INIT/ReadWeights_Mod.f90:567-1152), in
Fortran list-directed text:

    Cart                                  grid type
    nb   ! Number of blocks
    #.... Block ib .....                  per block:
    x0 y0 z0                              west/south/bottom coordinates
    nx, then nx lines dx                  (likewise ny/dy and nz/dz)
    n   NrW_FacesYZ, then n x (ix iy iz / area)     overrides of FU = dy dz
    m   NrMP_FacesYZ, then m x (ix iy iz / xS yS zS) cut-face midpoints
    ... the same for FV (= dx dz, ZX faces) and FW (= dx dy, XY faces) ...
    n   NrW_Cells, then n x (ix iy iz / volume)     overrides of VolC = dx dy dz
    NumBoundCell  NrB_Cells  nr_structbound ...     boundary cells (not needed here)

Indices are the block's own (faces ix0..ix1, cells ix0+1..ix1, ...).  Only what the
sound solve needs (spacings, face areas, volumes) is returned; the midpoint and
boundary-cell records are skipped.  Arrays follow the rest of this package: numpy
C-order (nz, ny, nx+1) etc., i.e. the reference's Fortran arrays in memory order.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class BlockGeometry:
    grid_type: str
    x0: float
    y0: float
    z0: float
    dx: np.ndarray
    dy: np.ndarray
    dz: np.ndarray
    FU: np.ndarray   # (nz, ny, nx+1)
    FV: np.ndarray   # (nz, ny+1, nx)
    FW: np.ndarray   # (nz+1, ny, nx)
    VolC: np.ndarray  # (nz, ny, nx)

    @property
    def shape(self):
        return len(self.dz), len(self.dy), len(self.dx)


def _num(tok: str) -> float:
    """A Fortran list-directed real (E or D exponent)."""
    return float(tok.replace("D", "E").replace("d", "e"))


class _Lines:
    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.i = 0

    def next(self) -> list[str]:
        while self.i < len(self.lines):
            s = self.lines[self.i].strip()
            self.i += 1
            if s:
                return s.split()
        raise ValueError("unexpected end of weight file")


def read_weight2(path: str, block: int = 1, offsets=(0, 0, 0)) -> BlockGeometry:
    """Read block `block` of a Cartesian .Weight2 file.  offsets = (ix0, iy0, iz0), the
    block's index origin from the .grid file (0 for a single-block domain)."""
    with open(path) as f:
        return parse_weight2(f.read(), block, offsets)


def parse_weight2(text: str, block: int = 1, offsets=(0, 0, 0)) -> BlockGeometry:
    L = _Lines(text)
    grid_type = L.next()[0].strip("'\"")
    if grid_type != "Cart":
        raise ValueError(f"grid type {grid_type!r}: only Cartesian blocks are supported")
    L.next()  # number of blocks
    while True:
        tok = L.next()
        line = " ".join(tok)
        if "#...." in line and "Block" in line:
            rest = line[line.index("Block") + len("Block"):].split()
            if int(rest[0]) == block:
                break
    x0, y0, z0 = (_num(t) for t in L.next()[:3])

    def spacing():
        n = int(L.next()[0])
        return np.array([_num(L.next()[0]) for _ in range(n)])
    dx, dy, dz = spacing(), spacing(), spacing()
    nx, ny, nz = len(dx), len(dy), len(dz)
    ix0, iy0, iz0 = offsets
    # full Cartesian areas and volumes (ReadWeights_Mod.f90:632-641, 760-786, 798-804, 889-895)
    FU = np.broadcast_to(dz[:, None, None] * dy[None, :, None], (nz, ny, nx + 1)).copy()
    FV = np.broadcast_to(dz[:, None, None] * dx[None, None, :], (nz, ny + 1, nx)).copy()
    FW = np.broadcast_to(dy[None, :, None] * dx[None, None, :], (nz + 1, ny, nx)).copy()
    V = (dz[:, None, None] * dy[None, :, None] * dx[None, None, :]).copy()

    def overrides(arr, di, dj, dk):
        n = int(L.next()[0])
        for _ in range(n):
            ix, iy, iz = (int(t) for t in L.next()[:3])
            val = _num(L.next()[0])
            i, j, k = ix - di, iy - dj, iz - dk
            if 0 <= k < arr.shape[0] and 0 <= j < arr.shape[1] and 0 <= i < arr.shape[2]:
                arr[k, j, i] = val

    def skip_midpoints():
        m = int(L.next()[0])
        for _ in range(2 * m):
            L.next()
    # faces: FU(ix0:ix1, iy0+1:iy1, iz0+1:iz1) -> numpy [iz-iz0-1, iy-iy0-1, ix-ix0]
    overrides(FU, ix0, iy0 + 1, iz0 + 1)
    skip_midpoints()
    overrides(FV, ix0 + 1, iy0, iz0 + 1)
    skip_midpoints()
    overrides(FW, ix0 + 1, iy0 + 1, iz0)
    skip_midpoints()
    overrides(V, ix0 + 1, iy0 + 1, iz0 + 1)
    return BlockGeometry(grid_type, x0, y0, z0, dx, dy, dz, FU, FV, FW, V)


def _f(v: float) -> str:
    return repr(float(v))


def format_weight2(g: BlockGeometry, block: int = 1, nblocks: int = 1) -> str:
    """One Cartesian block in the .Weight2 layout: entries that differ from the full
    Cartesian area/volume are written as overrides (the generator writes every cut or
    closed face and cell, WriteWeightBlk :647-760); no midpoint or boundary-cell records."""
    out = ["Cart", f"{nblocks}  ! Number of blocks"]
    out.append(f"#.... Block {block} ....................................................")
    out.append(f"{_f(g.x0)} {_f(g.y0)} {_f(g.z0)}")
    for d in (g.dx, g.dy, g.dz):
        out.append(str(len(d)))
        out.extend(_f(v) for v in d)
    nz, ny, nx = g.shape
    full = {
        "FU": np.broadcast_to(g.dz[:, None, None] * g.dy[None, :, None], (nz, ny, nx + 1)),
        "FV": np.broadcast_to(g.dz[:, None, None] * g.dx[None, None, :], (nz, ny + 1, nx)),
        "FW": np.broadcast_to(g.dy[None, :, None] * g.dx[None, None, :], (nz + 1, ny, nx)),
        "V": g.dz[:, None, None] * g.dy[None, :, None] * g.dx[None, None, :],
    }
    sections = (("FU", g.FU, (0, 1, 1), "NrW_FacesYZ", "NrMP_FacesYZ"),
                ("FV", g.FV, (1, 0, 1), "NrW_FacesZX", "NrMP_FacesZX"),
                ("FW", g.FW, (1, 1, 0), "NrW_FacesXY", "NrMP_FacesXY"),
                ("V", g.VolC, (1, 1, 1), "NrW_Cells", None))
    for key, arr, (di, dj, dk), name, mpname in sections:
        ks, js, is_ = np.nonzero(arr != full[key])
        out.append(f"{len(ks)}   {name}")
        for k, j, i in zip(ks, js, is_):
            out.append(f"{i + di} {j + dj} {k + dk}")
            out.append(_f(arr[k, j, i]))
        if mpname:
            out.append(f"0   {mpname}")
    out.append("0      NrB_Cells     4   nr_structbound")
    return "\n".join(out) + "\n"


def write_weight2(path: str, g: BlockGeometry) -> None:
    with open(path, "w") as f:
        f.write(format_weight2(g))


def geometry_of(problem) -> BlockGeometry:
    """The geometry part of a SoundProblem as a BlockGeometry."""
    return BlockGeometry("Cart", 0.0, 0.0, 0.0, problem.dxv.copy(), problem.dyv.copy(),
                         problem.dzv.copy(), problem.FU.copy(), problem.FV.copy(),
                         problem.FW.copy(), problem.VolC.copy())
