"""Seeded synthetic inputs of the benchmark configurations.

All structures are built by random sequential packing: atoms are drawn
uniformly in the target region and rejected when closer than ``dmin``
(1.0 Angstrom) to an atom already placed (minimum image for cells).
Coordinates are returned in bohr.

  molecule_c0      200 atoms, H/C/N/O 0.5/0.3/0.1/0.1, cube at 0.1 atoms/A^3
  periodic_c1      2,000 atoms, H/C/N/O/Si/S, 28 A cubic cell, +-5 % strain
  ellipsoid_c2     20,000 atoms, H .50 C .32 N .09 O .08 S .01, ellipsoid 1:1:2
  batch_c3         4,096 molecules of 10-300 atoms (log-uniform), 9 elements
  cells_c4         256 cells of 64-512 atoms, H/C/N/O/Si, 10-20 A edges
"""
import math
from typing import List, Tuple

import numpy as np

# Angstrom -> bohr: Bohr radius hbar / (m_e c alpha), CODATA 2018.
AATOAU = 1.0 / (6.62607015e-34 / (2.0 * math.pi) / (9.1093837015e-31 * 299792458.0 * 7.2973525693e-3) * 1.0e10)

PBE0_BJ = dict(s6=1.0, s8=1.2177, a1=0.4145, a2=4.8593, s9=1.0, alp=14.0)   # param.f90:364-366, atm
B3LYP_ZERO = dict(s6=1.0, s8=1.703, rs6=1.261, rs8=1.0, s9=1.0, alp=14.0)  # param.f90:857, atm


class _Grid:
    """Cell hash for the minimum-distance test (Cartesian or fractional)."""

    def __init__(self, cell: float):
        self.cell = cell
        self.cells = {}

    def key(self, x):
        return tuple(int(math.floor(c / self.cell)) for c in x)

    def neighbours(self, x):
        k = self.key(x)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    yield from self.cells.get((k[0] + dx, k[1] + dy, k[2] + dz), ())

    def add(self, x, idx):
        self.cells.setdefault(self.key(x), []).append(idx)


def _pack_region(rng, n, sample, dmin=1.0, max_tries=200000):
    grid = _Grid(dmin)
    pts = []
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > max_tries * max(1, n // 1000 + 1):
            raise RuntimeError("packing failed")
        x = sample()
        ok = True
        for j in grid.neighbours(x):
            d = x - pts[j]
            if d @ d < dmin * dmin:
                ok = False
                break
        if ok:
            grid.add(x, len(pts))
            pts.append(x)
    return np.array(pts)


def _elements(rng, n, symbols_z, probs):
    p = np.asarray(probs, dtype=float)
    return rng.choice(np.asarray(symbols_z), size=n, p=p / p.sum()).astype(np.int32)


def molecule_c0(seed=0, n=200, density=0.1):
    """Config 0: 200-atom molecule in a cube (H/C/N/O .5/.3/.1/.1)."""
    rng = np.random.default_rng(seed)
    edge = (n / density) ** (1.0 / 3.0)
    xyz = _pack_region(rng, n, lambda: rng.uniform(0.0, edge, 3))
    num = _elements(rng, n, [1, 6, 7, 8], [0.5, 0.3, 0.1, 0.1])
    return num, xyz * AATOAU


def _pack_cell(rng, n, lattice_aa, dmin=1.0):
    """Random sequential packing in a periodic cell (lattice rows, Angstrom)."""
    lat = np.asarray(lattice_aa, dtype=float)
    pts = []
    inv = np.linalg.inv(lat)
    # fractional grid spacing so that one cell spans at least dmin in space
    heights = 1.0 / np.linalg.norm(inv, axis=0)
    ncell = np.maximum(1, np.floor(heights / dmin).astype(int))
    cells = {}

    def key(f):
        return tuple((np.floor(f * ncell).astype(int) % ncell).tolist())

    while len(pts) < n:
        f = rng.uniform(0.0, 1.0, 3)
        x = f @ lat
        k = key(f)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    kk = ((k[0] + dx) % ncell[0], (k[1] + dy) % ncell[1], (k[2] + dz) % ncell[2])
                    for j in cells.get(kk, ()):
                        df = f - pts[j][0]
                        df -= np.round(df)
                        d = df @ lat
                        if d @ d < dmin * dmin:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            cells.setdefault(k, []).append(len(pts))
            pts.append((f, x))
    return np.array([p[1] for p in pts])


def periodic_c1(seed=1, n=2000, edge=28.0, strain=0.05):
    """Config 1: 2,000-atom triclinic cell (cubic 28 A, +-5 % off-diagonal)."""
    rng = np.random.default_rng(seed)
    lat = np.eye(3) * edge
    off = rng.uniform(-strain, strain, (3, 3))
    np.fill_diagonal(off, 0.0)
    lat = lat @ (np.eye(3) + off)
    xyz = _pack_cell(rng, n, lat)
    num = _elements(rng, n, [1, 6, 7, 8, 14, 16], [0.45, 0.30, 0.08, 0.08, 0.05, 0.04])
    return num, xyz * AATOAU, lat * AATOAU


def ellipsoid_c2(seed=2, n=20000, density=0.1, ratio=(1.0, 1.0, 2.0)):
    """Config 2: protein-like 20,000-atom molecule in an ellipsoid (1:1:2)."""
    rng = np.random.default_rng(seed)
    r = np.asarray(ratio, dtype=float)
    vol = n / density
    a = (vol / (4.0 / 3.0 * math.pi * r.prod())) ** (1.0 / 3.0)
    axes = a * r

    def sample():
        while True:
            u = rng.uniform(-1.0, 1.0, 3)
            if u @ u <= 1.0:
                return u * axes

    xyz = _pack_region(rng, n, sample)
    num = _elements(rng, n, [1, 6, 7, 8, 16], [0.50, 0.32, 0.09, 0.08, 0.01])
    return num, xyz * AATOAU


def batch_c3(seed=3, nmol=4096, nmin=10, nmax=300, density=0.1) -> List[Tuple[np.ndarray, np.ndarray, dict, int]]:
    """Config 3: molecules of log-uniform size; BJ (PBE0) and zero (B3LYP)
    damping alternate.  Returns (numbers, positions, param, kind) tuples with
    kind 0 = rational, 1 = zero."""
    rng = np.random.default_rng(seed)
    zs = [1, 5, 6, 7, 8, 9, 15, 16, 17]
    w = [0.40, 0.02, 0.30, 0.07, 0.09, 0.03, 0.03, 0.03, 0.03]
    sizes = np.exp(rng.uniform(math.log(nmin), math.log(nmax + 1), nmol)).astype(int)
    sizes = np.clip(sizes, nmin, nmax)
    out = []
    for m, n in enumerate(sizes):
        edge = (n / density) ** (1.0 / 3.0)
        xyz = _pack_region(rng, int(n), lambda: rng.uniform(0.0, edge, 3))
        num = _elements(rng, int(n), zs, w)
        if m % 2 == 0:
            out.append((num, xyz * AATOAU, dict(PBE0_BJ), 0))
        else:
            out.append((num, xyz * AATOAU, dict(B3LYP_ZERO), 1))
    return out


def cells_c4(seed=4, ncell=256, nmin=64, nmax=512, emin=10.0, emax=20.0):
    """Config 4: periodic cells, cubic to triclinic, H/C/N/O/Si."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ncell):
        edge = rng.uniform(emin, emax)
        lat = np.eye(3) * edge
        lat = lat @ (np.eye(3) + rng.uniform(-0.15, 0.15, (3, 3)) * rng.uniform(0, 1))
        vol = abs(np.linalg.det(lat))
        n = int(np.clip(round(vol * rng.uniform(0.05, 0.1)), nmin, nmax))
        n = min(n, int(vol * 0.12))
        n = max(n, 1)
        xyz = _pack_cell(rng, n, lat)
        num = _elements(rng, n, [1, 6, 7, 8, 14], [0.40, 0.30, 0.10, 0.12, 0.08])
        out.append((num, xyz * AATOAU, lat * AATOAU))
    return out
