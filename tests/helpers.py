"""Shared inputs of the test modules: damping parameters and random cells."""
import numpy as np

BJ = dict(s6=1.0, s8=1.2177, a1=0.4145, a2=4.8593, s9=1.0, alp=14.0)  # PBE0-D3(BJ)-ATM
ZERO = dict(s6=1.0, s8=1.703, rs6=1.261, rs8=1.0, s9=1.0, alp=14.0)   # B3LYP-D3(0)-ATM


def cell(seed, n, edge, min_dist=1.9):
    """n atoms (H/C/N/O/Si/S) at random fractional positions of a strained
    cubic cell of the given edge (bohr), no two closer than min_dist."""
    rng = np.random.default_rng(seed)
    lat = np.eye(3) * edge + rng.uniform(-0.08, 0.08, (3, 3)) * edge
    frac = []
    while len(frac) < n:
        f = rng.uniform(0, 1, 3)
        ok = True
        for g in frac:
            d = f - g
            d -= np.round(d)
            if np.linalg.norm(d @ lat) < min_dist:
                ok = False
                break
        if ok:
            frac.append(f)
    num = rng.choice([1, 6, 7, 8, 14, 16], size=n)
    return num, np.array(frac) @ lat, lat
