"""Reciprocal-space two-body term on the GPU (set_ewald_summation /
dftd3_set_model_ewald), the direct sum over the vectors (mesh < 0) and the
particle mesh (mesh >= 0), against the oracle's restatement of
fourier/ewald.f90, spme.f90 and decomposition.f90, which
tests/test_oracle_ewald.py pins (the same mesh on both sides).

Tolerances as for the real-space path: FP64 |dE| <= 1e-10 |E|, max |dg| <=
1e-9 Eh/bohr, max |dsigma| <= 1e-9 Eh; mixed precision (the reciprocal sum
stays FP64, the rest is mixed) |dE| <= 1e-8 Eh/atom, max |dg| <= 1e-7.
"""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import cell

pytestmark = pytest.mark.gpu

PBE_BJ = dict(s6=1.0, s9=0.0, alp=14.0, a1=0.4289, s8=0.7875, a2=4.4407)
PBE_ZERO = dict(s6=1.0, s9=0.0, alp=14.0, rs6=1.217, s8=0.722, rs8=1.0)
CUTOFF = (30.0, 60.0, 15.0, 0.0, 0.0)
PBC = [True, True, True]


def _step(msg):
    print(f"[gpu-ewald] {msg}", flush=True)
    sys.stdout.flush()


def cubic(alat, rep):
    frac = np.array([[0.05, 0.10, 0.15], [0.55, 0.35, 0.45]])
    num, xyz = [], []
    for ix in range(rep):
        for iy in range(rep):
            for iz in range(rep):
                shift = alat * np.array([ix, iy, iz], dtype=float)
                num += [6, 8]
                xyz += [alat * frac[0] + shift, alat * frac[1] + shift]
    return np.array(num), np.array(xyz), np.eye(3) * alat * rep


def _param(kind, par):
    from dftd3 import interface as I
    if kind == O.RATIONAL:
        return I.RationalDampingParam(**{k: par[k] for k in ("s6", "s8", "s9", "a1", "a2", "alp")})
    if kind == O.ZERO:
        return I.ZeroDampingParam(**{k: par[k] for k in ("s6", "s8", "s9", "rs6", "rs8", "alp")})
    return I.ModifiedZeroDampingParam(**{k: par[k] for k in ("s6", "s8", "s9", "rs6", "rs8", "alp", "bet")})


def _model(num, pos, lat, pbc, cut=CUTOFF, ghost=None, precision="fp64", **ew):
    from dftd3.interface import DispersionModel
    m = DispersionModel(np.asarray(num), np.asarray(pos), None if lat is None else np.asarray(lat),
                        None if pbc is None else np.asarray(pbc))
    m.set_realspace_cutoff(cut[1], cut[2], cut[0], cut[3], cut[4])
    if ghost is not None:
        m.set_ghost_atoms(np.flatnonzero(ghost))
    if precision != "fp64":
        m.set_precision(precision)
    m.set_ewald_summation(**ew)
    return m


def _check(res, ref, n, precision="fp64", grad=True):
    if precision == "fp64":
        assert abs(res["energy"] - ref["energy"]) <= 1e-10 * abs(ref["energy"]) + 1e-14
        if grad:
            assert np.abs(res["gradient"] - ref["gradient"]).max() <= 1e-9
            assert np.abs(res["virial"] - ref["virial"]).max() <= 1e-9
    else:
        assert abs(res["energy"] - ref["energy"]) <= 1e-8 * n
        if grad:
            assert np.abs(res["gradient"] - ref["gradient"]).max() <= 1e-7


def many_species(seed, n, edge, zs=(1, 3, 5, 6, 7, 8, 9, 11, 14, 15, 16)):
    """A strained cell holding every element of zs (more species than the
    reciprocal-space kernels keep in registers)."""
    _, xyz, lat = cell(seed, n, edge)
    return np.array([zs[k % len(zs)] for k in range(n)]), xyz, lat


CASES = [
    ("cubic-bj-full", lambda: cubic(8.0, 1), O.RATIONAL, PBE_BJ, dict(rank=10000, kcut=10.0, mesh=-1)),
    ("cubic-bj-default", lambda: cubic(8.0, 1), O.RATIONAL, PBE_BJ, dict(mesh=-1)),
    ("cubic-zero", lambda: cubic(8.0, 1), O.ZERO, PBE_ZERO, dict(kcut=12.0, mesh=-1)),
    ("supercell-bj-atm", lambda: cubic(8.0, 2), O.RATIONAL, dict(PBE_BJ, s9=1.0), dict(tolerance=1e-3, mesh=-1)),
    ("triclinic-bj-atm", lambda: cell(7, 10, 11.0), O.RATIONAL, dict(PBE_BJ, s9=1.0), dict(rank=10000, mesh=-1)),
    ("triclinic-zero-atm", lambda: cell(11, 24, 13.0), O.ZERO, dict(PBE_ZERO, s9=1.0),
     dict(rank=6, kcut=14.0, mesh=-1)),
    ("triclinic-bj-96", lambda: cell(3, 96, 19.0), O.RATIONAL, dict(PBE_BJ, s9=1.0), dict(mesh=-1)),
    # particle mesh
    ("mesh-cubic-bj-default", lambda: cubic(8.0, 1), O.RATIONAL, PBE_BJ, dict()),
    ("mesh-supercell-bj-64", lambda: cubic(8.0, 2), O.RATIONAL, PBE_BJ, dict(rank=10000, kcut=10.0, mesh=64)),
    ("mesh-triclinic-bj-atm-24", lambda: cell(7, 10, 11.0), O.RATIONAL, dict(PBE_BJ, s9=1.0),
     dict(rank=4, kcut=6.0, mesh=24)),
    ("mesh-triclinic-zero-atm", lambda: cell(11, 24, 13.0), O.ZERO, dict(PBE_ZERO, s9=1.0), dict(kcut=8.0, mesh=0)),
    # eleven species
    ("species11-bj-atm", lambda: many_species(13, 22, 12.5), O.RATIONAL, dict(PBE_BJ, s9=1.0),
     dict(rank=10000, kcut=8.0, mesh=-1)),
    ("mesh-species11-zero", lambda: many_species(14, 22, 12.5), O.ZERO, PBE_ZERO, dict(kcut=7.0, mesh=24)),
]


@pytest.mark.parametrize("name,make,kind,par,ew", CASES, ids=[c[0] for c in CASES])
def test_ewald_matches_oracle(name, make, kind, par, ew):
    num, xyz, lat = make()
    _step(f"{name}: n={len(num)}")
    ref = O.ewald(num, xyz, lat, PBC, kind, par, CUTOFF, **dict(dict(mesh=0), **ew))
    m = _model(num, xyz, lat, PBC, **ew)
    res = m.get_dispersion(_param(kind, par), grad=True)
    _check(res, ref, len(num))
    e_only = m.get_dispersion(_param(kind, par), grad=False)
    assert abs(e_only["energy"] - ref["energy"]) <= 1e-10 * abs(ref["energy"])


def test_ewald_ghost_atoms():
    num, xyz, lat = cell(5, 16, 12.0)
    ghost = np.zeros(len(num), dtype=np.int32)
    ghost[[1, 4, 9]] = 1
    par = dict(PBE_BJ, s9=1.0)
    _step("ghost atoms")
    for mesh in (-1, 32):
        ref = O.ewald(num, xyz, lat, PBC, O.RATIONAL, par, CUTOFF, ghost=ghost, rank=10000, kcut=8.0, mesh=mesh)
        res = _model(num, xyz, lat, PBC, ghost=ghost, rank=10000, kcut=8.0, mesh=mesh).get_dispersion(
            _param(O.RATIONAL, par), grad=True)
        _check(res, ref, len(num))


def test_ewald_mixed_precision():
    num, xyz, lat = cell(3, 96, 19.0)
    par = dict(PBE_BJ, s9=1.0)
    _step("mixed precision")
    ref = O.ewald(num, xyz, lat, PBC, O.RATIONAL, par, CUTOFF)
    res = _model(num, xyz, lat, PBC, precision="mixed", mesh=-1).get_dispersion(_param(O.RATIONAL, par), grad=True)
    _check(res, ref, len(num), precision="mixed")


def test_lowrank_real_space_when_not_3d_periodic():
    """A low-rank model without 3D periodicity sums in real space with the
    low-rank C6 (disp.f90:107, 142-150)."""
    num, xyz, lat = cell(9, 20, 12.0)
    par = dict(PBE_BJ, s9=1.0)
    for pbc in ([False] * 3, [True, True, False]):
        _step(f"real space low-rank pbc={pbc}")
        ref = O.ewald(num, xyz, lat, pbc, O.RATIONAL, par, CUTOFF, tolerance=1e-3)
        res = _model(num, xyz, lat, pbc, tolerance=1e-3).get_dispersion(_param(O.RATIONAL, par), grad=True)
        _check(res, ref, len(num))


def test_ewald_update_structure_and_converged_sum():
    """python/dftd3/test_interface.py:252-297: Ewald with kcut 10 equals the
    converged (400 bohr) real-space sum within 1e-7 and differs from the
    truncated one; new coordinates through update()."""
    from dftd3.interface import DispersionModel, RationalDampingParam
    num, xyz, lat = cubic(8.0, 1)
    param = RationalDampingParam(method="pbe", atm=False)
    _step("converged sum")
    disp = DispersionModel(num, xyz, lattice=lat, periodic=np.full(3, True))
    disp.set_realspace_cutoff(disp2=400.0, disp3=40.0, cn=40.0)
    converged = disp.get_dispersion(param, grad=False)["energy"]
    disp = DispersionModel(num, xyz, lattice=lat, periodic=np.full(3, True))
    truncated = disp.get_dispersion(param, grad=False)["energy"]
    disp = DispersionModel(num, xyz, lattice=lat, periodic=np.full(3, True))
    disp.set_ewald_summation(kcut=10.0, mesh=-1)
    ewald = disp.get_dispersion(param, grad=False)["energy"]
    mesh_model = DispersionModel(num, xyz, lattice=lat, periodic=np.full(3, True))
    mesh_model.set_ewald_summation(kcut=10.0)
    mesh = mesh_model.get_dispersion(param, grad=False)["energy"]
    assert ewald == pytest.approx(converged, abs=1e-7)
    assert mesh == pytest.approx(ewald, abs=1e-8)
    assert truncated != pytest.approx(converged, abs=1e-6)
    moved = xyz + np.array([[0.3, -0.2, 0.1], [0.0, 0.4, -0.3]])
    disp.update(moved)
    pdict = dict(s6=1.0, s8=0.7875, a1=0.4289, a2=4.4407, s9=0.0, alp=14.0)
    ref = O.ewald(num, moved, lat, PBC, O.RATIONAL, pdict, (40.0, 60.0, 40.0, 0.0, 0.0), grad=False, kcut=10.0)
    assert disp.get_dispersion(param, grad=False)["energy"] == pytest.approx(ref["energy"], rel=1e-10)
    # the particle mesh keeps its kernel table while the cell stays; new
    # coordinates, then a new cell
    mesh_model.update(moved)
    ref = O.ewald(num, moved, lat, PBC, O.RATIONAL, pdict, (40.0, 60.0, 40.0, 0.0, 0.0), kcut=10.0, mesh=0)
    _check(mesh_model.get_dispersion(param, grad=True), ref, len(num))
    lat2 = lat * 1.05
    mesh_model.update(moved, lat2)
    ref = O.ewald(num, moved, lat2, PBC, O.RATIONAL, pdict, (40.0, 60.0, 40.0, 0.0, 0.0), kcut=10.0, mesh=0)
    _check(mesh_model.get_dispersion(param, grad=True), ref, len(num))


def test_ewald_errors():
    from dftd3.interface import DispersionModel, ModifiedZeroDampingParam, RationalDampingParam
    num, xyz, lat = cubic(8.0, 1)
    _step("errors")
    disp = DispersionModel(num, xyz, lattice=lat, periodic=np.full(3, True))
    disp.set_ewald_summation()
    with pytest.raises(RuntimeError, match="does not support Ewald summation"):
        disp.get_dispersion(ModifiedZeroDampingParam(method="pbe"), grad=False)
    with pytest.raises(RuntimeError, match="Hessian is only available for the real space summation"):
        disp.get_hessian(RationalDampingParam(method="pbe"))
    with pytest.raises(RuntimeError, match="Pairwise analysis is only available for the real space summation"):
        disp.get_pairwise_dispersion(RationalDampingParam(method="pbe"))
    # the model stays usable
    res = disp.get_dispersion(RationalDampingParam(method="pbe"), grad=True)
    assert np.isfinite(res["energy"])
