"""Work partition on the GPU (DispersionModel.set_work_partition /
dftd3_set_model_work_partition, api.f90:459-480, partition.f90:57-148).

Every part is compared with the same part of the oracle (test_oracle_partition
pins the oracle's partition) at the tolerances of the whole calculation: FP64
|dE| <= 1e-10 |E| + 1e-14, max |dg| <= 1e-9 Eh/bohr, max |dsigma| <= 1e-9 Eh;
mixed |dE| <= 1e-8 Eh/atom, max |dg| <= 1e-7 Eh/bohr.  The parts of the GPU
sum to its unpartitioned result within 1e-12 (python/dftd3/test_interface.py
:440-466).
"""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import BJ, ZERO, cell

pytestmark = pytest.mark.gpu

PBC = [True, True, True]


def _step(msg):
    print(f"[gpu-partition] {msg}", flush=True)
    sys.stdout.flush()


def _water_dimer():
    num = np.array([8, 1, 1, 8, 1, 1])
    xyz = np.array([[-1.551007, -0.114520, 0.000000], [-1.934259, 0.762503, 0.000000],
                    [-0.599677, 0.040712, 0.000000], [1.350625, 0.111469, 0.000000],
                    [1.680398, -0.373741, -0.758561], [1.680398, -0.373741, 0.758561]]) / 0.52917721067
    return num, xyz


def _param(kind, par):
    from dftd3 import interface as I
    if kind == O.RATIONAL:
        return I.RationalDampingParam(**{k: par[k] for k in ("s6", "s8", "s9", "a1", "a2", "alp")})
    return I.ZeroDampingParam(**{k: par[k] for k in ("s6", "s8", "s9", "rs6", "rs8", "alp")})


def _model(num, xyz, lat=None, pbc=None, cut=O.DEFAULT_CUTOFF, part=0, nparts=1, precision="fp64", ewald=None):
    from dftd3.interface import DispersionModel
    m = DispersionModel(np.asarray(num), np.asarray(xyz), None if lat is None else np.asarray(lat),
                        None if pbc is None else np.asarray(pbc))
    m.set_realspace_cutoff(cut[1], cut[2], cut[0], cut[3], cut[4])
    if precision != "fp64":
        m.set_precision(precision)
    if ewald is not None:
        m.set_ewald_summation(**ewald)
    if nparts > 1:
        m.set_work_partition(part, nparts)
    return m


def _check(res, ref, n, precision="fp64"):
    if precision == "fp64":
        assert abs(res["energy"] - ref["energy"]) <= 1e-10 * abs(ref["energy"]) + 1e-14
        assert np.abs(res["gradient"] - ref["gradient"]).max() <= 1e-9
        assert np.abs(res["virial"] - ref["virial"]).max() <= 1e-9
    else:
        assert abs(res["energy"] - ref["energy"]) <= 1e-8 * n
        assert np.abs(res["gradient"] - ref["gradient"]).max() <= 1e-7


CASES = [
    ("water-dimer-bj-atm", lambda: (*_water_dimer(), None), None, O.RATIONAL, BJ, O.DEFAULT_CUTOFF, 3),
    ("water-dimer-zero-atm", lambda: (*_water_dimer(), None), None, O.ZERO, ZERO, O.DEFAULT_CUTOFF, 4),
    ("cell-10-bj-atm", lambda: cell(3, 10, 11.0), PBC, O.RATIONAL, BJ, (20.0, 30.0, 15.0, 0.0, 0.0), 4),
    # several 64-slot tiles per three-body list (diagonal and off-diagonal tile pairs)
    ("cell-60-bj-atm", lambda: cell(5, 60, 16.0), PBC, O.RATIONAL, BJ, (25.0, 30.0, 20.0, 0.0, 0.0), 3),
    ("cell-60-zero-atm-smooth", lambda: cell(5, 60, 16.0), PBC, O.ZERO, ZERO, (25.0, 30.0, 20.0, 2.0, 3.0), 5),
]


@pytest.mark.parametrize("name,make,pbc,kind,par,cut,nparts", CASES, ids=[c[0] for c in CASES])
def test_parts_match_oracle_and_sum_to_whole(name, make, pbc, kind, par, cut, nparts):
    num, xyz, lat = make()
    _step(f"{name}: n={len(num)} nparts={nparts}")
    whole = _model(num, xyz, lat, pbc, cut).get_dispersion(_param(kind, par), grad=True)
    tot_e, tot_g, tot_s = 0.0, np.zeros_like(whole["gradient"]), np.zeros((3, 3))
    for part in range(nparts):
        with O.partition(part, nparts):
            ref = O.dispersion(num, xyz, lat, pbc, kind, par, cut)
        res = _model(num, xyz, lat, pbc, cut, part, nparts).get_dispersion(_param(kind, par), grad=True)
        _check(res, ref, len(num))
        e_only = _model(num, xyz, lat, pbc, cut, part, nparts).get_dispersion(_param(kind, par), grad=False)
        assert abs(e_only["energy"] - ref["energy"]) <= 1e-10 * abs(ref["energy"]) + 1e-14
        tot_e += res["energy"]
        tot_g += res["gradient"]
        tot_s += res["virial"]
    assert tot_e == pytest.approx(whole["energy"], abs=1e-12)
    assert np.abs(tot_g - whole["gradient"]).max() <= 1e-12
    assert np.abs(tot_s - whole["virial"]).max() <= 1e-12


def test_mixed_precision_parts():
    num, xyz, lat = cell(5, 60, 16.0)
    cut = (25.0, 30.0, 20.0, 0.0, 0.0)
    _step("mixed precision parts")
    for part in range(3):
        with O.partition(part, 3):
            ref = O.dispersion(num, xyz, lat, PBC, O.RATIONAL, BJ, cut)
        res = _model(num, xyz, lat, PBC, cut, part, 3, precision="mixed").get_dispersion(
            _param(O.RATIONAL, BJ), grad=True)
        _check(res, ref, len(num), precision="mixed")


@pytest.mark.parametrize("periodic", [False, True], ids=["molecule", "cell"])
def test_hessian_parts(periodic):
    if periodic:
        num, xyz, lat = cell(7, 8, 10.0)
        pbc, cut = PBC, (15.0, 20.0, 12.0, 0.0, 0.0)
    else:
        (num, xyz), lat, pbc, cut = _water_dimer(), None, None, O.DEFAULT_CUTOFF
    _step(f"hessian parts periodic={periodic}")
    whole = _model(num, xyz, lat, pbc, cut).get_hessian(_param(O.RATIONAL, BJ))["hessian"]
    tot = np.zeros_like(whole)
    for part in range(3):
        with O.partition(part, 3):
            ref = O.hessian(num, xyz, lat, pbc, O.RATIONAL, BJ, cut)
        res = _model(num, xyz, lat, pbc, cut, part, 3).get_hessian(_param(O.RATIONAL, BJ))
        assert abs(res["energy"] - ref["energy"]) <= 1e-10 * abs(ref["energy"]) + 1e-14
        assert np.abs(res["hessian"] - ref["hessian"]).max() <= 1e-9
        tot += res["hessian"]
    assert np.abs(tot - whole).max() <= 1e-12


@pytest.mark.parametrize("mesh", [-1, 0, 32], ids=["direct", "mesh-default", "mesh-32"])
def test_ewald_parts(mesh):
    num, xyz, lat = cell(9, 12, 11.0)
    par = dict(BJ, s9=1.0)
    cut = (30.0, 60.0, 15.0, 0.0, 0.0)
    ew = dict(rank=10000, kcut=8.0, mesh=mesh)
    _step(f"ewald parts mesh={mesh}")
    whole = _model(num, xyz, lat, PBC, cut, ewald=ew).get_dispersion(_param(O.RATIONAL, par), grad=True)
    tot_e, tot_g = 0.0, np.zeros_like(whole["gradient"])
    for part in range(5):
        with O.partition(part, 5):
            ref = O.ewald(num, xyz, lat, PBC, O.RATIONAL, par, cut, **ew)
        res = _model(num, xyz, lat, PBC, cut, part, 5, ewald=ew).get_dispersion(_param(O.RATIONAL, par), grad=True)
        _check(res, ref, len(num))
        tot_e += res["energy"]
        tot_g += res["gradient"]
    assert tot_e == pytest.approx(whole["energy"], abs=1e-12)
    assert np.abs(tot_g - whole["gradient"]).max() <= 1e-12


def test_invalid_partition_and_pairwise():
    """api-test.c:992-1035: invalid partitions are errors and leave the model
    serial; the pairwise analysis is not partitioned (api.f90 passes no
    partition to it)."""
    num, xyz = _water_dimer()
    _step("invalid partition")
    m = _model(num, xyz)
    whole = m.get_dispersion(_param(O.RATIONAL, BJ), grad=False)["energy"]
    for part, nparts in ((-1, 3), (3, 3), (0, 0), (1, 1)):
        with pytest.raises(RuntimeError, match="Invalid dispersion work partition"):
            m.set_work_partition(part, nparts)
    assert m.get_dispersion(_param(O.RATIONAL, BJ), grad=False)["energy"] == pytest.approx(whole, rel=1e-14)
    parted = _model(num, xyz, part=1, nparts=3)
    serial = _model(num, xyz)
    pw = parted.get_pairwise_dispersion(_param(O.RATIONAL, BJ))
    ref = serial.get_pairwise_dispersion(_param(O.RATIONAL, BJ))
    for k in ("additive pairwise energy", "non-additive pairwise energy"):
        assert np.abs(pw[k] - ref[k]).max() <= 1e-15
