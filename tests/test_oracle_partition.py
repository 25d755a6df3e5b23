"""The oracle's work partition (partition.f90:57-148, api.f90:459-480).

Pinned by the reference's own tests: the parts sum to the whole calculation
(python/dftd3/test_interface.py:440-466 at 1e-12 for energy, gradient, virial
and Hessian; test/unit/test_partition.f90:447-624 for the dispersion, the
Hessian and the Ewald sum), an explicit serial partition is the unpartitioned
calculation (test_partition.f90:112-166), and the invalid partitions of
test/api/api-test.c:992-1035 are rejected.  The ownership rule itself is
pinned per part independently of the loops that apply it: a part's two-body
energy is the sum of the unpartitioned pairwise matrix (validated against the
CLI goldens) over the pairs the part owns, and the single triple of three
atoms belongs to the part of its pair (3, 2).
"""
import numpy as np
import pytest

import oracle as O
from helpers import BJ, ZERO, cell

PBC = [True, True, True]
NPARTS = 3


def _water_dimer():
    num = np.array([8, 1, 1, 8, 1, 1])
    xyz = np.array([[-1.551007, -0.114520, 0.000000], [-1.934259, 0.762503, 0.000000],
                    [-0.599677, 0.040712, 0.000000], [1.350625, 0.111469, 0.000000],
                    [1.680398, -0.373741, -0.758561], [1.680398, -0.373741, 0.758561]]) / 0.52917721067
    return num, xyz


def _sum_parts(fn, keys, nparts=NPARTS):
    tot = None
    for part in range(nparts):
        with O.partition(part, nparts):
            res = fn()
        if tot is None:
            tot = {k: np.array(res[k], dtype=float) for k in keys}
        else:
            for k in keys:
                tot[k] = tot[k] + np.array(res[k], dtype=float)
    return tot


def test_invalid_partitions_are_rejected():
    for part, nparts in ((-1, 3), (3, 3), (0, 0), (1, 1)):
        assert not O.set_partition(part, nparts)
    assert O.set_partition(0, 1)


def test_serial_partition_is_the_whole_calculation():
    num, xyz = _water_dimer()
    ref = O.dispersion(num, xyz, kind=O.RATIONAL, param=BJ)
    with O.partition(0, 1):
        res = O.dispersion(num, xyz, kind=O.RATIONAL, param=BJ)
    # equal up to the merge order of the oracle's per-thread sums
    assert res["energy"] == pytest.approx(ref["energy"], rel=1e-14)
    assert np.abs(res["gradient"] - ref["gradient"]).max() <= 1e-16
    assert np.abs(res["virial"] - ref["virial"]).max() <= 1e-15


@pytest.mark.parametrize("kind,par", [(O.RATIONAL, BJ), (O.ZERO, ZERO)], ids=["bj-atm", "zero-atm"])
def test_parts_sum_to_whole_molecule(kind, par):
    num, xyz = _water_dimer()
    ref = O.dispersion(num, xyz, kind=kind, param=par)
    tot = _sum_parts(lambda: O.dispersion(num, xyz, kind=kind, param=par), ("energy", "gradient", "virial"))
    assert tot["energy"] == pytest.approx(ref["energy"], abs=1e-12)
    assert np.abs(tot["gradient"] - ref["gradient"]).max() <= 1e-12
    assert np.abs(tot["virial"] - ref["virial"]).max() <= 1e-12


def test_parts_sum_to_whole_periodic_cell():
    """Self images, several images per pair and triples with two images of
    one atom, 4 parts."""
    num, xyz, lat = cell(3, 10, 11.0)
    cut = (20.0, 30.0, 15.0, 0.0, 0.0)
    ref = O.dispersion(num, xyz, lat, PBC, O.RATIONAL, BJ, cut)
    tot = _sum_parts(lambda: O.dispersion(num, xyz, lat, PBC, O.RATIONAL, BJ, cut),
                     ("energy", "gradient", "virial"), nparts=4)
    assert tot["energy"] == pytest.approx(ref["energy"], rel=1e-13, abs=1e-14)
    assert np.abs(tot["gradient"] - ref["gradient"]).max() <= 1e-12
    assert np.abs(tot["virial"] - ref["virial"]).max() <= 1e-12


def test_parts_sum_to_whole_hessian():
    num, xyz = _water_dimer()
    ref = O.hessian(num, xyz, kind=O.RATIONAL, param=BJ)
    tot = _sum_parts(lambda: O.hessian(num, xyz, kind=O.RATIONAL, param=BJ), ("energy", "hessian"))
    assert tot["energy"] == pytest.approx(ref["energy"], abs=1e-12)
    assert np.abs(tot["hessian"] - ref["hessian"]).max() <= 1e-12


@pytest.mark.parametrize("mesh", [-1, 0], ids=["direct", "mesh"])
def test_parts_sum_to_whole_ewald(mesh):
    num = np.array([6, 8, 6, 8])
    xyz = np.array([[0.4, 0.8, 1.2], [4.4, 2.8, 3.6], [4.4, 4.8, 5.2], [0.2, 6.8, 7.6]])
    lat = np.eye(3) * 8.0
    par = dict(BJ, s9=0.0)
    cut = (40.0, 60.0, 40.0, 0.0, 0.0)
    kw = dict(rank=10000, kcut=6.0, mesh=mesh)
    ref = O.ewald(num, xyz, lat, PBC, O.RATIONAL, par, cut, **kw)
    tot = _sum_parts(lambda: O.ewald(num, xyz, lat, PBC, O.RATIONAL, par, cut, **kw),
                     ("energy", "gradient", "virial"), nparts=5)
    assert tot["energy"] == pytest.approx(ref["energy"], rel=1e-12)
    assert np.abs(tot["gradient"] - ref["gradient"]).max() <= 1e-12
    assert np.abs(tot["virial"] - ref["virial"]).max() <= 1e-11


def _owns(i, j, part, nparts):
    """partition.f90:127-148 for 1-based i >= j."""
    return ((i - 1) * i // 2 + (j - 1)) % nparts == part


def test_two_body_part_is_its_pairs_of_the_pairwise_matrix():
    num, xyz = _water_dimer()
    par = dict(BJ, s9=0.0)
    e2, _ = O.pairwise(num, xyz, kind=O.RATIONAL, param=par)
    n = len(num)
    for part in range(NPARTS):
        want = 0.0
        for i in range(1, n + 1):
            for j in range(1, i + 1):
                if _owns(i, j, part, NPARTS):
                    want += e2[i - 1, j - 1] * (1.0 if i == j else 2.0)
        with O.partition(part, NPARTS):
            got = O.dispersion(num, xyz, kind=O.RATIONAL, param=par, grad=False)["energy"]
        assert got == pytest.approx(want, rel=1e-12, abs=1e-16)


def test_three_body_term_of_three_atoms_belongs_to_pair_3_2():
    num = np.array([6, 7, 8])
    xyz = np.array([[0.0, 0.0, 0.0], [2.6, 0.0, 0.0], [1.1, 2.3, 0.4]])
    e_atm = (O.dispersion(num, xyz, kind=O.RATIONAL, param=BJ, grad=False)["energy"]
             - O.dispersion(num, xyz, kind=O.RATIONAL, param=dict(BJ, s9=0.0), grad=False)["energy"])
    assert abs(e_atm) > 1e-9
    for part in range(NPARTS):
        with O.partition(part, NPARTS):
            with_atm = O.dispersion(num, xyz, kind=O.RATIONAL, param=BJ, grad=False)["energy"]
            without = O.dispersion(num, xyz, kind=O.RATIONAL, param=dict(BJ, s9=0.0), grad=False)["energy"]
        want = e_atm if _owns(3, 2, part, NPARTS) else 0.0
        assert with_atm - without == pytest.approx(want, rel=1e-12, abs=1e-18)
