"""Pairwise-resolved energies on the GPU (dftd3_get_pairwise_dispersion,
d3hip_get_pairwise_dispersion_batch): the reference's golden matrices, the
oracle on molecules and small periodic cells (self images, smooth cutoff
widths, 2D, ghost atoms), both damping kinds and both precisions, and the
batch entry point against the energy path.

Tolerances: FP64 within 1e-10 of the largest element of the oracle's matrix
(the elementwise form of the 1e-10 relative energy bar).  Mixed precision
(FP32 terms, FP64 sums): every row sum, the atomic energy of that row's atom,
within 1e-8 Eh (the mixed-precision energy bar per atom), and every element
within 1e-5 of the largest (FP32 powers (r0/r)^14 of the zero damping give
single large elements ~2e-6 relative error).
"""
import sys

import numpy as np
import pytest
import torch

import oracle as O
from helpers import BJ, ZERO, cell

pytestmark = pytest.mark.gpu


def _step(msg):
    print(f"[gpu-pairwise] {msg}", flush=True)
    sys.stdout.flush()


def _param(kind, par):
    from dftd3.interface import RationalDampingParam, ZeroDampingParam
    if kind == "rational":
        return RationalDampingParam(**{k: par[k] for k in ("s6", "s8", "s9", "a1", "a2", "alp")})
    return ZeroDampingParam(**{k: par[k] for k in ("s6", "s8", "s9", "rs6", "rs8", "alp")})


def _model(num, pos, lat=None, pbc=None, cut=None, precision="fp64", ghost=None):
    from dftd3.interface import DispersionModel
    m = DispersionModel(np.asarray(num), np.asarray(pos), None if lat is None else np.asarray(lat),
                        None if pbc is None else np.asarray(pbc))
    if cut is not None:
        m.set_realspace_cutoff(cut[1], cut[2], cut[0], cut[3], cut[4])
    if precision != "fp64":
        m.set_precision(precision)
    if ghost is not None:
        m.set_ghost_atoms(ghost)
    return m


def _pairwise(m, kind, par):
    r = m.get_pairwise_dispersion(_param(kind, par))
    return r["additive pairwise energy"], r["non-additive pairwise energy"]


def _close(got, ref, rel):
    scale = max(np.abs(ref).max(), 1e-300)
    np.testing.assert_allclose(got, ref, rtol=0, atol=rel * scale)


def _check(got, ref, precision):
    if precision == "fp64":
        _close(got, ref, 1e-10)
    else:
        _close(got, ref, 1e-5)
        np.testing.assert_allclose(got.sum(axis=1), ref.sum(axis=1), rtol=0, atol=1e-8)


@pytest.mark.parametrize("name", ["validation_11-pairwise-d3-bj", "validation_12-pairwise-d3-zero",
                                  "validation_13-pairwise-d3-bjm", "python_pair_resolved"])
def test_pairwise_goldens(pairwise_goldens, name):
    f = pairwise_goldens[name]
    _step(f"golden {name}")
    e2, e3 = _pairwise(_model(f["numbers"], f["positions"], f["lattice"], f["periodic"]), f["kind"], f["param"])
    np.testing.assert_allclose(e2, np.array(f["expected"]["pair2"]), rtol=0, atol=f["atol"])
    np.testing.assert_allclose(e3, np.array(f["expected"]["pair3"]), rtol=0, atol=f["atol"])
    kind = O.RATIONAL if f["kind"] == "rational" else O.ZERO
    r2, r3 = O.pairwise(f["numbers"], f["positions"], f["lattice"], f["periodic"], kind, f["param"], f["cutoff"])
    _close(e2, r2, 1e-10)
    _close(e3, r3, 1e-10)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", ["rational", "zero"])
def test_pairwise_molecule_vs_oracle(kind, precision):
    from dftd3 import synthetic
    _step(f"molecule c0 {kind} {precision}")
    num, pos = synthetic.molecule_c0()
    par = BJ if kind == "rational" else ZERO
    e2, e3 = _pairwise(_model(num, pos, precision=precision), kind, par)
    r2, r3 = O.pairwise(num, pos, None, None, O.RATIONAL if kind == "rational" else O.ZERO, par,
                        O.DEFAULT_CUTOFF)
    _check(e2, r2, precision)
    _check(e3, r3, precision)
    np.testing.assert_array_equal(e2, e2.T)  # exactly symmetric, as the reference's
    np.testing.assert_array_equal(e3, e3.T)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", ["rational", "zero"])
def test_pairwise_small_cells_vs_oracle(kind, precision):
    """Cells smaller than the cutoffs: self-image pairs on the diagonal,
    triples with two images of one atom; smooth cutoff widths; 3D and 2D."""
    par = BJ if kind == "rational" else ZERO
    cut = (40.0, 60.0, 40.0, 3.0, 4.0)
    for seed, n, edge, pbc in [(41, 4, 9.0, [True] * 3), (42, 5, 10.0, [True, True, False]),
                               (43, 9, 11.5, [True] * 3)]:
        _step(f"cell seed={seed} n={n} {kind} {precision}")
        num, x, lat = cell(seed, n, edge)
        e2, e3 = _pairwise(_model(num, x, lat, pbc, cut=cut, precision=precision), kind, par)
        r2, r3 = O.pairwise(num, x, lat, pbc, O.RATIONAL if kind == "rational" else O.ZERO, par, cut)
        _check(e2, r2, precision)
        _check(e3, r3, precision)
        np.testing.assert_array_equal(e2, e2.T)
        np.testing.assert_array_equal(e3, e3.T)


def test_pairwise_ghost_atoms():
    from dftd3 import synthetic
    _step("ghost atoms")
    num, pos = synthetic.molecule_c0(seed=5, n=40)
    ghost = np.array([2, 7, 31])
    mask = np.zeros(len(num), dtype=np.int32)
    mask[ghost] = 1
    e2, e3 = _pairwise(_model(num, pos, ghost=ghost), "rational", BJ)
    r2, r3 = O.pairwise(num, pos, None, None, O.RATIONAL, BJ, O.DEFAULT_CUTOFF, ghost=mask)
    _close(e2, r2, 1e-10)
    _close(e3, r3, 1e-10)
    assert not e2[ghost].any() and not e3[:, ghost].any()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pairwise_batch_entry_point(precision):
    """Molecules and cells of both damping kinds, with and without the
    three-body term, in one call: every block against the oracle, and the
    block sums against the batch energy entry point."""
    from dftd3 import synthetic, torch_api
    systems = []
    for k in range(4):
        num, pos = synthetic.molecule_c0(seed=60 + k, n=9 + 17 * k)
        kind = "rational" if k % 2 == 0 else "zero"
        par = dict(BJ if kind == "rational" else ZERO, s9=0.0 if k == 2 else 1.0)
        systems.append(dict(numbers=num, positions=pos, kind=kind, param=par))
    for k in range(2):
        num, pos, lat = cell(70 + k, 6 + 3 * k, 10.5)
        systems.append(dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, param=BJ))
    spec, pos = torch_api.batch_spec(systems)
    pos = torch.as_tensor(pos, dtype=torch.float64, device="cuda").contiguous()
    _step(f"batch of {len(systems)} {precision}")
    out = torch_api.run_pairwise(spec, pos, precision=precision)
    ref = torch_api.run_batch(spec, pos, precision=precision, grad=False)
    torch.cuda.synchronize()
    b2, b3 = torch_api.blocks(out, "pair2"), torch_api.blocks(out, "pair3")
    for k, s in enumerate(systems):
        kind = O.RATIONAL if s.get("kind", "rational") == "rational" else O.ZERO
        r2, r3 = O.pairwise(s["numbers"], s["positions"], s.get("lattice"), s.get("periodic"), kind, s["param"],
                            O.DEFAULT_CUTOFF)
        _check(b2[k].cpu().numpy(), r2, precision)
        _check(b3[k].cpu().numpy(), r3, precision)
        if s["param"]["s9"] == 0.0:
            assert not b3[k].any()
        e = ref["energy"][k].item()
        assert abs(b2[k].sum().item() + b3[k].sum().item() - e) <= (1e-11 if precision == "fp64" else 1e-7) * abs(e)
