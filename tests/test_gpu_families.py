"""Damping families on the GPU (modified rational, modified zero, optimized
power, CSO, Z) through dftd3.interface and the batch entry point, against
the reference's validation goldens and the oracle.

Tolerances as for the rational and zero damping: FP64 |dE| <= 1e-10 |E|,
max |dg| <= 1e-9 Eh/bohr, max |dsigma| <= 1e-9 Eh; mixed precision
|dE| <= 1e-8 Eh/atom, max |dg| <= 1e-7 Eh/bohr; pair matrices as in
test_gpu_pairwise.py.
"""
import sys

import numpy as np
import pytest
import torch

import oracle as O
from helpers import cell

pytestmark = pytest.mark.gpu

KINDS = {"rational": O.RATIONAL, "zero": O.ZERO, "mzero": O.MZERO, "optimizedpower": O.OPTIMIZEDPOWER,
         "cso": O.CSO, "z": O.ZDAMP}
PARAMS = {
    "mzero": dict(s6=1.0, s8=1.841686, rs6=1.279637, rs8=1.0, bet=0.014370, s9=1.0, alp=14.0),
    "optimizedpower": dict(s6=1.0, s8=0.51581, a1=0.575, a2=3.0, bet=8.0, s9=1.0, alp=14.0),
    "cso": dict(s6=1.0, a1=0.2, a2=2.5, a3=0.0, a4=6.25, s9=1.0, alp=14.0),
    "z": dict(s6=1.0, s8=1.0, a1=200770.0, s9=1.0, alp=14.0),
}


def _step(msg):
    print(f"[gpu-families] {msg}", flush=True)
    sys.stdout.flush()


def _param(kind, par):
    from dftd3 import interface as I
    keys = dict(rational=("s6", "s8", "s9", "a1", "a2", "alp"), zero=("s6", "s8", "s9", "rs6", "rs8", "alp"),
                mzero=("s6", "s8", "s9", "rs6", "rs8", "alp", "bet"),
                optimizedpower=("s6", "s8", "s9", "a1", "a2", "alp", "bet"),
                cso=("s6", "s9", "a1", "a2", "a3", "a4", "alp"), z=("s6", "s8", "s9", "a1", "alp"))
    cls = dict(rational=I.RationalDampingParam, zero=I.ZeroDampingParam, mzero=I.ModifiedZeroDampingParam,
               optimizedpower=I.OptimizedPowerDampingParam, cso=I.CSODampingParam, z=I.ZDampingParam)[kind]
    return cls(**{k: par[k] for k in keys[kind] if k in par})


def _model(num, pos, lat=None, pbc=None, cut=None, precision="fp64"):
    from dftd3.interface import DispersionModel
    m = DispersionModel(np.asarray(num), np.asarray(pos), None if lat is None else np.asarray(lat),
                        None if pbc is None else np.asarray(pbc))
    if cut is not None:
        m.set_realspace_cutoff(cut[1], cut[2], cut[0], cut[3], cut[4])
    if precision != "fp64":
        m.set_precision(precision)
    return m


def _check(res, ref, n, precision):
    if precision == "fp64":
        assert abs(res["energy"] - ref["energy"]) <= 1e-10 * abs(ref["energy"]) + 1e-14
        assert np.abs(res["gradient"] - ref["gradient"]).max() <= 1e-9
        assert np.abs(res["virial"] - ref["virial"]).max() <= 1e-9
    else:
        assert abs(res["energy"] - ref["energy"]) <= 1e-8 * n
        assert np.abs(res["gradient"] - ref["gradient"]).max() <= 1e-7


@pytest.mark.parametrize("name", ["validation_03-energy-d3-bjm", "validation_04-energy-d3-zerom",
                                  "validation_05-energy-d3-op", "validation_08-gradient-d3-bjm",
                                  "validation_09-gradient-d3-zerom", "validation_10-gradient-d3-op",
                                  "validation_14-pairwise-d3-zerom", "validation_15-pairwise-d3-op",
                                  "validation_16-db-d3-bj", "validation_17-db-d3-zero",
                                  "validation_18-db-d3-bjm", "validation_19-db-d3-zerom",
                                  "validation_20-db-d3-op"])
def test_family_goldens(family_goldens, name):
    f = family_goldens[name]
    _step(f"golden {name}")
    e = f["expected"]
    m = _model(f["numbers"], f["positions"], f["lattice"], f["periodic"])
    par = _param(f["kind"], f["param"])
    res = m.get_dispersion(par, grad=True)
    atol = 1e-12 if f["version"] == "1.2.1" else f["atol"]
    assert abs(float(res["energy"]) - e["energy"]) <= atol
    ref = O.dispersion(f["numbers"], f["positions"], f["lattice"], f["periodic"], KINDS[f["kind"]], f["param"],
                       f["cutoff"], wrap_cell=f["wrap"])
    _check(dict(energy=float(res["energy"]), gradient=res["gradient"], virial=res["virial"]), ref,
           len(f["numbers"]), "fp64")
    if "gradient" in e:
        np.testing.assert_allclose(res["gradient"], np.array(e["gradient"]), rtol=0, atol=f["atol"])
        np.testing.assert_allclose(res["virial"], np.array(e["virial"]), rtol=0, atol=f["atol"])
    if "pair2" in e:
        pw = m.get_pairwise_dispersion(par)
        np.testing.assert_allclose(pw["additive pairwise energy"], np.array(e["pair2"]), rtol=0, atol=f["atol"])
        np.testing.assert_allclose(pw["non-additive pairwise energy"], np.array(e["pair3"]), rtol=0,
                                   atol=f["atol"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", list(PARAMS))
def test_family_molecule_and_cells_vs_oracle(kind, precision):
    """Config-0 molecule and small periodic cells (self images, smooth
    cutoff widths) of every family, energy, gradient and virial."""
    from dftd3 import synthetic
    par = PARAMS[kind]
    num, pos = synthetic.molecule_c0(seed=3, n=60)
    cases = [(num, pos, None, None, None)]
    for seed, n, edge in [(81, 5, 10.0), (82, 9, 12.0)]:
        cn, cx, lat = cell(seed, n, edge)
        cases.append((cn, cx, lat, [True] * 3, (40.0, 60.0, 40.0, 3.0, 4.0)))
    for n_, x, lat, pbc, cut in cases:
        _step(f"{kind} {precision} n={len(n_)} periodic={pbc is not None}")
        m = _model(n_, x, lat, pbc, cut, precision)
        res = m.get_dispersion(_param(kind, par), grad=True)
        ref = O.dispersion(n_, x, lat, pbc, KINDS[kind], par, cut or O.DEFAULT_CUTOFF)
        _check(dict(energy=float(res["energy"]), gradient=res["gradient"], virial=res["virial"]), ref, len(n_),
               precision)
        if precision == "fp64":
            pw = m.get_pairwise_dispersion(_param(kind, par))
            r2, r3 = O.pairwise(n_, x, lat, pbc, KINDS[kind], par, cut or O.DEFAULT_CUTOFF)
            scale = np.abs(r2).max()
            np.testing.assert_allclose(pw["additive pairwise energy"], r2, rtol=0, atol=1e-10 * scale)
            np.testing.assert_allclose(pw["non-additive pairwise energy"], r3, rtol=0,
                                       atol=1e-10 * np.abs(r3).max())


def test_family_batch_with_per_system_species_ids():
    """One batch of every family; the Z-damped systems list their elements in
    different first-appearance orders, so their species identifiers differ
    from the batch-wide species numbering."""
    from dftd3 import synthetic, torch_api
    systems = []
    for k, kind in enumerate(["z", "mzero", "optimizedpower", "cso", "z", "rational", "zero"]):
        num, pos = synthetic.molecule_c0(seed=90 + k, n=15 + 6 * k)
        if k == 4:  # reverse: another element first
            num, pos = num[::-1].copy(), pos[::-1].copy()
        par = PARAMS.get(kind) or (dict(s6=1.0, s8=1.2177, a1=0.4145, a2=4.8593, s9=1.0) if kind == "rational"
                                   else dict(s6=1.0, s8=1.703, rs6=1.261, rs8=1.0, s9=1.0))
        systems.append(dict(numbers=num, positions=pos, kind=kind, param=par))
    spec, pos = torch_api.batch_spec(systems)
    pos = torch.as_tensor(pos, dtype=torch.float64, device="cuda").contiguous()
    for precision in ("fp64", "mixed"):
        _step(f"batch {precision}")
        out = torch_api.run_batch(spec, pos, precision=precision)
        torch.cuda.synchronize()
        off = 0
        for k, s in enumerate(systems):
            n = len(s["numbers"])
            ref = O.dispersion(s["numbers"], s["positions"], None, None, KINDS[s["kind"]], s["param"],
                               O.DEFAULT_CUTOFF)
            sig = out["sigma"][k].cpu().numpy()
            _check(dict(energy=out["energy"][k].item(), gradient=out["gradient"][off:off + n].cpu().numpy(),
                        virial=sig), ref, n, precision)
            off += n
