"""Analytic Hessian on the GPU (dftd3_get_dispersion_hessian,
d3hip_get_dispersion_hessian_batch) against the oracle's restatement of
disp.f90:236-308 and, as in python/dftd3/test_interface.py:359-442, against
central differences of the GPU gradient.

Tolerance: every element within 1e-10 of the largest element of the
oracle's Hessian (FP64; the three-body terms are summed with atomics).
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
    "rational": dict(s6=1.0, s8=1.2177, a1=0.4145, a2=4.8593, s9=1.0, alp=14.0),
    "zero": dict(s6=1.0, s8=1.703, rs6=1.261, rs8=1.0, s9=1.0, alp=14.0),
    "mzero": dict(s6=1.0, s8=1.841686, rs6=1.279637, rs8=1.0, bet=0.014370, s9=1.0, alp=14.0),
    "optimizedpower": dict(s6=1.0, s8=0.51581, a1=0.575, a2=3.0, bet=8.0, s9=1.0, alp=14.0),
    "cso": dict(s6=1.0, a1=0.2, a2=2.5, a3=0.0, a4=6.25, s9=1.0, alp=14.0),
    "z": dict(s6=1.0, s8=1.0, a1=200770.0, s9=1.0, alp=14.0),
}
WATER = (np.array([8, 1, 1, 8, 1, 1]),
         np.array([[-2.85, -0.05, +0.00], [-1.05, +0.00, +0.00], [-3.40, +1.65, +0.00],
                   [+2.65, +0.00, +0.00], [+3.30, +1.00, +1.45], [+3.30, +1.00, -1.45]]))


def _step(msg):
    print(f"[gpu-hessian] {msg}", flush=True)
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


def _close(h, ref):
    np.testing.assert_allclose(h, ref, rtol=0, atol=1e-10 * np.abs(ref).max())


@pytest.mark.parametrize("ghost", [None, [2, 3]])
def test_hessian_water_dimer(ghost):
    """The reference's Python test case: PBE0-D3(BJ)-ATM loaded by name;
    symmetric, no NaN with ghosts, equal to central differences (step 1e-6)
    of the GPU gradient within 1e-7, and to the oracle within 1e-10."""
    from dftd3.interface import DispersionModel, RationalDampingParam
    _step(f"water dimer ghost={ghost}")
    num, pos = WATER
    model = DispersionModel(num, pos)
    mask = None
    if ghost is not None:
        model.set_ghost_atoms(ghost)
        mask = np.zeros(len(num), dtype=np.int32)
        mask[ghost] = 1
    param = RationalDampingParam(method="pbe0", atm=True)
    res = model.get_hessian(param)
    h = res["hessian"]
    assert h.shape == (18, 18) and not np.isnan(h).any()
    np.testing.assert_allclose(h, h.T, rtol=0, atol=1e-12)
    ref = O.hessian(num, pos, None, None, O.RATIONAL, PBE0_ATM, O.DEFAULT_CUTOFF, ghost=mask)
    _close(h, ref["hessian"])
    assert abs(float(res["energy"]) - ref["energy"]) <= 1e-10 * abs(ref["energy"])
    step = 1.0e-6
    fd = np.zeros_like(h)
    for i in range(len(num)):
        for c in range(3):
            x = pos.copy()
            x[i, c] += step
            model.update(x)
            gp = model.get_dispersion(param, grad=True)["gradient"]
            x[i, c] -= 2 * step
            model.update(x)
            gm = model.get_dispersion(param, grad=True)["gradient"]
            fd[3 * i + c] = 0.5 * (gp - gm).ravel() / step
    model.update(pos)
    np.testing.assert_allclose(h, fd, rtol=0, atol=1e-7)


PBE0_ATM = dict(s6=1.0, s8=1.2177, a1=0.4145, a2=4.8593, s9=1.0, alp=14.0)


@pytest.mark.parametrize("kind", list(PARAMS))
def test_hessian_families_vs_oracle(kind):
    from dftd3 import synthetic
    from dftd3.interface import DispersionModel
    num, pos = synthetic.molecule_c0(seed=21, n=24)
    _step(f"molecule {kind}")
    h = DispersionModel(num, pos).get_hessian(_param(kind, PARAMS[kind]))["hessian"]
    _close(h, O.hessian(num, pos, None, None, KINDS[kind], PARAMS[kind])["hessian"])


@pytest.mark.parametrize("kind", ["rational", "zero", "mzero"])
def test_hessian_periodic_cells_vs_oracle(kind):
    """Cells smaller than the cutoffs (self images, triples with repeated
    atoms) with smooth cutoff widths, 3D and 2D."""
    from dftd3.interface import DispersionModel
    cut = (20.0, 25.0, 12.0, 2.0, 3.0)
    for seed, n, edge, pbc in [(7, 3, 8.0, [True] * 3), (8, 4, 9.0, [True, True, False])]:
        _step(f"cell {kind} seed={seed} pbc={pbc}")
        num, x, lat = cell(seed, n, edge)
        m = DispersionModel(num, x, lat, np.array(pbc))
        m.set_realspace_cutoff(cut[1], cut[2], cut[0], cut[3], cut[4])
        h = m.get_hessian(_param(kind, PARAMS[kind]))["hessian"]
        _close(h, O.hessian(num, x, lat, pbc, KINDS[kind], PARAMS[kind], cut)["hessian"])


def test_hessian_batch_entry_point():
    from dftd3 import synthetic, torch_api
    systems = []
    for k, kind in enumerate(["rational", "zero", "z", "optimizedpower"]):
        num, pos = synthetic.molecule_c0(seed=30 + k, n=6 + 5 * k)
        systems.append(dict(numbers=num, positions=pos, kind=kind, param=dict(PARAMS[kind], s9=0.0 if k == 1 else 1.0)))
    num, pos, lat = cell(9, 4, 9.5)
    systems.append(dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, kind="rational",
                        param=PARAMS["rational"]))
    spec, pos = torch_api.batch_spec(systems)
    pos = torch.as_tensor(pos, dtype=torch.float64, device="cuda").contiguous()
    _step("batch")
    out = torch_api.run_hessian(spec, pos)
    torch.cuda.synchronize()
    blocks = torch_api.hessian_blocks(out)
    for k, s in enumerate(systems):
        ref = O.hessian(s["numbers"], s["positions"], s.get("lattice"), s.get("periodic"), KINDS[s["kind"]],
                        s["param"])
        _close(blocks[k].cpu().numpy(), ref["hessian"])
        assert abs(out["energy"][k].item() - ref["energy"]) <= 1e-10 * abs(ref["energy"])
