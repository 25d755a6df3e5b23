"""GPU parity of the HIP library against the CPU oracle and the reference's
golden vectors.

Bars (checked against the FP64 oracle on the same inputs):
  plain FP64 path:  |dE| <= 1e-10 |E| (+1e-14), max |dg| <= 1e-9 Eh/bohr,
                    max |d sigma| <= 1e-9 Eh
  mixed path:       |dE| <= 1e-8 Eh per atom, max |dg| <= 1e-7 Eh/bohr
"""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import BJ, ZERO, cell as _cell

pytestmark = pytest.mark.gpu

E_REL, G_ABS, S_ABS = 1e-10, 1e-9, 1e-9
MIX_E_PER_ATOM, MIX_G = 1e-8, 1e-7


def _step(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _oracle(numbers, pos, lat=None, pbc=None, kind="rational", param=BJ, cutoff=O.DEFAULT_CUTOFF, grad=True,
            ghost=None, wrap=False):
    return O.dispersion(numbers, pos, lat, pbc, O.RATIONAL if kind == "rational" else O.ZERO, param, cutoff,
                        grad=grad, ghost=ghost, wrap_cell=wrap)


def _check(res, ref, nat, mixed=False):
    e, er = float(res["energy"]), ref["energy"]
    if mixed:
        assert abs(e - er) <= MIX_E_PER_ATOM * nat, (e, er)
    else:
        assert abs(e - er) <= E_REL * abs(er) + 1e-14, (e, er, abs(e - er) / abs(er))
    if "gradient" in res and "gradient" in ref:
        dg = np.abs(np.asarray(res["gradient"]) - ref["gradient"]).max()
        assert dg <= (MIX_G if mixed else G_ABS), dg
        ds = np.abs(np.asarray(res["virial"]) - ref["virial"]).max()
        assert ds <= (MIX_G * nat if mixed else S_ABS), ds


def _batch(systems, precision="fp64", grad=True):
    from dftd3.torch_api import get_dispersion_batch
    out = get_dispersion_batch(systems, precision=precision, grad=grad)
    res = []
    off = 0
    for k, s in enumerate(systems):
        n = len(s["numbers"])
        r = dict(energy=out["energy"][k].item())
        if grad:
            r["gradient"] = out["gradient"][off:off + n].cpu().numpy()
            r["virial"] = out["sigma"][k].cpu().numpy()
        off += n
        res.append(r)
    return res


@pytest.mark.parametrize("name", ["ta90_3d_bj", "bp86_zero_c4_2d", "bp86_zero_c4_vac3d", "gh185_bj_2d",
                                  "gh185_bj_vac3d"])
def test_reference_tight_goldens(goldens, name):
    f = goldens[name]
    _step(f"tight golden {name}")
    cut = dict(cn=f["cutoff"][0], disp2=f["cutoff"][1], disp3=f["cutoff"][2])
    sysd = dict(numbers=f["numbers"], positions=f["positions"], lattice=f["lattice"], periodic=f["periodic"],
                kind=f["kind"], param=f["param"], cutoff=cut)
    r = _batch([sysd], grad=f["grad"])[0]
    ref = f["expected"]["energy"]
    assert abs(r["energy"] - ref) <= max(f["atol"], E_REL * abs(ref)), (r["energy"], ref)
    if "gradient_norm" in f["expected"]:
        assert abs(np.linalg.norm(r["gradient"]) - f["expected"]["gradient_norm"]) <= G_ABS
    orc = _run_golden_oracle(f)
    _check(r, orc, len(f["numbers"]))


def _run_golden_oracle(f):
    return _oracle(f["numbers"], f["positions"], f["lattice"], f["periodic"], f["kind"], f["param"], f["cutoff"],
                   grad=f["grad"], wrap=False)


@pytest.mark.parametrize("name", ["validation_01-energy-d3-bj", "validation_02-energy-d3-zero",
                                  "validation_06-gradient-d3-bj", "validation_07-gradient-d3-zero"])
def test_validation_goldens_through_interface(goldens, name):
    from dftd3.interface import DispersionModel, RationalDampingParam, ZeroDampingParam
    f = goldens[name]
    _step(f"validation {name}")
    model = DispersionModel(np.array(f["numbers"]), np.array(f["positions"]),
                            None if f["lattice"] is None else np.array(f["lattice"]),
                            None if f["periodic"] is None else np.array(f["periodic"]))
    cls = RationalDampingParam if f["kind"] == "rational" else ZeroDampingParam
    res = model.get_dispersion(cls(method=f["method"], atm=f["atm"]), grad=f["grad"])
    e = f["expected"]
    assert abs(float(res["energy"]) - e["energy"]) <= f["atol"]
    if "gradient" in e:
        np.testing.assert_allclose(res["gradient"], np.array(e["gradient"]), atol=f["atol"], rtol=0)
        np.testing.assert_allclose(res["virial"], np.array(e["virial"]), atol=f["atol"], rtol=0)
    _check(res, _run_golden_oracle(dict(f, wrap=True)), len(f["numbers"]))


@pytest.mark.parametrize("atm", [True, False])
@pytest.mark.parametrize("kind,method", [("rational", "pbe0"), ("zero", "b3lyp"), ("mrational", "pbe"),
                                         ("mzero", "bp"), ("optimizedpower", "b97d"), ("z", "")])
def test_python_interface_model(python_model, kind, method, atm):
    """python/dftd3/test_interface.py:236-356 through dftd3.interface, and
    against the oracle on the same parameters."""
    from dftd3 import interface as I
    from test_oracle_golden import KINDS, python_model_param
    _step(f"python model {kind} {method} atm={atm}")
    m = I.DispersionModel(np.array(python_model["numbers"]), np.array(python_model["positions"]))
    cls = dict(rational=I.RationalDampingParam, zero=I.ZeroDampingParam, mrational=I.ModifiedRationalDampingParam,
               mzero=I.ModifiedZeroDampingParam, optimizedpower=I.OptimizedPowerDampingParam,
               z=I.ZDampingParam)[kind]
    if method:
        param = cls(method=method, atm=atm)
    else:  # ZDampingParam(a1=..., s9=1.0 if atm else 0.0)
        param = cls(**python_model["explicit"][kind], s9=1.0 if atm else 0.0)
    res = m.get_dispersion(param, grad=False)
    ref = python_model["energies"][f"{kind}:{method}"]["atm" if atm else "noatm"]
    assert float(res["energy"]) == pytest.approx(ref, rel=python_model["rel"])
    assert set(res) == {"energy"} and res["energy"].shape == ()
    okind, p = python_model_param(python_model, f"{kind}:{method}", atm)
    orc = O.dispersion(python_model["numbers"], python_model["positions"], None, None, KINDS[okind], p, grad=False)
    assert abs(float(res["energy"]) - orc["energy"]) <= E_REL * abs(orc["energy"])


def test_config0_molecule_fp64():
    from dftd3 import synthetic
    from dftd3.interface import DispersionModel, RationalDampingParam
    _step("config 0, 200-atom molecule, BJ + ATM, FP64")
    num, pos = synthetic.molecule_c0()
    model = DispersionModel(num, pos)
    res = model.get_dispersion(RationalDampingParam(**BJ), grad=True)
    assert res["gradient"].shape == (200, 3) and res["virial"].shape == (3, 3)
    _check(res, _oracle(num, pos), 200)


def test_config0_molecule_mixed():
    from dftd3 import synthetic
    from dftd3.interface import DispersionModel, RationalDampingParam
    _step("config 0, mixed precision")
    num, pos = synthetic.molecule_c0()
    model = DispersionModel(num, pos)
    model.set_precision("mixed")
    res = model.get_dispersion(RationalDampingParam(**BJ), grad=True)
    _check(res, _oracle(num, pos), 200, mixed=True)


@pytest.mark.parametrize("kind", ["rational", "zero"])
def test_energy_only_path(kind):
    from dftd3 import synthetic
    _step(f"energy only {kind}")
    num, pos = synthetic.molecule_c0(seed=5, n=80)
    p = BJ if kind == "rational" else ZERO
    r = _batch([dict(numbers=num, positions=pos, kind=kind, param=p)], grad=False)[0]
    _check(r, _oracle(num, pos, kind=kind, param=p, grad=False), 80)


def test_zero_damping_atm_ghost_and_smooth_cutoff():
    from dftd3 import synthetic
    from dftd3.interface import DispersionModel, ZeroDampingParam
    _step("zero damping, ghosts, smooth cutoffs")
    num, pos = synthetic.molecule_c0(seed=9, n=60)
    model = DispersionModel(num, pos)
    model.set_ghost_atoms([0, 5, 17])
    model.set_realspace_cutoff(disp2=14.0, disp3=10.0, cn=25.0, width2=4.0, width3=3.0)
    res = model.get_dispersion(ZeroDampingParam(**ZERO), grad=True)
    ghost = np.zeros(60, dtype=np.int32)
    ghost[[0, 5, 17]] = 1
    ref = _oracle(num, pos, kind="zero", param=ZERO, cutoff=(25.0, 14.0, 10.0, 4.0, 3.0), ghost=ghost)
    _check(res, ref, 60)


@pytest.mark.parametrize("kind", ["rational", "zero"])
def test_periodic_cell_default_cutoffs(kind):
    from dftd3.interface import DispersionModel, RationalDampingParam, ZeroDampingParam
    _step(f"periodic 16-atom cell {kind} + ATM")
    num, pos, lat = _cell(21, 16, 11.0)
    model = DispersionModel(num, pos, lat, np.array([True, True, True]))
    p = BJ if kind == "rational" else ZERO
    cls = RationalDampingParam if kind == "rational" else ZeroDampingParam
    res = model.get_dispersion(cls(**p), grad=True)
    _check(res, _oracle(num, pos, lat, [True] * 3, kind, p, wrap=True), 16)


@pytest.mark.parametrize("natoms,edge", [(1, 6.5), (2, 7.5)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_small_cell_self_image_triples(natoms, edge, precision):
    """Cells much smaller than the three-body cutoff: most triples are built
    from images of one atom (weights 1/6, 1/2 of the reference's
    triple_scale) and long slot lists (many 64-slot tiles per atom)."""
    _step(f"{natoms}-atom cell, edge {edge}, {precision}")
    num, pos, lat = _cell(40 + natoms, natoms, edge)
    r = _batch([dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, param=BJ)], precision=precision)[0]
    _check(r, _oracle(num, pos, lat, [True] * 3, "rational", BJ), natoms, mixed=precision == "mixed")


def test_periodic_mixed_long_lists():
    """Mixed path on a dense periodic cell (about 80 tiles per slot list)."""
    _step("32-atom cell, mixed, BJ + ATM")
    num, pos, lat = _cell(51, 32, 12.0)
    r = _batch([dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, param=BJ)], precision="mixed")[0]
    _check(r, _oracle(num, pos, lat, [True] * 3, "rational", BJ), 32, mixed=True)


@pytest.mark.parametrize("alp", [14.0, 16.0, 13.5])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_three_body_damping_exponent(alp, precision):
    """The three-body damping power (r0/r)^((alp + 2)/3) for the default
    alp = 14 (the FP64 kernel's Newton-refined power) and for other values
    (the general power), periodic cell and molecule, against the oracle."""
    _step(f"alp {alp} {precision}")
    p = dict(BJ, alp=alp)
    num, pos, lat = _cell(61, 20, 11.0)
    mol_num, mol_pos, _ = _cell(62, 24, 14.0)
    rs = _batch([dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, param=p),
                 dict(numbers=mol_num, positions=mol_pos, param=p)], precision=precision)
    _check(rs[0], _oracle(num, pos, lat, [True] * 3, "rational", p), 20, mixed=precision == "mixed")
    _check(rs[1], _oracle(mol_num, mol_pos, None, None, "rational", p), 24, mixed=precision == "mixed")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_elongated_molecule_tile_culling(precision):
    """A 640-atom rod 420 bohr long: most 64-atom tiles of the spatial order
    lie beyond the CN, two-body and three-body cutoffs of an atom and are
    skipped by the pair loops (CN, two-body, three-body lists, chain rule);
    energy, gradient and virial must be the oracle's."""
    _step(f"640-atom rod {precision}")
    rng = np.random.default_rng(17)
    n = 640
    num = rng.choice([1, 6, 7, 8], size=n, p=[0.5, 0.3, 0.1, 0.1])
    pos = np.column_stack([np.linspace(0.0, 420.0, n) + rng.uniform(-0.2, 0.2, n),
                           rng.uniform(-4.0, 4.0, n), rng.uniform(-4.0, 4.0, n)])
    sys_ = dict(numbers=num, positions=pos, param=BJ)
    r = _batch([sys_], precision=precision)[0]
    _check(r, _oracle(num, pos), n, mixed=precision == "mixed")


@pytest.mark.parametrize("scale", [1.0, 1.0 + 3e-7, 1.0 - 3e-7])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_three_body_cutoff_band(precision, scale):
    """Slot pairs at the three-body cutoff: a 3 x 3 x 4 grid of spacing 3
    bohr (times 1, 1 + 3e-7, 1 - 3e-7) with a 6-bohr three-body cutoff, so
    that many r_jk^2 fall exactly at 36 (taken: r^2 <= c^2, atm.f90:723-724)
    or within 2e-6 of it on either side; a triple taken or dropped wrongly
    moves the energy by ~1e-5 Eh."""
    _step(f"three-body cutoff band {precision} scale {scale}")
    rng = np.random.default_rng(5)
    g = np.array([[x, y, z] for x in range(3) for y in range(3) for z in range(4)], dtype=float)
    pos = 3.0 * scale * g
    num = rng.choice([1, 6], size=len(pos))
    sys_ = dict(numbers=num, positions=pos, param=BJ, cutoff=dict(disp3=6.0))
    r = _batch([sys_], precision=precision)[0]
    _check(r, _oracle(num, pos, cutoff=(O.DEFAULT_CUTOFF[0], O.DEFAULT_CUTOFF[1], 6.0, 0.0, 0.0)), len(pos),
           mixed=precision == "mixed")


def test_periodic_2d_slab():
    _step("2D slab BJ + ATM")
    num, pos, lat = _cell(22, 10, 9.0)
    lat[2] = [0.0, 0.0, 30.0]
    pos[:, 2] = np.random.default_rng(1).uniform(0, 6, len(pos))
    r = _batch([dict(numbers=num, positions=pos, lattice=lat, periodic=[True, True, False], param=BJ)])[0]
    _check(r, _oracle(num, pos, lat, [True, True, False], "rational", BJ), 10)


@pytest.mark.parametrize("kind", ["rational", "zero"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_periodic_1d_chain(kind, precision):
    """One periodic direction (cutoff.f90:157-163): a chain of 12 atoms with
    a 15-bohr repeat, BJ and zero damping with ATM."""
    _step(f"1D chain {kind} {precision}")
    rng = np.random.default_rng(8)
    num = rng.choice([1, 6, 7, 8, 16], size=12)
    pos = np.column_stack([np.linspace(0.0, 13.6, 12), rng.uniform(0, 5, 12), rng.uniform(0, 5, 12)])
    lat = np.array([[15.0, 0.3, -0.2], [0.0, 40.0, 0.0], [0.0, 0.0, 40.0]])
    p = BJ if kind == "rational" else ZERO
    r = _batch([dict(numbers=num, positions=pos, lattice=lat, periodic=[True, False, False], kind=kind, param=p)],
               precision=precision)[0]
    _check(r, _oracle(num, pos, lat, [True, False, False], kind, p), 12, mixed=precision == "mixed")


def test_term_counts_match_oracle_enumeration():
    """GPU work-unit counts (used for pair-/triple-terms/s) equal the
    oracle's enumeration of the reference's loops."""
    import torch
    from dftd3 import synthetic, torch_api
    _step("term counts")
    num0, pos0 = synthetic.molecule_c0(seed=13, n=70)
    num1, pos1, lat1 = _cell(23, 14, 9.5)
    systems = [dict(numbers=num0, positions=pos0, param=BJ),
               dict(numbers=num1, positions=pos1, lattice=lat1, periodic=[True] * 3, param=BJ)]
    spec, pos = torch_api.batch_spec(systems)
    counts = torch_api.count_terms(spec, torch.as_tensor(pos, device="cuda"))
    assert tuple(counts[0]) == O.count_terms(pos0)
    assert tuple(counts[1]) == O.count_terms(pos1, lat1, [True] * 3)


def test_batch_mixed_systems():
    _step("batch of 6 mixed systems")
    from dftd3 import synthetic
    systems, refs = [], []
    for k in range(6):
        if k % 3 == 2:
            num, pos, lat = _cell(30 + k, 12, 10.0)
            s = dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3)
        else:
            num, pos = synthetic.molecule_c0(seed=40 + k, n=15 + 9 * k)
            s = dict(numbers=num, positions=pos)
        kind = "rational" if k % 2 == 0 else "zero"
        s.update(kind=kind, param=BJ if kind == "rational" else ZERO)
        systems.append(s)
        refs.append(_oracle(s["numbers"], s["positions"], s.get("lattice"), s.get("periodic"), kind, s["param"]))
    for prec in ("fp64", "mixed"):
        for r, ref, s in zip(_batch(systems, precision=prec), refs, systems):
            _check(r, ref, len(s["numbers"]), mixed=prec == "mixed")


def test_many_elements_in_one_batch():
    """No limit on the distinct elements of a batch (model.f90:87-181
    allocates per species): three molecules drawing on 40 elements (H to Zr),
    one of them holding 30 distinct elements, against the oracle."""
    _step("40 distinct elements in one batch")
    rng = np.random.default_rng(12)
    systems = []
    for k, (zs, n) in enumerate([(np.arange(1, 31), 36), (np.arange(20, 41), 24), (np.array([1, 6, 7, 8]), 12)]):
        num = np.concatenate([zs, rng.choice(zs, size=n - len(zs))]) if n > len(zs) else zs[:n]
        edge = (n / 0.01) ** (1.0 / 3.0)  # 0.01 atoms per bohr^3, no two closer than 2.5 bohr
        pts = []
        while len(pts) < n:
            x = rng.uniform(0, edge, 3)
            if all(np.linalg.norm(x - q) > 2.5 for q in pts):
                pts.append(x)
        systems.append(dict(numbers=num, positions=np.array(pts), param=BJ))
    res = _batch(systems)
    assert len(set(np.concatenate([s["numbers"] for s in systems]).tolist())) == 40
    for s, r in zip(systems, res):
        _check(r, _oracle(s["numbers"], s["positions"]), len(s["numbers"]))
