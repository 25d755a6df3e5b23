"""Pins of the CPU oracle (oracle/d3_oracle.c) -- CPU only.

1. Reference golden vectors: the 100-epsilon periodic regressions of the
   reference's unit tests, the CLI validation JSONs (abs 1e-6) and the
   Python interface energies (rel 1e-6), see tests/golden/make_fixtures.py.
2. What the mathematics fixes: dimer closed forms for CN, C6 and the two-body
   term, the equilateral ATM triangle, finite differences of the energy for
   gradient and virial, the virial theorem for molecules, translation and
   rotation invariance, and the coverage of the lattice-image enumeration.
"""
import numpy as np
import pytest

import oracle as O

EPS = np.finfo(float).eps
TABLES = np.load(O.TABLES)


KINDS = {"rational": O.RATIONAL, "zero": O.ZERO, "mzero": O.MZERO, "optimizedpower": O.OPTIMIZEDPOWER,
         "cso": O.CSO, "z": O.ZDAMP}


def _kind(name):
    return KINDS[name]


def _run(f, **kw):
    return O.dispersion(f["numbers"], f["positions"], f["lattice"], f["periodic"], _kind(f["kind"]), f["param"],
                        f["cutoff"], grad=kw.pop("grad", f["grad"]), wrap_cell=f["wrap"], **kw)


TIGHT = ["ta90_3d_bj", "bp86_zero_c4_2d", "bp86_zero_c4_vac3d", "gh185_bj_2d", "gh185_bj_vac3d"]
LOOSE = ["validation_01-energy-d3-bj", "validation_02-energy-d3-zero", "validation_06-gradient-d3-bj",
         "validation_07-gradient-d3-zero"]


@pytest.mark.parametrize("name", TIGHT)
def test_tight_goldens(goldens, name):
    f = goldens[name]
    r = _run(f)
    assert abs(r["energy"] - f["expected"]["energy"]) <= f["atol"]
    if "gradient_norm" in f["expected"]:
        assert abs(np.linalg.norm(r["gradient"]) - f["expected"]["gradient_norm"]) <= f["atol"]


@pytest.mark.parametrize("name", LOOSE)
def test_validation_goldens(goldens, name):
    f = goldens[name]
    r = _run(f)
    e = f["expected"]
    assert abs(r["energy"] - e["energy"]) <= f["atol"]
    if "gradient" in e:
        np.testing.assert_allclose(r["gradient"], np.array(e["gradient"]), rtol=0, atol=f["atol"])
        np.testing.assert_allclose(r["virial"], np.array(e["virial"]), rtol=0, atol=f["atol"])


def python_model_param(python_model, key, atm):
    """Damping kind and parameters of an energy of the reference's 16-atom
    Python test model: loaded by method name from the parameter database
    (`load_param(method, atm)`: s9 = 1 with the three-body term, else 0), or
    given explicitly (Z damping: a1 with the other keywords at their
    defaults, s9 = 1 or 0)."""
    import json
    import os
    kind, method = key.split(":")
    if method:
        params = json.load(open(os.path.join(os.path.dirname(O.TABLES), "d3_params.json")))
        p = dict(params["defaults"])
        p.update(params[kind][params["aliases"][method]])
    else:
        p = dict(s6=1.0, s8=1.0, alp=14.0)
        p.update(python_model["explicit"][kind])
    p["s9"] = 1.0 if atm else 0.0
    return ("rational" if kind == "mrational" else kind), p


PY_KEYS = ["rational:pbe0", "zero:b3lyp", "mrational:pbe", "mzero:bp", "optimizedpower:b97d", "z:"]


@pytest.mark.parametrize("key", PY_KEYS)
@pytest.mark.parametrize("atm", [True, False])
def test_python_interface_energies(python_model, key, atm):
    """python/dftd3/test_interface.py:236-356: BJ, zero, modified rational,
    modified zero, optimized power and Z damping on the 16-atom model."""
    kind, p = python_model_param(python_model, key, atm)
    ref = python_model["energies"][key]["atm" if atm else "noatm"]
    r = O.dispersion(python_model["numbers"], python_model["positions"], None, None, _kind(kind), p, grad=False)
    assert r["energy"] == pytest.approx(ref, rel=python_model["rel"])


def _c6_single_ref(z1, z2):
    a, b = max(z1, z2), min(z1, z2)
    ic = b + a * (a - 1) // 2 - 1
    return TABLES["c6ref"][ic, 0, 0]


def _r4r2(z):
    return np.sqrt(0.5 * TABLES["r4_over_r2"][z - 1] * np.sqrt(z))


def test_cn_dimer_closed_form():
    r = 2.1
    cn = O.coordination_number([1, 1], [[0, 0, 0], [0, 0, r]])
    rc = 2 * 4.0 / 3.0 * 0.32 * O.ORACLE_AATOAU
    expect = 1.0 / (1.0 + np.exp(-16.0 * (rc / r - 1.0)))
    np.testing.assert_allclose(cn, [expect, expect], rtol=1e-14)


def test_weights_sum_to_one_and_single_reference():
    num = [1, 6, 7, 8, 14, 16, 2]
    cn = [0.3, 3.7, 2.2, 1.1, 3.9, 2.0, 0.4]
    gw, gwd = O.weights(num, cn)
    np.testing.assert_allclose(gw.sum(axis=1), 1.0, rtol=1e-14)
    np.testing.assert_allclose(gwd.sum(axis=1), 0.0, atol=1e-13)
    # He has a single reference: weight one, derivative zero
    assert gw[6, 0] == pytest.approx(1.0, abs=1e-15) and abs(gwd[6, 0]) < 1e-15


def test_weights_exceptional_fallback():
    # CN far above every reference CN: all Gaussians underflow; the reference
    # falls back to the highest reference CN (model.f90:260-266)
    gw, gwd = O.weights([6], [60.0])
    refcn = TABLES["refcn"][5, :TABLES["nref"][5]]
    expect = (refcn == refcn.max()).astype(float)
    np.testing.assert_array_equal(gw[0, :len(expect)], expect)
    np.testing.assert_array_equal(gwd[0], 0.0)


def test_weight_derivative_finite_difference():
    num = [6, 7, 8, 1, 14]
    cn = np.array([2.7, 1.4, 0.8, 0.95, 3.1])
    gw, gwd = O.weights(num, cn)
    h = 1e-6
    gp, _ = O.weights(num, cn + h)
    gm, _ = O.weights(num, cn - h)
    np.testing.assert_allclose(gwd, (gp - gm) / (2 * h), atol=1e-8)


def test_c6_single_reference_elements():
    # He and Ne have one reference system: C6 is the table entry
    num = [2, 10]
    gw, gwd = O.weights(num, [0.0, 0.0])
    c6, dc6 = O.atomic_c6(num, gw, gwd)
    assert c6[0, 1] == pytest.approx(_c6_single_ref(2, 10), rel=1e-15)
    assert c6[0, 0] == pytest.approx(_c6_single_ref(2, 2), rel=1e-15)
    np.testing.assert_allclose(dc6, 0.0, atol=1e-12)


@pytest.mark.parametrize("kind", ["rational", "zero"])
def test_two_body_dimer_closed_form(kind):
    r = 6.3
    num = [2, 10]
    pos = np.array([[0.0, 0.0, 0.0], [0.3, -0.4, r]])
    d = np.linalg.norm(pos[1] - pos[0])
    c6 = _c6_single_ref(2, 10)
    rrij = 3 * _r4r2(2) * _r4r2(10)
    if kind == "rational":
        p = dict(s6=0.9, s8=1.3, a1=0.41, a2=4.9, s9=0.0, alp=14.0)
        r0 = p["a1"] * np.sqrt(rrij) + p["a2"]
        e = lambda x: -c6 * (p["s6"] / (x ** 6 + r0 ** 6) + p["s8"] * rrij / (x ** 8 + r0 ** 8))
        de = lambda x: -c6 * (-p["s6"] * 6 * x ** 5 / (x ** 6 + r0 ** 6) ** 2
                              - p["s8"] * rrij * 8 * x ** 7 / (x ** 8 + r0 ** 8) ** 2)
    else:
        p = dict(s6=1.0, s8=0.7, rs6=1.2, rs8=1.0, s9=0.0, alp=14.0)
        a, b = max(2, 10), min(2, 10)
        r0 = O.ORACLE_AATOAU * TABLES["vdwrad_aa"][b + a * (a - 1) // 2 - 1]

        def e(x):
            f6 = 1 / (1 + 6 * (p["rs6"] * r0 / x) ** 14)
            f8 = 1 / (1 + 6 * (p["rs8"] * r0 / x) ** 16)
            return -c6 * (p["s6"] * f6 / x ** 6 + p["s8"] * rrij * f8 / x ** 8)

        de = lambda x: (e(x + 1e-6) - e(x - 1e-6)) / 2e-6
    res = O.dispersion(num, pos, kind=_kind(kind), param=p)
    assert res["energy"] == pytest.approx(e(d), rel=1e-13)
    u = (pos[0] - pos[1]) / d
    np.testing.assert_allclose(res["gradient"][0], de(d) * u, rtol=1e-7, atol=1e-15)
    np.testing.assert_allclose(res["gradient"][1], -de(d) * u, rtol=1e-7, atol=1e-15)


def test_atm_equilateral_closed_form():
    d = 5.5
    pos = np.array([[0.0, 0.0, 0.0], [d, 0.0, 0.0], [d / 2, d * np.sqrt(3) / 2, 0.0]])
    p = dict(s6=0.0, s8=0.0, a1=0.4, a2=5.0, s9=1.0, alp=14.0)
    res = O.dispersion([2, 2, 2], pos, param=p)
    c6 = _c6_single_ref(2, 2)
    r0 = (4.0 / 3.0 * O.ORACLE_AATOAU * TABLES["vdwrad_aa"][2 + 2 * 1 // 2 - 1]) ** 3  # pair (2, 2)
    fdmp = 1 / (1 + 6 * (r0 / d ** 3) ** (16.0 / 3.0))
    expect = c6 ** 1.5 * (1 + 3 * 0.5 ** 3) / d ** 9 * fdmp
    assert res["energy"] == pytest.approx(expect, rel=1e-13)
    np.testing.assert_allclose(res["gradient"].sum(axis=0), 0.0, atol=1e-17)


def _small_molecule(seed=7, n=14):
    rng = np.random.default_rng(seed)
    num = rng.choice([1, 6, 7, 8, 16], size=n)
    pos = []
    while len(pos) < n:
        x = rng.uniform(0, 9.0, 3)
        if all(np.linalg.norm(x - y) > 2.0 for y in pos):
            pos.append(x)
    return num, np.array(pos)


PARAMS = {
    "rational": dict(s6=1.0, s8=1.2177, a1=0.4145, a2=4.8593, s9=1.0, alp=14.0),
    "zero": dict(s6=1.0, s8=1.703, rs6=1.261, rs8=1.0, s9=1.0, alp=14.0),
    # param.f90:1168-1169 (blyp), 1219-1220 (tpss), 1266-1267 (pbe0 with a nonzero sigmoid), z as
    # test_d3z_mb01 of test/unit/test_param.f90
    "mzero": dict(s6=1.0, s8=1.841686, rs6=1.279637, rs8=1.0, bet=0.014370, s9=1.0, alp=14.0),
    "optimizedpower": dict(s6=1.0, s8=0.51581, a1=0.575, a2=3.0, bet=8.0, s9=1.0, alp=14.0),
    "cso": dict(s6=1.0, a1=0.2, a2=2.5, a3=0.0, a4=6.25, s9=1.0, alp=14.0),
    "z": dict(s6=1.0, s8=1.0, a1=200770.0, s9=1.0, alp=14.0),
}
ALL_KINDS = list(PARAMS)


@pytest.mark.parametrize("kind", ALL_KINDS)
def test_gradient_finite_difference_molecule(kind):
    num, pos = _small_molecule()
    p = PARAMS[kind]
    res = O.dispersion(num, pos, kind=_kind(kind), param=p)
    h = 1e-4
    fd = np.zeros_like(pos)
    for i in range(len(pos)):
        for c in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, c] += h
            pm[i, c] -= h
            fd[i, c] = (O.dispersion(num, pp, kind=_kind(kind), param=p, grad=False)["energy"]
                        - O.dispersion(num, pm, kind=_kind(kind), param=p, grad=False)["energy"]) / (2 * h)
    np.testing.assert_allclose(res["gradient"], fd, rtol=0, atol=2e-9)
    # translation invariance and the virial theorem of a molecule
    np.testing.assert_allclose(res["gradient"].sum(axis=0), 0.0, atol=1e-14)
    virial = np.einsum("ia,ib->ab", res["gradient"], pos)  # sum_i g_ia x_ib = sigma(a,b)
    np.testing.assert_allclose(res["virial"].T, virial, atol=1e-13)


def test_rotation_invariance():
    num, pos = _small_molecule(11)
    q, _ = np.linalg.qr(np.random.default_rng(3).normal(size=(3, 3)))
    a = O.dispersion(num, pos, param=PARAMS["rational"])
    b = O.dispersion(num, pos @ q.T, param=PARAMS["rational"])
    assert b["energy"] == pytest.approx(a["energy"], rel=1e-13)
    np.testing.assert_allclose(b["gradient"], a["gradient"] @ q.T, atol=1e-13)


def _small_cell(seed=5):
    rng = np.random.default_rng(seed)
    lat = np.diag([9.5, 10.2, 8.8]) + rng.uniform(-0.6, 0.6, (3, 3))
    num = rng.choice([1, 6, 8, 14], size=8)
    frac = []
    while len(frac) < 8:
        f = rng.uniform(0, 1, 3)
        ok = True
        for g in frac:
            d = f - g
            d -= np.round(d)
            if np.linalg.norm(d @ lat) < 2.2:
                ok = False
        if ok:
            frac.append(f)
    return num, np.array(frac) @ lat, lat


@pytest.mark.parametrize("kind", ALL_KINDS)
def test_periodic_gradient_and_strain_virial(kind):
    num, pos, lat = _small_cell()
    p = PARAMS[kind]
    cut = (20.0, 25.0, 12.0, 0.0, 0.0)  # keeps the FD affordable, same code path
    res = O.dispersion(num, pos, lat, [True] * 3, _kind(kind), p, cut)
    e = lambda x, l: O.dispersion(num, x, l, [True] * 3, _kind(kind), p, cut, grad=False)["energy"]
    h = 1e-4
    for i, c in [(0, 0), (3, 1), (7, 2)]:
        pp, pm = pos.copy(), pos.copy()
        pp[i, c] += h
        pm[i, c] -= h
        assert res["gradient"][i, c] == pytest.approx((e(pp, lat) - e(pm, lat)) / (2 * h), abs=2e-9)
    # sigma_ab = dE/d eps_ab with positions and lattice strained by (1 + eps)
    # (test_periodic_3d.f90:229-262)
    sig = np.zeros((3, 3))
    he = 1e-5
    for a in range(3):
        for b in range(3):
            ep = np.eye(3)
            ep[a, b] += he
            em = np.eye(3)
            em[a, b] -= he
            sig[a, b] = (e(pos @ ep.T, lat @ ep.T) - e(pos @ em.T, lat @ em.T)) / (2 * he)
    np.testing.assert_allclose(res["virial"].T, sig, atol=5e-9)


def test_lattice_points_cover_cutoff_sphere():
    rng = np.random.default_rng(0)
    for trial in range(5):
        lat = np.diag(rng.uniform(5, 12, 3)) + rng.uniform(-2, 2, (3, 3))
        rthr = rng.uniform(8, 30)
        tr = O.lattice_points([1, 1, 1], lat, rthr)
        n = np.array(np.meshgrid(*[np.arange(-15, 16)] * 3, indexing="ij")).reshape(3, -1).T
        allt = n @ lat
        inside = allt[np.linalg.norm(allt, axis=1) <= rthr]
        have = {tuple(np.round(t, 9)) for t in tr}
        assert all(tuple(np.round(t, 9)) in have for t in inside)
        # symmetric and ordered ix outer, iz inner: the first image is -rep
        np.testing.assert_array_equal(tr[::-1], -tr)


def test_lattice_points_molecule_and_2d():
    assert O.lattice_points([0, 0, 0], np.eye(3), 40.0).tolist() == [[0.0, 0.0, 0.0]]
    lat = np.array([[4.7, 0.0, 0.0], [-2.3, 4.0, 0.0], [0.0, 0.0, 1.0]])
    tr = O.lattice_points([1, 1, 0], lat, 30.0)
    assert np.all(tr[:, 2] == 0.0)
    n = np.array(np.meshgrid(np.arange(-20, 21), np.arange(-20, 21), indexing="ij")).reshape(2, -1).T
    allt = n @ lat[:2]
    inside = allt[np.linalg.norm(allt, axis=1) <= 30.0]
    have = {tuple(np.round(t, 9)) for t in tr}
    assert all(tuple(np.round(t, 9)) in have for t in inside)


def test_lattice_points_1d():
    """One periodic direction (cutoff.f90:157-163): rep = ceiling(r / |a_k|)
    along the periodic lattice vector, 0 along the others; the images are
    the multiples n a_k, |n| <= rep, in increasing n, and cover every such
    translation within the cutoff."""
    lat = np.array([[4.3, 0.7, -0.2], [0.5, 11.0, 0.3], [-0.4, 0.2, 13.0]])
    for axis in range(3):
        pbc = [k == axis for k in range(3)]
        for rthr in (9.0, 40.0, 60.0):
            tr = O.lattice_points(pbc, lat, rthr)
            rep = int(np.ceil(rthr / np.linalg.norm(lat[axis])))
            expect = np.outer(np.arange(-rep, rep + 1), lat[axis])
            np.testing.assert_allclose(tr, expect, rtol=0, atol=1e-12)
            n = np.arange(-200, 201)
            inside = [k for k in n if np.linalg.norm(k * lat[axis]) <= rthr]
            assert min(inside) >= -rep and max(inside) <= rep


def test_1d_chain_equals_3d_with_vacuum():
    """A system periodic along one lattice vector equals the 3D-periodic
    system whose other two lattice vectors are longer than every cutoff plus
    the system's extent (the 3D enumeration is pinned by the Ta90 golden);
    BJ and zero damping with ATM, energy, gradient and virial."""
    from helpers import BJ, ZERO
    rng = np.random.default_rng(5)
    a = np.array([7.5, 0.4, -0.3])
    num = rng.choice([1, 6, 7, 8], size=8)
    pos = np.column_stack([rng.uniform(0, 7.5, 8), rng.uniform(0, 5, 8), rng.uniform(0, 5, 8)])
    lat1 = np.array([a, [0.0, 150.0, 0.0], [0.0, 0.0, 150.0]])
    for kind, par in ((O.RATIONAL, BJ), (O.ZERO, ZERO)):
        r1 = O.dispersion(num, pos, lat1, [True, False, False], kind, par)
        r3 = O.dispersion(num, pos, lat1, [True, True, True], kind, par)
        assert abs(r1["energy"] - r3["energy"]) <= 1e-13 * abs(r3["energy"])
        np.testing.assert_allclose(r1["gradient"], r3["gradient"], rtol=0, atol=1e-14)
        np.testing.assert_allclose(r1["virial"], r3["virial"], rtol=0, atol=1e-13)
        # and differs from the molecule (the images along a interact)
        r0 = O.dispersion(num, pos, None, None, kind, par, grad=False)
        assert abs(r1["energy"] - r0["energy"]) > 1e-4


def test_cso_sigmoid_against_published_form():
    """CSO damping (cso.f90:300-320; Schroeder, Creon, Schwabe, JCTC 2015):
    E = -C6 (s6 + a1 sf(r)) / (r^6 + (a3 r0 + a4)^6) with the sigmoid
    sf(r) = 1 / (1 + exp(r - a2 r0)) and r0 = sqrt(3 r4r2_i r4r2_j).  The
    rational damping with s8 = 0, a1 = a3, a2 = a4 is -C6 s6 / (r^6 + (a3 r0
    + a4)^6) with the same C6 (pinned by the goldens and its dimer closed
    form), so a dimer's sf(r) = (E_cso / E_rat - s6) / a1 follows from two
    oracle energies; it must be the published sigmoid with r0 from the r4r2
    table: s6 + a1 at short range, s6 alone at long range, a1 / 2 added at
    r = a2 r0."""
    s6, a1, a2, a3, a4 = 1.0, 0.86, 2.5, 0.0, 6.25
    r0 = np.sqrt(3.0 * _r4r2(6) * _r4r2(8))
    for r in (0.8, a2 * r0 - 1.5, a2 * r0, a2 * r0 + 2.0, 30.0, 55.0):
        pos = [[0.0, 0.0, 0.0], [0.0, 0.3 * r, np.sqrt(0.91) * r]]
        e_cso = O.dispersion([6, 8], pos, kind=O.CSO, param=dict(s6=s6, a1=a1, a2=a2, a3=a3, a4=a4, s9=0.0),
                             grad=False)["energy"]
        e_rat = O.dispersion([6, 8], pos, kind=O.RATIONAL, param=dict(s6=1.0, s8=0.0, a1=a3, a2=a4, s9=0.0),
                             grad=False)["energy"]
        sf = (e_cso / e_rat - s6) / a1
        assert abs(sf - 1.0 / (1.0 + np.exp(r - a2 * r0))) <= 1e-12, (r, sf)
        if r < 1.0:  # short range: s6 + a1
            assert sf > 1.0 - 1e-4
        if r > 50.0:  # long range: s6 alone
            assert abs(sf) <= 1e-15


def test_wrap_to_central_cell():
    lat = np.array([[10.0, 0.0, 0.0], [1.0, 9.0, 0.0], [0.5, 0.5, 8.0]])
    pos = np.array([[-3.0, 2.0, 1.0], [25.0, -7.0, 19.0]])
    w = O.wrap(pos, lat, [1, 1, 1])
    frac = w @ np.linalg.inv(lat)
    assert np.all(frac >= -1e-12) and np.all(frac < 1.0)
    shift = (w - pos) @ np.linalg.inv(lat)
    np.testing.assert_allclose(shift, np.round(shift), atol=1e-12)


def test_too_close():
    assert O.too_close([[0, 0, 0], [0, 0, 1e-10]])
    assert not O.too_close([[0, 0, 0], [0, 0, 1e-8]])


def test_lattice_adjoint_identity_on_the_oracle():
    """dftd3.autograd.lattice_adjoint turns the oracle's gradient and virial
    into dE/dL (Cartesian positions fixed); central differences of the
    oracle energy in the lattice (h = 1e-5 bohr) agree."""
    import torch
    from dftd3.autograd import lattice_adjoint
    from helpers import BJ, cell
    for k, (n, edge) in enumerate([(4, 11.0), (5, 12.0)]):
        num, x, lat = cell(31 + k, n, edge)
        r = O.dispersion(num, x, lat, [True] * 3, O.RATIONAL, BJ, O.DEFAULT_CUTOFF)
        sigma_lib = torch.as_tensor(r["virial"].T.copy()).reshape(1, 3, 3).transpose(1, 2)  # [b][a] layout
        d = lattice_adjoint(torch.as_tensor(r["gradient"]), sigma_lib, torch.as_tensor(x),
                            torch.as_tensor(lat).reshape(1, 3, 3), torch.zeros(n, dtype=torch.long))[0].numpy()
        h = 1e-5
        fd = np.zeros((3, 3))
        for a in range(3):
            for b in range(3):
                e = []
                for sgn in (1.0, -1.0):
                    lt = lat.copy()
                    lt[a, b] += sgn * h
                    e.append(O.dispersion(num, x, lt, [True] * 3, O.RATIONAL, BJ, O.DEFAULT_CUTOFF,
                                          grad=False)["energy"])
                fd[a, b] = (e[0] - e[1]) / (2 * h)
        assert np.abs(d - fd).max() <= 1e-9 * np.abs(fd).max() + 1e-12


# ------------------------------------------------------------------ pairwise
PAIRWISE = ["validation_11-pairwise-d3-bj", "validation_12-pairwise-d3-zero", "validation_13-pairwise-d3-bjm",
            "python_pair_resolved"]


@pytest.mark.parametrize("name", PAIRWISE)
def test_pairwise_goldens(pairwise_goldens, name):
    """Pair energy matrices of the CLI validation (abs 1e-6) and of the
    Python interface test (abs 1e-8); the total energy of the CLI cases."""
    f = pairwise_goldens[name]
    e2, e3 = O.pairwise(f["numbers"], f["positions"], f["lattice"], f["periodic"], _kind(f["kind"]), f["param"],
                        f["cutoff"])
    np.testing.assert_allclose(e2, np.array(f["expected"]["pair2"]), rtol=0, atol=f["atol"])
    np.testing.assert_allclose(e3, np.array(f["expected"]["pair3"]), rtol=0, atol=f["atol"])
    if "energy" in f["expected"]:
        assert abs(e2.sum() + e3.sum() - f["expected"]["energy"]) <= f["atol"]


def test_pairwise_ta90_sum(goldens):
    """The pair energies of the Ta90 cell (3D, 729 images of the 60-bohr
    cutoff, self-image pairs on the diagonal) sum to the 100-epsilon energy."""
    f = goldens["ta90_3d_bj"]
    e2, e3 = O.pairwise(f["numbers"], f["positions"], f["lattice"], f["periodic"], _kind(f["kind"]), f["param"],
                        f["cutoff"])
    assert abs(e2.sum() + e3.sum() - f["expected"]["energy"]) <= f["atol"]
    assert not e3.any()


@pytest.mark.parametrize("kind", ["rational", "zero"])
def test_pairwise_rows_are_atomic_energies(kind):
    """Row x of the pair matrices sums to the atomic energy of x of the energy
    branch (atm.f90:500-621, rational.f90:282-374): each two-body term is
    split in halves between the pair's atoms and each triple term in thirds
    between its atoms, so a wrong weight on the diagonal (self images), on
    repeated indices or on one of the three pairs of a triple fails here.
    Small periodic cells (self images, two atoms of the same index in a
    triple) with smooth cutoff widths, and a 2D cell."""
    from helpers import BJ, ZERO, cell
    par = BJ if kind == "rational" else ZERO
    cut = (40.0, 60.0, 40.0, 3.0, 4.0)
    cases = [cell(41, 4, 9.0) + ([True] * 3,), cell(42, 5, 10.0) + ([True, True, False],)]
    for num, x, lat, pbc in cases:
        e2, e3 = O.pairwise(num, x, lat, pbc, _kind(kind), par, cut)
        r = O.dispersion(num, x, lat, pbc, _kind(kind), par, cut, grad=False)
        p0 = dict(par, s9=0.0)
        r2 = O.dispersion(num, x, lat, pbc, _kind(kind), p0, cut, grad=False)
        scale = np.abs(r["energies"]).max()
        np.testing.assert_allclose(e2.sum(axis=1), r2["energies"], rtol=0, atol=1e-13 * scale)
        np.testing.assert_allclose(e3.sum(axis=1), r["energies"] - r2["energies"], rtol=0, atol=1e-13 * scale)
        np.testing.assert_array_equal(e2, e2.T)
        np.testing.assert_array_equal(e3, e3.T)
        assert np.abs(np.diag(e2)).min() > 0 and np.abs(np.diag(e3)).min() > 0  # self images present


def test_pairwise_ghost_atoms():
    """A ghost atom (model.f90:166-174) has no pair energy with anything."""
    from helpers import BJ
    from dftd3 import synthetic
    num, x = synthetic.molecule_c0(seed=5, n=12)
    ghost = np.zeros(12, dtype=np.int32)
    ghost[[2, 7]] = 1
    e2, e3 = O.pairwise(num, x, None, None, O.RATIONAL, BJ, O.DEFAULT_CUTOFF, ghost=ghost)
    for g in (2, 7):
        assert not e2[g].any() and not e2[:, g].any() and not e3[g].any() and not e3[:, g].any()
    r = O.dispersion(num, x, None, None, O.RATIONAL, BJ, O.DEFAULT_CUTOFF, grad=False, ghost=ghost)
    assert abs(e2.sum() + e3.sum() - r["energy"]) <= 1e-14 * abs(r["energy"])


# ------------------------------------------------------- damping families
FAMILY = ["validation_03-energy-d3-bjm", "validation_04-energy-d3-zerom", "validation_05-energy-d3-op",
          "validation_08-gradient-d3-bjm", "validation_09-gradient-d3-zerom", "validation_10-gradient-d3-op",
          "validation_14-pairwise-d3-zerom", "validation_15-pairwise-d3-op", "validation_16-db-d3-bj",
          "validation_17-db-d3-zero", "validation_18-db-d3-bjm", "validation_19-db-d3-zerom",
          "validation_20-db-d3-op"]


@pytest.mark.parametrize("name", FAMILY)
def test_family_goldens(family_goldens, name):
    """Modified rational, modified zero and optimized power damping (with ATM
    in 16-20) against the CLI validation set at its 1e-6; the cases written
    by v1.2.1 (16, 18-20) agree to rounding and are held to 1e-12."""
    f = family_goldens[name]
    e = f["expected"]
    atol = 1e-12 if f["version"] == "1.2.1" else f["atol"]
    r = _run(f)
    assert abs(r["energy"] - e["energy"]) <= atol
    if "gradient" in e:
        np.testing.assert_allclose(r["gradient"], np.array(e["gradient"]), rtol=0, atol=f["atol"])
        np.testing.assert_allclose(r["virial"], np.array(e["virial"]), rtol=0, atol=f["atol"])
    if "pair2" in e:
        e2, e3 = O.pairwise(f["numbers"], f["positions"], f["lattice"], f["periodic"], _kind(f["kind"]),
                            f["param"], f["cutoff"])
        np.testing.assert_allclose(e2, np.array(e["pair2"]), rtol=0, atol=f["atol"])
        np.testing.assert_allclose(e3, np.array(e["pair3"]), rtol=0, atol=f["atol"])


REDUCTIONS = [
    # modified zero with bet = 0 is the zero damping (interface.py:336-340)
    ("mzero", dict(PARAMS["mzero"], bet=0.0), "zero", dict(PARAMS["mzero"], bet=0.0)),
    # optimized power with bet = 0 is the rational damping (interface.py:386-388)
    ("optimizedpower", dict(PARAMS["optimizedpower"], bet=0.0), "rational", dict(PARAMS["optimizedpower"])),
    # CSO with a1 = 0: sig = s6 and t6 = 1/(r^6 + (a3 sqrt(rrij) + a4)^6), the rational
    # damping with s8 = 0, a1 = a3, a2 = a4
    ("cso", dict(PARAMS["cso"], a1=0.0, a3=0.3, a4=4.0), "rational",
     dict(s6=1.0, s8=0.0, a1=0.3, a2=4.0, s9=1.0, alp=14.0)),
    # Z with a1 = 0: no damping, the rational damping with a1 = a2 = 0
    ("z", dict(PARAMS["z"], a1=0.0), "rational", dict(s6=1.0, s8=1.0, a1=0.0, a2=0.0, s9=1.0, alp=14.0)),
]


@pytest.mark.parametrize("kind,par,ref_kind,ref_par", REDUCTIONS)
def test_family_reductions(kind, par, ref_kind, ref_par):
    """Special cases that reduce a family to a pinned one: energies,
    gradients and virials agree to rounding (molecule and periodic cell)."""
    num, pos = _small_molecule(5)
    cases = [(num, pos, None, None, O.DEFAULT_CUTOFF)]
    cnum, cpos, lat = _small_cell(9)
    cases.append((cnum, cpos, lat, [True] * 3, (20.0, 25.0, 12.0, 0.0, 0.0)))
    for n_, x, l, pbc, cut in cases:
        a = O.dispersion(n_, x, l, pbc, _kind(kind), par, cut)
        b = O.dispersion(n_, x, l, pbc, _kind(ref_kind), ref_par, cut)
        assert a["energy"] == pytest.approx(b["energy"], rel=1e-12)
        np.testing.assert_allclose(a["gradient"], b["gradient"], rtol=0, atol=1e-12 * np.abs(b["gradient"]).max())
        np.testing.assert_allclose(a["virial"], b["virial"], rtol=0, atol=1e-12 * np.abs(b["virial"]).max())


def test_z_damping_species_ids():
    """The Z damping's r0 = a1 / (id_i + id_j) uses the species identifiers of
    the structure (elements numbered by first appearance), not atomic
    numbers: reordering the atoms so that another element appears first
    changes the energy, and relabelling that keeps the first-appearance
    order does not."""
    num, pos = _small_molecule(13)
    p = PARAMS["z"]
    e = lambda n_, x: O.dispersion(n_, x, None, None, O.ZDAMP, p, O.DEFAULT_CUTOFF, grad=False)["energy"]
    base = e(num, pos)
    # same structure, atoms listed in reverse: the identifiers change unless the
    # first-appearance order of the elements is unchanged
    rev = e(num[::-1].copy(), pos[::-1].copy())
    first = list(dict.fromkeys(num.tolist()))
    first_rev = list(dict.fromkeys(num[::-1].tolist()))
    if first != first_rev:
        assert abs(rev - base) > 1e-10 * abs(base)
    # a permutation that keeps the first atom of every element in place keeps the ids
    perm = np.arange(len(num))
    last = len(num) - 1
    for k in range(len(num) - 2, 0, -1):
        if num[k] == num[last] and num[:k].tolist().count(num[k]) > 0:
            perm[[k, last]] = perm[[last, k]]
            break
    assert e(num[perm], pos[perm]) == pytest.approx(base, rel=1e-13)


# ------------------------------------------------------------------ Hessian
WATER_DIMER = (np.array([8, 1, 1, 8, 1, 1]),
               np.array([[-2.85, -0.05, +0.00], [-1.05, +0.00, +0.00], [-3.40, +1.65, +0.00],
                         [+2.65, +0.00, +0.00], [+3.30, +1.00, +1.45], [+3.30, +1.00, -1.45]]))
PBE0_ATM = dict(s6=1.0, s8=1.2177, a1=0.4145, a2=4.8593, s9=1.0, alp=14.0)  # param.f90:364-366, atm


def _fd_hessian(num, pos, lat, pbc, kind, par, cut, step, ghost=None):
    n = len(num)
    h = np.zeros((3 * n, 3 * n))
    for i in range(n):
        for c in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[i, c] += step
            pm[i, c] -= step
            gp = O.dispersion(num, pp, lat, pbc, kind, par, cut, ghost=ghost)["gradient"]
            gm = O.dispersion(num, pm, lat, pbc, kind, par, cut, ghost=ghost)["gradient"]
            h[3 * i + c] = (gp - gm).ravel() / (2 * step)
    return h


@pytest.mark.parametrize("ghost", [None, [2, 3]])
def test_hessian_python_interface_case(ghost):
    """python/dftd3/test_interface.py:359-442: the water dimer with PBE0-D3(BJ)
    + ATM (and with ghost atoms 2, 3): symmetric to 1e-12, equal to central
    differences of the analytic gradient (step 1e-6) within 1e-7, no NaN."""
    num, pos = WATER_DIMER
    mask = None
    if ghost is not None:
        mask = np.zeros(len(num), dtype=np.int32)
        mask[ghost] = 1
    r = O.hessian(num, pos, None, None, O.RATIONAL, PBE0_ATM, O.DEFAULT_CUTOFF, ghost=mask)
    h = r["hessian"]
    assert not np.isnan(h).any()
    np.testing.assert_allclose(h, h.T, rtol=0, atol=1e-12)
    fd = _fd_hessian(num, pos, None, None, O.RATIONAL, PBE0_ATM, O.DEFAULT_CUTOFF, 1e-6, ghost=mask)
    np.testing.assert_allclose(h, fd, rtol=0, atol=1e-7)
    e = O.dispersion(num, pos, None, None, O.RATIONAL, PBE0_ATM, O.DEFAULT_CUTOFF, grad=False, ghost=mask)["energy"]
    assert r["energy"] == pytest.approx(e, rel=1e-14)


@pytest.mark.parametrize("kind", ALL_KINDS)
def test_hessian_finite_difference_of_gradient(kind):
    """Analytic Hessian of every damping family (two-body, ATM, CN and the
    dCN/dR products) against central differences of the analytic gradient
    (h = 1e-4 bohr, truncation ~1e-10): a molecule, and for rational and zero
    damping a small periodic cell with self images and smooth cutoff widths."""
    from dftd3 import synthetic
    num, pos = synthetic.molecule_c0(seed=4, n=8)
    p = PARAMS[kind]
    h = O.hessian(num, pos, None, None, _kind(kind), p)["hessian"]
    np.testing.assert_allclose(h, h.T, rtol=0, atol=1e-15)
    fd = _fd_hessian(num, pos, None, None, _kind(kind), p, O.DEFAULT_CUTOFF, 1e-4)
    np.testing.assert_allclose(h, fd, rtol=0, atol=1e-6 * np.abs(h).max())
    if kind in ("rational", "zero"):
        from helpers import cell
        cnum, cpos, lat = cell(7, 3, 8.0)
        cut = (20.0, 25.0, 12.0, 2.0, 3.0)
        h = O.hessian(cnum, cpos, lat, [True] * 3, _kind(kind), p, cut)["hessian"]
        fd = _fd_hessian(cnum, cpos, lat, [True] * 3, _kind(kind), p, cut, 1e-4)
        np.testing.assert_allclose(h, fd, rtol=0, atol=1e-6 * np.abs(h).max())
