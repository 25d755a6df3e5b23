"""The benchmark workloads through their own generators (dftd3.synthetic) on
the GPU.

* 20,000-atom ellipsoid (protein-like H/C/N/O/S): a 1,500-atom ellipsoid of
  the same density and axis ratio against the oracle (FP64 and mixed), and
  the full system through properties the oracle cannot afford at that size:
  term counts equal those of the distance graphs (pairs within the cutoffs,
  triangles of the three-body graph), the gradient sums to zero
  (translation invariance), the energy is invariant under a rotation and
  the gradient rotates with it, the virial is symmetric and equals
  sum_i g_i x_i^T of a molecule, mixed agrees with FP64 within the mixed
  bars.
* 2,000-atom triclinic cell (H/C/N/O/Si/S): a 200-atom cell of the same
  density and strain against the oracle (FP64 and mixed, term counts), and
  the full cell: term counts, gradient sum, mixed against FP64.
* 4,096 molecules of 10-300 atoms, BJ and zero damping alternating: a
  64-molecule batch of the same generator (B, F, P, Cl included) through a
  captured and replayed plan against the oracle, FP64 and mixed.
* 256 periodic cells of 64-512 atoms through autograd: four cells of 64-128
  atoms; position and lattice adjoints against the oracle's analytic
  gradient and virial and against central finite differences
  (h = 1e-4 bohr, FP64) of the energy with smooth two- and three-body
  cutoffs and the CN cutoff in a gap of the pair distances.
* Repeat runs of the FP64 path: the three-body kernel adds FP64 partial sums
  with atomics, so two evaluations may differ in the last bits; they agree
  to 1e-14 of the energy and of the largest gradient element.

Bars as in test_gpu_parity: FP64 |dE| <= 1e-10 |E|, max |dg| <= 1e-9 Eh/bohr,
max |d sigma| <= 1e-9 Eh; mixed |dE| <= 1e-8 Eh per atom, max |dg| <= 1e-7
Eh/bohr.
"""
import functools
import sys

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

E_REL, G_ABS, S_ABS = 1e-10, 1e-9, 1e-9
MIX_E_PER_ATOM, MIX_G = 1e-8, 1e-7


def _step(msg):
    print(f"[gpu-configs] {msg}", flush=True)
    sys.stdout.flush()


def _check(res, ref, nat, mixed):
    e, er = float(res["energy"]), ref["energy"]
    if mixed:
        assert abs(e - er) <= MIX_E_PER_ATOM * nat, (e, er)
    else:
        assert abs(e - er) <= E_REL * abs(er), (e, er, abs(e - er) / abs(er))
    dg = np.abs(np.asarray(res["gradient"]) - ref["gradient"]).max()
    assert dg <= (MIX_G if mixed else G_ABS), dg
    ds = np.abs(np.asarray(res["virial"]) - ref["virial"]).max()
    assert ds <= (MIX_G * nat if mixed else S_ABS), ds
    return dg, ds


def _run(systems, precision, grad=True):
    from dftd3 import torch_api
    spec, pos = torch_api.batch_spec(systems)
    p = torch.as_tensor(pos, dtype=torch.float64, device="cuda").contiguous()
    out = torch_api.run_batch(spec, p, precision=precision, grad=grad)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def _counts(systems):
    from dftd3 import torch_api
    spec, pos = torch_api.batch_spec(systems)
    return torch_api.count_terms(spec, torch.as_tensor(pos, dtype=torch.float64, device="cuda").contiguous())


# ---------------------------------------------------------------- 20,000-atom ellipsoid

@functools.lru_cache(maxsize=None)
def _c2_small():
    from dftd3 import synthetic
    num, pos = synthetic.ellipsoid_c2(n=1500)
    return num, pos, O.dispersion(num, pos, param=synthetic.PBE0_BJ)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ellipsoid_1500_against_oracle(precision):
    from dftd3 import synthetic
    from dftd3.interface import DispersionModel, RationalDampingParam
    _step(f"ellipsoid 1,500 atoms (20,000-atom generator), {precision}")
    num, pos, ref = _c2_small()
    model = DispersionModel(num, pos)
    if precision == "mixed":
        model.set_precision("mixed")
    res = model.get_dispersion(RationalDampingParam(**synthetic.PBE0_BJ), grad=True)
    dg, ds = _check(res, ref, len(num), precision == "mixed")
    _step(f"  dE={abs(float(res['energy']) - ref['energy']):.2e} dg={dg:.2e} dsigma={ds:.2e}")
    if precision == "fp64":
        sysd = dict(numbers=num, positions=pos, param=synthetic.PBE0_BJ)
        assert tuple(_counts([sysd])[0]) == O.count_terms(pos)


@functools.lru_cache(maxsize=None)
def _c2_full():
    from dftd3 import synthetic
    num, pos = synthetic.ellipsoid_c2()
    sysd = dict(numbers=num, positions=pos, param=synthetic.PBE0_BJ)
    return sysd, _run([sysd], "fp64")


def _graph_counts(positions, cut_cn=40.0, cut2=60.0, cut3=40.0):
    """Term counts of a molecule from its distance graphs, independent of the
    library: pairs (i > j) with eps <= r^2 <= c^2 for the CN and two-body
    cutoffs, and triples as the triangles of the three-body graph,
    trace(A^3) / 6 by an FP64 matrix product (exact for these integers).
    r^2 is summed in the reference's order (x, y, z, no fused products)."""
    eps = np.finfo(float).eps
    x = torch.as_tensor(positions, dtype=torch.float64, device="cuda")
    n = x.shape[0]
    adj = torch.zeros(n, n, dtype=torch.float64, device="cuda")
    n_cn = n_2 = 0
    for i0 in range(0, n, 2048):
        d = x[i0:i0 + 2048, None, :] - x[None, :, :]
        r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
        lower = torch.arange(i0, i0 + r2.shape[0], device="cuda")[:, None] > torch.arange(n, device="cuda")[None, :]
        n_cn += int(((r2 >= eps) & (r2 <= cut_cn * cut_cn) & lower).sum())
        n_2 += int(((r2 >= eps) & (r2 <= cut2 * cut2) & lower).sum())
        adj[i0:i0 + r2.shape[0]] = ((r2 >= eps) & (r2 <= cut3 * cut3)).double()
        del d, r2, lower
    n_3 = int(round(float((adj * (adj @ adj)).sum()) / 6.0))
    return n_cn, n_2, n_3


def test_ellipsoid_20000_term_counts():
    _step("ellipsoid 20,000 atoms: term counts against the distance graphs")
    sysd, _ = _c2_full()
    gpu = tuple(int(c) for c in _counts([sysd])[0])
    _step(f"  CN {gpu[0]}, pairs {gpu[1]}, triples {gpu[2]}")
    assert gpu == _graph_counts(sysd["positions"])


def test_graph_counts_against_oracle():
    """The graph count above equals the oracle's enumeration on the
    1,500-atom ellipsoid."""
    from dftd3 import synthetic
    _step("distance-graph term counts against the oracle's enumeration")
    num, pos = synthetic.ellipsoid_c2(n=1500)
    assert _graph_counts(pos) == O.count_terms(pos)


def test_ellipsoid_20000_invariances():
    """FP64: forces sum to zero, the virial of a molecule is sum_i g_i x_i^T
    and symmetric, and a rotated copy has the same energy and the rotated
    gradient."""
    _step("ellipsoid 20,000 atoms: FP64 invariances")
    sysd, out = _c2_full()
    g = out["gradient"]
    gmax = np.abs(g).max()
    assert np.abs(g.sum(axis=0)).max() <= 1e-12 * len(g) * gmax
    sig = out["sigma"][0]
    gx = g.T @ sysd["positions"]  # sum_i g_ia x_ib
    assert np.abs(sig - sig.T).max() <= 1e-10
    assert np.abs(sig - gx.T).max() <= 1e-9 * max(1.0, np.abs(sig).max())
    rot = _rotation(3)
    centre = sysd["positions"].mean(axis=0)
    moved = dict(sysd, positions=(sysd["positions"] - centre) @ rot.T + centre)
    out_r = _run([moved], "fp64")
    e, er = out["energy"][0], out_r["energy"][0]
    assert abs(e - er) <= 1e-11 * abs(e), (e, er)
    assert np.abs(out_r["gradient"] - g @ rot.T).max() <= 1e-11


def test_ellipsoid_20000_mixed_against_fp64():
    _step("ellipsoid 20,000 atoms: mixed against FP64")
    sysd, ref = _c2_full()
    out = _run([sysd], "mixed")
    n = len(sysd["numbers"])
    de = abs(out["energy"][0] - ref["energy"][0]) / n
    dg = np.abs(out["gradient"] - ref["gradient"]).max()
    _step(f"  dE/atom={de:.2e} dg={dg:.2e}")
    assert de <= MIX_E_PER_ATOM and dg <= MIX_G
    assert np.abs(out["gradient"].sum(axis=0)).max() <= MIX_G


def _rotation(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    q = q @ np.diag(np.sign(np.diag(r)))
    return q if np.linalg.det(q) > 0 else -q


# ---------------------------------------------------------------- 2,000-atom cell

def _c1_small_geometry():
    from dftd3 import synthetic
    edge = 28.0 * (200 / 2000) ** (1.0 / 3.0)  # same density as the 2,000-atom cell
    return synthetic.periodic_c1(n=200, edge=edge)


@functools.lru_cache(maxsize=None)
def _c1_small():
    from dftd3 import synthetic
    num, pos, lat = _c1_small_geometry()
    # the interface wraps positions into the central cell, as the reference's does
    return num, pos, lat, O.dispersion(num, pos, lat, [True] * 3, param=synthetic.PBE0_BJ, wrap_cell=True)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cell_200_against_oracle(precision):
    from dftd3 import synthetic
    from dftd3.interface import DispersionModel, RationalDampingParam
    _step(f"triclinic cell 200 atoms (2,000-atom generator), {precision}")
    num, pos, lat, ref = _c1_small()
    model = DispersionModel(num, pos, lat, np.array([True, True, True]))
    if precision == "mixed":
        model.set_precision("mixed")
    res = model.get_dispersion(RationalDampingParam(**synthetic.PBE0_BJ), grad=True)
    dg, ds = _check(res, ref, len(num), precision == "mixed")
    _step(f"  dE={abs(float(res['energy']) - ref['energy']):.2e} dg={dg:.2e} dsigma={ds:.2e}")
    if precision == "fp64":
        sysd = dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, param=synthetic.PBE0_BJ)
        assert tuple(_counts([sysd])[0]) == O.count_terms(pos, lat, [True] * 3)


def test_cell_2000_full():
    """Full 2,000-atom cell: term counts against the oracle's enumeration,
    forces sum to zero, mixed against FP64 within the mixed bars."""
    from dftd3 import synthetic
    _step("triclinic cell 2,000 atoms: counts, gradient sum, mixed vs FP64")
    num, pos, lat = synthetic.periodic_c1()
    sysd = dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, param=synthetic.PBE0_BJ)
    gpu = tuple(int(c) for c in _counts([sysd])[0])
    assert gpu == O.count_terms(pos, lat, [True] * 3)
    ref = _run([sysd], "fp64")
    out = _run([sysd], "mixed")
    g = ref["gradient"]
    assert np.abs(g.sum(axis=0)).max() <= 1e-12 * len(g) * np.abs(g).max()
    assert np.abs(ref["sigma"][0] - ref["sigma"][0].T).max() <= 1e-10
    de = abs(out["energy"][0] - ref["energy"][0]) / len(num)
    dg = np.abs(out["gradient"] - g).max()
    _step(f"  dE/atom={de:.2e} dg={dg:.2e}")
    assert de <= MIX_E_PER_ATOM and dg <= MIX_G


# ---------------------------------------------------------------- 4,096 molecules

@functools.lru_cache(maxsize=None)
def _c3_subset():
    from dftd3 import synthetic
    mols = synthetic.batch_c3(nmol=64)
    systems, refs = [], []
    for num, pos, param, kind in mols:
        k = "rational" if kind == 0 else "zero"
        systems.append(dict(numbers=num, positions=pos, kind=k, param=param))
        refs.append(O.dispersion(num, pos, kind=O.RATIONAL if kind == 0 else O.ZERO, param=param))
    return systems, refs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_molecule_batch_plan_against_oracle(precision):
    from dftd3 import torch_api
    _step(f"64 molecules (4,096-molecule generator), captured plan, {precision}")
    systems, refs = _c3_subset()
    elements = set(np.concatenate([s["numbers"] for s in systems]).tolist())
    assert {5, 9, 15, 17} <= elements and {s["kind"] for s in systems} == {"rational", "zero"}
    spec, pos = torch_api.batch_spec(systems)
    p = torch.as_tensor(pos, dtype=torch.float64, device="cuda").contiguous()
    plan = torch_api.BatchPlan(spec, p, precision=precision)
    plan.autotune(reps=1)
    plan.capture()
    for rep in range(2):  # replays with eager work and a synchronisation between them
        out = plan.replay()
        torch.cuda.synchronize()
        energy = out["energy"].cpu().numpy()
        grad = out["gradient"].cpu().numpy()
        sigma = out["sigma"].cpu().numpy()
        off = 0
        for k, (s, ref) in enumerate(zip(systems, refs)):
            n = len(s["numbers"])
            res = dict(energy=energy[k], gradient=grad[off:off + n], virial=sigma[k])
            _check(res, ref, n, precision == "mixed")
            off += n
    assert not plan.overflowed()


# ---------------------------------------------------------------- 256 cells, autograd

@functools.lru_cache(maxsize=None)
def _c4_subset():
    from dftd3 import synthetic
    cells = synthetic.cells_c4(ncell=4, nmax=128)
    return [dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, param=synthetic.PBE0_BJ)
            for num, pos, lat in cells]


def test_cells_autograd_against_oracle():
    from dftd3 import torch_api
    from dftd3.autograd import dispersion_energy
    _step("4 cells (256-cell generator): autograd adjoints vs oracle gradient and virial")
    systems = _c4_subset()
    assert all(64 <= len(s["numbers"]) <= 128 for s in systems)
    spec, pos = torch_api.batch_spec(systems)
    p = torch.as_tensor(pos, dtype=torch.float64, device="cuda").requires_grad_(True)
    lat = torch.as_tensor(np.stack([s["lattice"] for s in systems]), dtype=torch.float64,
                          device="cuda").requires_grad_(True)
    e = dispersion_energy(spec, p, lat)
    gpos, glat = torch.autograd.grad(e.sum(), (p, lat))
    gpos, glat, p64, l64 = (t.detach().cpu().numpy() for t in (gpos, glat, p, lat))
    off = 0
    for k, s in enumerate(systems):
        n = len(s["numbers"])
        ref = O.dispersion(s["numbers"], s["positions"], s["lattice"], s["periodic"], param=s["param"])
        g = gpos[off:off + n]
        # the virial from the adjoints: sigma = G^T X + (dE/dL)^T L
        sig = g.T @ p64[off:off + n] + glat[k].T @ l64[k]
        res = dict(energy=float(e[k].detach()), gradient=g, virial=sig.T)
        dg, ds = _check(res, ref, n, False)
        _step(f"  cell {k}: {n} atoms, dg={dg:.2e} dsigma={ds:.2e}")
        off += n

    # float32 inputs run the mixed kernels; FP64 on the same (rounded) inputs
    # is their reference
    p32 = p.detach().float().requires_grad_(True)
    l32 = lat.detach().float()
    e32 = dispersion_energy(spec, p32, l32)
    g32, = torch.autograd.grad(e32.sum(), p32)
    p64r = p32.detach().double().requires_grad_(True)
    e64r = dispersion_energy(spec, p64r, l32.double())
    g64r, = torch.autograd.grad(e64r.sum(), p64r)
    natoms = torch.as_tensor(spec.natoms, dtype=torch.float64, device="cuda")
    assert (torch.abs(e32.double() - e64r.detach()) <= MIX_E_PER_ATOM * natoms).all()
    assert torch.abs(g32.double() - g64r).max().item() <= MIX_G


def _cn_cutoff_in_gap(positions, lattice, h, centre=40.0):
    """A coordination-number cutoff near `centre` that no pair distance
    (lattice images included) comes within the largest displacement a step h
    of a position or lattice component can cause: the counting function does
    not vanish at the cutoff (it tends to 1 / (1 + e^16)), so a pair that
    crosses it changes the energy by a step."""
    x = np.asarray(positions)
    lat = np.asarray(lattice)
    rep = np.ceil((centre + 3.0) * np.linalg.norm(np.linalg.inv(lat), axis=0)).astype(int)
    n = np.array(np.meshgrid(*[np.arange(-r, r + 1) for r in rep], indexing="ij")).reshape(3, -1).T
    d = np.linalg.norm(x[:, None, None, :] - x[None, :, None, :] - (n @ lat)[None, None, :, :], axis=-1).ravel()
    d = np.sort(d[(d > centre - 2.0) & (d < centre + 2.0)])
    # a step h of lattice component (b, c) moves image n by n_b h
    margin = 1.01 * np.abs(n).max() * h
    gaps = np.diff(d)
    k = int(np.argmax(gaps))
    assert gaps[k] > 2 * margin, "no gap in the pair distances near the cutoff"
    return 0.5 * (d[k] + d[k + 1])


def test_cells_autograd_finite_differences():
    """Central differences (h = 1e-4 bohr) of the FP64 energy in positions and
    lattice components against the autograd adjoints.  The energy must be
    differentiable across the step: the two- and three-body terms use smooth
    cutoffs (widths 8 and 6 bohr at 60 and 40 bohr) and the CN cutoff lies in
    a gap of the pair distances (a step of a lattice component moves every
    image distance, so hard cutoffs are crossed by many terms at once).  A
    coordinate whose step still changes the counted terms is replaced by the
    next one."""
    from dftd3 import torch_api
    from dftd3.autograd import dispersion_energy
    _step("4 cells: autograd adjoints vs central finite differences")
    h = 1e-4
    rng = np.random.default_rng(11)
    for k, s0 in enumerate(_c4_subset()[:2]):
        cut = dict(disp2=60.0, disp3=40.0, width2=8.0, width3=6.0,
                   cn=_cn_cutoff_in_gap(s0["positions"], s0["lattice"], h))
        s = dict(s0, cutoff=cut)
        spec, pos = torch_api.batch_spec([s])
        p = torch.as_tensor(pos, dtype=torch.float64, device="cuda").requires_grad_(True)
        lat = torch.as_tensor(s["lattice"][None], dtype=torch.float64, device="cuda").requires_grad_(True)
        e = dispersion_energy(spec, p, lat)
        gpos, glat = (t.cpu().numpy() for t in torch.autograd.grad(e.sum(), (p, lat)))

        def energy_and_counts(xyz, lattice):
            sp = dict(s, positions=xyz, lattice=lattice)
            return _run([sp], "fp64", grad=False)["energy"][0], tuple(_counts([sp])[0])

        base = tuple(_counts([s])[0])
        n = len(s["numbers"])
        checked = 0
        for idx in rng.permutation(3 * n):
            if checked == 8:
                break
            i, a = divmod(int(idx), 3)
            xp, xm = s["positions"].copy(), s["positions"].copy()
            xp[i, a] += h
            xm[i, a] -= h
            ep, cp = energy_and_counts(xp, s["lattice"])
            em, cm = energy_and_counts(xm, s["lattice"])
            if cp[0] != base[0] or cm[0] != base[0]:
                continue
            fd = (ep - em) / (2 * h)
            assert abs(fd - gpos[i, a]) <= 1e-8 + 1e-6 * abs(gpos[i, a]), (k, i, a, fd, gpos[i, a])
            checked += 1
        assert checked == 8
        for b in range(3):
            for c in range(3):
                lp, lm = s["lattice"].copy(), s["lattice"].copy()
                lp[b, c] += h
                lm[b, c] -= h
                ep, cp = energy_and_counts(s["positions"], lp)
                em, cm = energy_and_counts(s["positions"], lm)
                assert cp[0] == base[0] and cm[0] == base[0]  # no CN pair crosses its cutoff
                fd = (ep - em) / (2 * h)
                assert abs(fd - glat[0, b, c]) <= 1e-8 + 1e-6 * abs(glat[0, b, c]), (k, b, c, fd, glat[0, b, c])
        _step(f"  cell {k}: {n} atoms, 8 coordinates and 9 lattice components (CN cutoff {cut['cn']:.4f})")


# ---------------------------------------------------------------- repeat runs

@pytest.mark.parametrize("which", ["molecule", "cell"])
def test_fp64_repeat_runs(which):
    from dftd3 import synthetic
    _step(f"FP64 repeat runs, {which}")
    if which == "molecule":
        num, pos = synthetic.molecule_c0()
        sysd = dict(numbers=num, positions=pos, param=synthetic.PBE0_BJ)
    else:
        num, pos, lat = _c1_small_geometry()
        sysd = dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, param=synthetic.PBE0_BJ)
    a = _run([sysd], "fp64")
    for _ in range(3):
        b = _run([sysd], "fp64")
        assert abs(a["energy"][0] - b["energy"][0]) <= 1e-14 * abs(a["energy"][0])
        gmax = np.abs(a["gradient"]).max()
        assert np.abs(a["gradient"] - b["gradient"]).max() <= 1e-14 * gmax
        assert np.abs(a["sigma"] - b["sigma"]).max() <= 1e-14 * np.abs(a["sigma"]).max()
