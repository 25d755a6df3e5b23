"""Differentiable batches (dftd3.autograd): energies from the forward pass,
position and lattice adjoints from the backward pass.

Checks: energies and position adjoints against the oracle's energy and
analytic gradient; lattice adjoints against central finite differences of
the energy (h = 1e-5 bohr, FP64, cutoffs placed in gaps of the pair
distances) with the Cartesian positions held fixed; the
virial rebuilt from the adjoints (G^T X + (dE/dL)^T L) against the oracle's
virial; float32 inputs (mixed precision) against the FP64 results within the
mixed-precision bars.
"""
import sys

import numpy as np
import pytest
import torch

import oracle as O
from helpers import BJ, ZERO, cell

pytestmark = pytest.mark.gpu


def _step(msg):
    print(f"[gpu-autograd] {msg}", flush=True)
    sys.stdout.flush()


def _cells(seed=7):
    systems = []
    for k, (n, edge) in enumerate([(12, 10.0), (20, 12.5), (9, 9.0)]):
        num, pos, lat = cell(seed + k, n, edge)
        kind = "rational" if k % 2 == 0 else "zero"
        systems.append(dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, kind=kind,
                            param=BJ if kind == "rational" else ZERO))
    return systems


def _oracle(s):
    return O.dispersion(s["numbers"], s["positions"], s["lattice"], s["periodic"],
                        O.RATIONAL if s["kind"] == "rational" else O.ZERO, s["param"], O.DEFAULT_CUTOFF)


def _inputs(systems, dtype):
    from dftd3 import torch_api
    spec, pos = torch_api.batch_spec(systems)
    p = torch.as_tensor(pos, dtype=dtype, device="cuda").requires_grad_(True)
    lat = torch.as_tensor(np.stack([s["lattice"] for s in systems]), dtype=dtype, device="cuda")
    return spec, p, lat.requires_grad_(True)


def test_energy_and_position_adjoint_against_oracle():
    from dftd3.autograd import dispersion_energy
    _step("fp64 energies and position adjoints vs oracle")
    systems = _cells()
    spec, pos, lat = _inputs(systems, torch.float64)
    e = dispersion_energy(spec, pos, lat)
    w = torch.tensor([1.0, -0.5, 2.0], dtype=torch.float64, device="cuda")  # non-trivial upstream gradient
    gpos, = torch.autograd.grad((e * w).sum(), pos, retain_graph=True)
    off = 0
    for k, s in enumerate(systems):
        ref = _oracle(s)
        n = len(s["numbers"])
        assert abs(e[k].item() - ref["energy"]) <= 1e-10 * abs(ref["energy"]) + 1e-14
        np.testing.assert_allclose(gpos[off:off + n].cpu().numpy(), w[k].item() * ref["gradient"], rtol=0, atol=1e-9)
        off += n


def _gap_cutoff(systems, target, margin=2e-4):
    """A cutoff near `target` with no image-pair distance of any system within
    `margin`: the hard cutoffs make the energy jump when a pair crosses them,
    and a lattice step h moves an image pair by up to |n| h."""
    d = []
    for s in systems:
        x = np.asarray(s["positions"], dtype=float)
        tr = O.lattice_points(s["periodic"], s["lattice"], target + 1.0)
        v = x[:, None, None, :] + tr[None, None, :, :] - x[None, :, None, :]
        d.append(np.sqrt((v ** 2).sum(-1)).ravel())
    d = np.sort(np.concatenate(d))
    for step in range(2000):
        for c in (target - step * 1e-3, target + step * 1e-3):
            k = np.searchsorted(d, c)
            lo = d[k - 1] if k > 0 else -np.inf
            hi = d[k] if k < len(d) else np.inf
            if c - lo > margin and hi - c > margin:
                return c
    raise AssertionError("no gap in the pair distances")


def test_lattice_adjoint_against_finite_differences():
    from dftd3 import torch_api
    from dftd3.autograd import dispersion_energy
    _step("fp64 lattice adjoint vs central differences")
    systems = []
    for k, (n, edge) in enumerate([(4, 11.0), (5, 12.0)]):
        num, xyz, lt = cell(31 + k, n, edge)
        systems.append(dict(numbers=num, positions=xyz, lattice=lt, periodic=[True] * 3, kind="rational",
                            param=BJ))
    cut = dict(cn=_gap_cutoff(systems, 40.0), disp2=_gap_cutoff(systems, 60.0), disp3=_gap_cutoff(systems, 40.0))
    for s in systems:
        s["cutoff"] = cut
    spec, pos, lat = _inputs(systems, torch.float64)
    e = dispersion_energy(spec, pos, lat)
    glat, = torch.autograd.grad(e.sum(), lat)
    h = 1e-5  # moves an image pair by at most about 1e-4 bohr, inside the cutoff gaps
    fd = np.zeros((len(systems), 3, 3))
    base = lat.detach().cpu().numpy()
    p64 = pos.detach().contiguous()
    for k in range(len(systems)):
        for a in range(3):
            for b in range(3):
                vals = []
                for sgn in (1.0, -1.0):
                    lt = base.copy()
                    lt[k, a, b] += sgn * h
                    sp = torch_api.BatchSpec(spec.natoms, spec.numbers, lt, spec.periodic, spec.params, spec.cutoffs)
                    vals.append(torch_api.run_batch(sp, p64, precision="fp64", grad=False)["energy"][k].item())
                fd[k, a, b] = (vals[0] - vals[1]) / (2 * h)
    g = glat.cpu().numpy()
    scale = np.abs(fd).max()
    np.testing.assert_allclose(g, fd, rtol=0, atol=1e-6 * scale + 1e-10)


def test_virial_rebuilt_from_adjoints():
    from dftd3.autograd import forces_and_stress
    _step("virial from adjoints vs oracle virial")
    systems = _cells(seed=43)
    spec, pos, lat = _inputs(systems, torch.float64)
    energy, forces, stress = forces_and_stress(spec, pos, lat)
    off = 0
    for k, s in enumerate(systems):
        ref = _oracle(s)
        n = len(s["numbers"])
        vol = abs(np.linalg.det(s["lattice"]))
        np.testing.assert_allclose(-forces[off:off + n].cpu().numpy(), ref["gradient"], rtol=0, atol=1e-9)
        np.testing.assert_allclose(stress[k].cpu().numpy() * vol, ref["virial"], rtol=0, atol=1e-9)
        off += n


def test_float32_inputs_run_the_mixed_path():
    from dftd3.autograd import dispersion_energy
    _step("float32 inputs, mixed precision")
    systems = _cells(seed=57)
    spec, pos, lat = _inputs(systems, torch.float32)
    e = dispersion_energy(spec, pos, lat)
    assert e.dtype == torch.float32
    gpos, glat = torch.autograd.grad(e.sum(), (pos, lat))
    assert gpos.dtype == torch.float32 and glat.dtype == torch.float32
    spec64, pos64, lat64 = _inputs(systems, torch.float64)
    # same (float32-rounded) coordinates in FP64
    with torch.no_grad():
        pos64.copy_(pos.double())
        lat64.copy_(lat.double())
    e64 = dispersion_energy(spec64, pos64, lat64)
    g64, l64 = torch.autograd.grad(e64.sum(), (pos64, lat64))
    nat = torch.as_tensor(spec.natoms, dtype=torch.float64, device="cuda")
    assert ((e.double() - e64).abs() / nat).max().item() <= 1e-8
    assert (gpos.double() - g64).abs().max().item() <= 1e-7
    assert (glat.double() - l64).abs().max().item() <= 1e-6 * max(1.0, l64.abs().max().item())


def test_molecules_without_lattice():
    from dftd3 import synthetic, torch_api
    from dftd3.autograd import dispersion_energy
    _step("molecules, positions only")
    num, xyz = synthetic.molecule_c0(seed=3, n=40)
    spec, pos = torch_api.batch_spec([dict(numbers=num, positions=xyz, param=BJ)])
    p = torch.as_tensor(pos, dtype=torch.float64, device="cuda").requires_grad_(True)
    e = dispersion_energy(spec, p)
    g, = torch.autograd.grad(e.sum(), p)
    ref = O.dispersion(num, xyz, None, None, O.RATIONAL, BJ, O.DEFAULT_CUTOFF)
    assert abs(e.item() - ref["energy"]) <= 1e-10 * abs(ref["energy"])
    np.testing.assert_allclose(g.cpu().numpy(), ref["gradient"], rtol=0, atol=1e-9)


def test_one_evaluation_per_step_and_lattice_cache(monkeypatch):
    """With inputs that require gradients the forward pass evaluates energy,
    gradient and virial once and the backward pass evaluates nothing; without
    them the forward pass runs the energy-only kernels.  A lattice tensor
    changed in place is read again; the same tensor unchanged is not."""
    from dftd3 import autograd, torch_api
    _step("one evaluation per step, lattice cache")
    calls = []
    orig = torch_api.run_batch

    def counting(*a, **k):
        calls.append(k.get("grad", True))
        return orig(*a, **k)

    monkeypatch.setattr(torch_api, "run_batch", counting)
    systems = _cells(seed=61)
    spec, pos, lat = _inputs(systems, torch.float64)
    e = autograd.dispersion_energy(spec, pos, lat)
    gpos, glat = torch.autograd.grad(e.sum(), (pos, lat))
    assert calls == [True]
    with torch.no_grad():
        e_plain = autograd.dispersion_energy(spec, pos.detach(), lat.detach())
    assert calls == [True, False]
    torch.testing.assert_close(e_plain, e.detach(), rtol=1e-14, atol=0)
    # the cached host copy follows in-place changes of the lattice
    lat2 = lat.detach().clone()
    e1 = autograd.dispersion_energy(spec, pos.detach(), lat2)
    with torch.no_grad():
        lat2.mul_(1.02)
    e2 = autograd.dispersion_energy(spec, pos.detach(), lat2)
    e_ref = autograd.dispersion_energy(spec, pos.detach(), lat2.clone())
    assert (e1 - e2).abs().max().item() > 1e-6
    torch.testing.assert_close(e2, e_ref, rtol=1e-14, atol=0)
    # gradients of the single evaluation against the oracle
    off = 0
    for k, s in enumerate(systems):
        n = len(s["numbers"])
        ref = _oracle(s)
        np.testing.assert_allclose(gpos[off:off + n].cpu().numpy(), ref["gradient"], rtol=0, atol=1e-9)
        off += n
