"""Planned batches (include/d3hip.h, plans): repeated evaluations without host
synchronisation, every launch shape of the three-body kernel, HIP-graph
capture and replay, and the slot-capacity guard.

A plan evaluates the same kernels as d3hip_get_dispersion_batch, so its
results are compared with that entry point on the same inputs (same
arithmetic; only the order of the atomic additions differs) and, for one
case, with the oracle.
"""
import sys

import numpy as np
import pytest
import torch

import oracle as O
from helpers import BJ, ZERO, cell

pytestmark = pytest.mark.gpu


def _step(msg):
    print(f"[gpu-plan] {msg}", flush=True)
    sys.stdout.flush()


def _systems(seed=3, nmol=7, ncell=2):
    from dftd3 import synthetic
    systems = []
    for k in range(nmol):
        num, pos = synthetic.molecule_c0(seed=seed + k, n=12 + 23 * k)
        kind = "rational" if k % 2 == 0 else "zero"
        systems.append(dict(numbers=num, positions=pos, kind=kind, param=BJ if kind == "rational" else ZERO))
    for k in range(ncell):
        num, pos, lat = cell(seed + 50 + k, 10 + 4 * k, 10.5)
        systems.append(dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, param=BJ))
    return systems


def _spec(systems):
    from dftd3 import torch_api
    spec, pos = torch_api.batch_spec(systems)
    return spec, torch.as_tensor(pos, dtype=torch.float64, device="cuda").contiguous()


def _close(a, b, grad=True, rtol=1e-11, atol=1e-12):
    ea, eb = a["energy"].cpu().numpy(), b["energy"].cpu().numpy()
    np.testing.assert_allclose(ea, eb, rtol=rtol, atol=atol)
    if grad:
        np.testing.assert_allclose(a["gradient"].cpu().numpy(), b["gradient"].cpu().numpy(), rtol=0,
                                   atol=atol + rtol * float(np.abs(b["gradient"].cpu().numpy()).max()))
        np.testing.assert_allclose(a["sigma"].cpu().numpy(), b["sigma"].cpu().numpy(), rtol=0,
                                   atol=atol + rtol * float(np.abs(b["sigma"].cpu().numpy()).max()))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_plan_matches_batch_entry_point(precision):
    from dftd3 import torch_api
    _step(f"plan vs batch entry point, {precision}")
    spec, pos = _spec(_systems())
    ref = torch_api.run_batch(spec, pos, precision=precision)
    plan = torch_api.BatchPlan(spec, pos, precision=precision)
    out = plan.execute()
    _close(out, ref)
    assert not plan.overflowed()


def test_plan_fp64_against_oracle():
    from dftd3 import torch_api
    _step("plan fp64 vs oracle")
    systems = _systems(nmol=3, ncell=1)
    spec, pos = _spec(systems)
    out = torch_api.BatchPlan(spec, pos, precision="fp64").execute()
    off = 0
    for k, s in enumerate(systems):
        n = len(s["numbers"])
        kind = s.get("kind", "rational")
        ref = O.dispersion(s["numbers"], s["positions"], s.get("lattice"), s.get("periodic"),
                           O.RATIONAL if kind == "rational" else O.ZERO, s["param"], O.DEFAULT_CUTOFF)
        assert abs(out["energy"][k].item() - ref["energy"]) <= 1e-10 * abs(ref["energy"]) + 1e-14
        assert np.abs(out["gradient"][off:off + n].cpu().numpy() - ref["gradient"]).max() <= 1e-9
        assert np.abs(out["sigma"][k].cpu().numpy() - ref["virial"]).max() <= 1e-9
        off += n


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_every_launch_shape_gives_the_same_result(precision):
    from dftd3 import library, torch_api
    spec, pos = _spec(_systems(seed=11))
    ref = torch_api.run_batch(spec, pos, precision=precision)
    plan = torch_api.BatchPlan(spec, pos, precision=precision)
    for waves, bpc in library.plan_candidate_shapes():
        _step(f"shape waves={waves} blocks/CU={bpc} {precision}")
        plan.set_shape(waves, bpc)
        assert plan.shape == (waves, bpc)
        out = {k: torch.zeros_like(v) for k, v in plan.out.items()}
        _close(plan.execute(out=out), ref)


def test_initial_launch_shape_follows_the_work():
    """A new plan starts with the shape of a call without a plan: four
    wavefronts and one block per work item for small work (a 200-atom
    molecule, about 400 items), a persistent grid above 8,192 items (two
    wavefronts and 8 blocks per CU; one and 16 for mixed-precision lists
    with lattice images)."""
    from dftd3 import synthetic, torch_api
    _step("initial launch shapes")
    num, pos = synthetic.molecule_c0()
    small = [dict(numbers=num, positions=pos, param=BJ)]
    num, pos = synthetic.molecule_c0(seed=5, n=1300)
    molecule = [dict(numbers=num, positions=pos, param=BJ)]
    num, pos, lat = cell(41, 320, 28.0)
    crystal = [dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, param=BJ)]
    for systems, fp64, mixed in ((small, (4, 0), (4, 0)), (molecule, (2, 8), (2, 8)), (crystal, (2, 8), (1, 16))):
        spec, p = _spec(systems)
        for precision, want in (("fp64", fp64), ("mixed", mixed)):
            plan = torch_api.BatchPlan(spec, p, precision=precision)
            assert plan.shape == want, (len(systems[0]["numbers"]), precision, plan.shape)
            del plan


def test_autotune_keeps_a_candidate_and_results():
    from dftd3 import library, torch_api
    _step("autotune")
    spec, pos = _spec(_systems(seed=17))
    ref = torch_api.run_batch(spec, pos, precision="mixed")
    plan = torch_api.BatchPlan(spec, pos, precision="mixed")
    best, times = plan.autotune(reps=2)
    assert best in library.plan_candidate_shapes() and plan.shape == best
    assert len(times) == len(library.plan_candidate_shapes()) and all(t > 0 for t in times.values())
    _close(plan.execute(), ref)


@pytest.mark.parametrize("grad", [True, False])
def test_graph_replays_follow_new_coordinates(grad):
    """Capture once, then replay with new coordinates written into the
    plan's positions; eager work and synchronisations run between the
    replays, and every replay is checked."""
    from dftd3 import torch_api
    _step(f"graph capture and replays, grad={grad}")
    systems = _systems(seed=23)
    spec, pos = _spec(systems)
    plan = torch_api.BatchPlan(spec, pos, precision="mixed", grad=grad)
    plan.capture()
    gen = torch.Generator(device="cuda").manual_seed(5)
    for rep in range(4):
        new = pos + 0.05 * torch.randn(pos.shape, dtype=torch.float64, device="cuda", generator=gen)
        plan.positions.copy_(new)
        out = plan.replay()
        torch.cuda.synchronize()
        got = {k: v.clone() for k, v in out.items()}
        # eager work between the replays
        scratch = torch.randn(1 << 16, device="cuda").square().sum()
        eager = plan.execute(new.contiguous(), out={k: torch.zeros_like(v) for k, v in out.items()})
        ref = torch_api.run_batch(spec, new.contiguous(), precision="mixed", grad=grad)
        torch.cuda.synchronize()
        assert torch.isfinite(scratch)
        # the replay is the plan's own evaluation (same storage order)
        _close(got, eager, grad=grad)
        # the batch entry point orders the atoms of the new coordinates afresh,
        # which changes the FP32 rounding of single triples (~1 ulp)
        _close(got, ref, grad=grad, rtol=1e-7, atol=1e-10)
    assert not plan.overflowed()


def test_slot_capacity_guard():
    """A 2D slab planned with its atoms far apart along the non-periodic axis
    gets lists sized for that geometry; compressing the slab overflows them:
    the plan reports it, and the outputs of the system with a three-body
    term are NaN while a system without one in the same batch is exact."""
    from dftd3 import torch_api
    _step("slot capacity guard")
    num = np.array([6, 7, 8, 16, 1, 6])
    lat = np.diag([10.0, 10.5, 600.0])
    spread = np.array([[0.0, 0.0, 0.0], [3.0, 2.0, 90.0], [6.0, 7.0, 180.0], [1.0, 8.0, 270.0],
                       [8.0, 4.0, 360.0], [4.0, 5.0, 450.0]])
    packed = spread.copy()
    packed[:, 2] = 2.5 * np.arange(6)
    sysd = dict(numbers=num, positions=spread, lattice=lat, periodic=[True, True, False], param=BJ)
    # second system of the batch: the same slab without the three-body term
    sys2 = dict(sysd, param=dict(BJ, s9=0.0))
    spec, pos = _spec([sysd, sys2])
    plan = torch_api.BatchPlan(spec, pos, precision="fp64")
    out = plan.execute()
    torch.cuda.synchronize()
    assert not plan.overflowed()
    assert torch.isfinite(out["energy"]).all() and torch.isfinite(out["gradient"]).all()
    packed_pos = np.concatenate([packed, packed])
    packed_t = torch.as_tensor(packed_pos, dtype=torch.float64, device="cuda").contiguous()
    out = plan.execute(packed_t)
    torch.cuda.synchronize()
    assert plan.overflowed()
    # the three-body system: every output NaN
    assert torch.isnan(out["energy"][0]).item()
    assert torch.isnan(out["gradient"][:6]).all().item()
    assert torch.isnan(out["sigma"][0]).all().item()
    # the two-body-only system: exact
    ref = torch_api.run_batch(spec, packed_t, precision="fp64")
    assert torch.isfinite(ref["energy"]).all()
    e2 = ref["energy"][1].item()
    assert abs(out["energy"][1].item() - e2) <= 1e-12 * abs(e2)
    assert torch.allclose(out["gradient"][6:], ref["gradient"][6:], rtol=0, atol=1e-13)
    # a later evaluation that fits is complete again; the flag stays raised
    out = plan.execute(torch.as_tensor(np.concatenate([spread, spread]), dtype=torch.float64,
                                       device="cuda").contiguous())
    torch.cuda.synchronize()
    assert torch.isfinite(out["energy"]).all() and plan.overflowed()
