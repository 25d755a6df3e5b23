"""Allocation-free entry point (include/d3hip.h,
d3hip_get_dispersion_batch_noalloc) through the C ABI, against
d3hip_get_dispersion_batch on the same inputs.

* Two calls in a row, the second with a smaller batch (the work area keeps
  the first call's contents), FP64 and mixed.  The same kernels run on both
  paths, so the results agree to rounding (the three-body kernel's FP64
  atomics add in a varying order): |dE| <= 1e-12 |E|, gradient and virial
  within 1e-12 of their largest element.
* The device's free memory is the same before and after a call.
* A batch larger than the work area is rejected with the library's message
  before any output is written; the next call works.
"""
import sys

import numpy as np
import pytest
import torch

from helpers import BJ, ZERO, cell

pytestmark = pytest.mark.gpu


def _step(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)
    sys.stdout.flush()


def _systems(sizes, seed, with_cell=True):
    from dftd3 import synthetic
    out = []
    for k, n in enumerate(sizes):
        num, pos = synthetic.molecule_c0(seed=seed + k, n=n)
        out.append(dict(numbers=num, positions=pos, kind="rational" if k % 2 == 0 else "zero",
                        param=BJ if k % 2 == 0 else ZERO))
    if with_cell:
        num, pos, lat = cell(seed, 12, 10.0)
        out.append(dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, param=BJ))
    return out


def _outputs(spec, dev):
    nat = int(np.sum(spec.natoms))
    return dict(energy=torch.full((len(spec),), 12345.0, dtype=torch.float64, device=dev),
                gradient=torch.full((nat, 3), 12345.0, dtype=torch.float64, device=dev),
                sigma=torch.full((len(spec), 3, 3), 12345.0, dtype=torch.float64, device=dev))


def _call(noalloc, spec, pos, out, precision):
    from dftd3 import library, torch_api
    prec = {"fp64": library.PRECISION_FP64, "mixed": library.PRECISION_MIXED}[precision]
    stream = torch.cuda.current_stream().cuda_stream
    ptrs = (pos.data_ptr(), out["energy"].data_ptr(), out["gradient"].data_ptr(), out["sigma"].data_ptr(), stream)
    if noalloc:
        library.get_dispersion_batch_noalloc(spec.natoms, spec.numbers, spec.lattices, spec.periodic, spec.params,
                                             spec.cutoffs, prec, *ptrs)
    else:
        library.get_dispersion_batch(torch_api.workspace(pos.device), spec.natoms, spec.numbers, spec.lattices,
                                     spec.periodic, spec.params, spec.cutoffs, prec, *ptrs)
    torch.cuda.synchronize()


def _agree(a, b):
    ea, eb = a["energy"].cpu().numpy(), b["energy"].cpu().numpy()
    assert np.all(np.abs(ea - eb) <= 1e-12 * np.abs(eb)), (ea, eb)
    for key in ("gradient", "sigma"):
        x, y = a[key].cpu().numpy(), b[key].cpu().numpy()
        assert np.abs(x - y).max() <= 1e-12 * np.abs(y).max(), key


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_batch_entry_two_calls(precision):
    from dftd3 import torch_api
    dev = torch.device("cuda")
    for k, (sizes, cellq) in enumerate([((60, 45, 30), True), ((24,), False)]):
        _step(f"{precision}: call {k + 1}, molecules {sizes}, cell {cellq}")
        systems = _systems(sizes, seed=20 + k, with_cell=cellq)
        spec, pos = torch_api.batch_spec(systems)
        p = torch.as_tensor(pos, dtype=torch.float64, device=dev).contiguous()
        ref, out = _outputs(spec, dev), _outputs(spec, dev)
        _call(False, spec, p, ref, precision)
        free0, _ = torch.cuda.mem_get_info()
        _call(True, spec, p, out, precision)
        free1, _ = torch.cuda.mem_get_info()
        assert free1 == free0, (free0, free1)
        _agree(out, ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_persistent_three_body_grid(precision):
    """A 320-atom cell (about 10,000 three-body work items, above the
    threshold of the persistent launch: its work counter lives in the work
    area too) through both entry points."""
    from dftd3 import torch_api
    dev = torch.device("cuda")
    _step(f"{precision}: 320-atom cell, persistent three-body grid")
    num, pos, lat = cell(41, 320, 28.0)
    spec, pos = torch_api.batch_spec([dict(numbers=num, positions=pos, lattice=lat, periodic=[True] * 3, param=BJ)])
    p = torch.as_tensor(pos, dtype=torch.float64, device=dev).contiguous()
    ref, out = _outputs(spec, dev), _outputs(spec, dev)
    _call(False, spec, p, ref, precision)
    free0, _ = torch.cuda.mem_get_info()
    _call(True, spec, p, out, precision)
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, (free0, free1)
    _agree(out, ref)


def test_noalloc_rejects_a_batch_beyond_the_work_area():
    from dftd3 import synthetic, torch_api
    dev = torch.device("cuda")
    _step("1,400-atom molecule (FP64 needs about 140 MB)")
    num, pos = synthetic.molecule_c0(seed=3, n=1400)
    spec, pos = torch_api.batch_spec([dict(numbers=num, positions=pos, param=BJ)])
    p = torch.as_tensor(pos, dtype=torch.float64, device=dev).contiguous()
    out = _outputs(spec, dev)
    with pytest.raises(RuntimeError, match="fixed work area of the allocation-free entry point"):
        _call(True, spec, p, out, "fp64")
    torch.cuda.synchronize()
    for v in out.values():
        assert torch.all(v == 12345.0).item()
    # the entry point stays usable
    systems = _systems((30,), seed=7, with_cell=False)
    spec, pos = torch_api.batch_spec(systems)
    p = torch.as_tensor(pos, dtype=torch.float64, device=dev).contiguous()
    ref, out = _outputs(spec, dev), _outputs(spec, dev)
    _call(False, spec, p, ref, "fp64")
    _call(True, spec, p, out, "fp64")
    _agree(out, ref)
