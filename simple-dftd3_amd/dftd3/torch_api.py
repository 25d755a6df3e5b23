"""Device-tensor entry points: batches of systems on PyTorch-ROCm tensors.

``get_dispersion_batch`` evaluates many independent molecules or cells in one
pass of the HIP kernels (include/d3hip.h, d3hip_get_dispersion_batch).  The
positions stay on the GPU; only the small per-system metadata (atomic
numbers, lattices, parameters) is read on the host.  Results are float64
tensors on the same device.
"""
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import library

_KIND = {"rational": 0, "bj": 0, "zero": 1}
_workspaces: Dict[int, library.ModelHandle] = {}


def workspace(device: Optional[torch.device] = None) -> library.ModelHandle:
    """Per-device workspace (device buffers) of the batched entry point."""
    dev = torch.device(device or "cuda")
    index = dev.index if dev.index is not None else torch.cuda.current_device()
    if index not in _workspaces:
        with torch.cuda.device(index):
            _workspaces[index] = library.new_workspace()
    return _workspaces[index]


def param_row(kind: str, param: dict) -> List[float]:
    """[kind, s6, s8, s9, a1|rs6, a2|rs8, alp] of the C ABI."""
    k = _KIND[kind]
    if k == 0:
        return [0.0, param.get("s6", 1.0), param["s8"], param.get("s9", 1.0), param["a1"], param["a2"],
                param.get("alp", 14.0)]
    return [1.0, param.get("s6", 1.0), param["s8"], param.get("s9", 1.0), param["rs6"], param.get("rs8", 1.0),
            param.get("alp", 14.0)]


def get_dispersion_batch(systems: Sequence[dict], precision: str = "fp64", grad: bool = True,
                         device: Optional[torch.device] = None) -> dict:
    """Evaluate a batch of systems.

    Each system is a dict with ``numbers`` (n,), ``positions`` (n, 3) in bohr
    (numpy array or tensor), optional ``lattice`` (3, 3) and ``periodic``
    (3,), ``kind`` ('rational' or 'zero'), ``param`` (dict of damping
    parameters) and optional ``cutoff`` (dict with disp2, disp3, cn,
    width2, width3).  Positions are used as given (not wrapped).

    Returns ``energy`` (nsys,), and with ``grad`` also ``gradient`` (N, 3)
    over the concatenated atoms and ``sigma`` (nsys, 3, 3) laid out like the
    reference's ``virial``.
    """
    dev = torch.device(device or "cuda")
    nsys = len(systems)
    natoms = np.array([len(s["numbers"]) for s in systems], dtype=np.int32)
    numbers = np.ascontiguousarray(np.concatenate([np.asarray(s["numbers"], dtype=np.int32) for s in systems]))
    periodic_any = any(s.get("lattice") is not None for s in systems)
    lattices = np.zeros((nsys, 3, 3)) if periodic_any else None
    periodic = np.zeros((nsys, 3), dtype=bool) if periodic_any else None
    for k, s in enumerate(systems):
        if s.get("lattice") is not None:
            lattices[k] = np.asarray(s["lattice"], dtype=float).reshape(3, 3)
            pbc = s.get("periodic")
            periodic[k] = True if pbc is None else np.asarray(pbc, dtype=bool).reshape(3)
    params = np.ascontiguousarray([param_row(s.get("kind", "rational"), s["param"]) for s in systems], dtype=float)
    cutoffs = None
    if any("cutoff" in s for s in systems):
        cutoffs = np.zeros((nsys, 5))
        for k, s in enumerate(systems):
            c = dict(disp2=60.0, disp3=40.0, cn=40.0, width2=0.0, width3=0.0)
            c.update(s.get("cutoff", {}))
            cutoffs[k] = [c["disp2"], c["disp3"], c["cn"], c["width2"], c["width3"]]
    pos = []
    for s in systems:
        p = s["positions"]
        p = p if isinstance(p, torch.Tensor) else torch.as_tensor(np.asarray(p, dtype=float))
        pos.append(p.to(device=dev, dtype=torch.float64).reshape(-1, 3))
    positions = torch.cat(pos).contiguous()
    return get_dispersion_batch_tensors(natoms, numbers, lattices, periodic, params, cutoffs, positions,
                                        precision=precision, grad=grad)


def get_dispersion_batch_tensors(natoms: np.ndarray, numbers: np.ndarray, lattices: Optional[np.ndarray],
                                 periodic: Optional[np.ndarray], params: np.ndarray, cutoffs: Optional[np.ndarray],
                                 positions: torch.Tensor, precision: str = "fp64", grad: bool = True,
                                 out: Optional[dict] = None) -> dict:
    """Low-level batched call on prepared host metadata and a device tensor
    of concatenated positions (float64, contiguous).  ``out`` may supply
    preallocated result tensors (energy, gradient, sigma)."""
    if positions.dtype != torch.float64 or not positions.is_contiguous() or not positions.is_cuda:
        raise ValueError("positions must be a contiguous float64 tensor on a GPU")
    nsys = len(natoms)
    dev = positions.device
    prec = {"fp64": library.PRECISION_FP64, "mixed": library.PRECISION_MIXED}[precision]
    if out is None:
        out = dict(energy=torch.empty(nsys, dtype=torch.float64, device=dev))
        if grad:
            out["gradient"] = torch.empty(int(np.sum(natoms)), 3, dtype=torch.float64, device=dev)
            out["sigma"] = torch.empty(nsys, 3, 3, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    with torch.cuda.device(dev):
        library.get_dispersion_batch(
            workspace(dev), np.ascontiguousarray(natoms, dtype=np.int32),
            np.ascontiguousarray(numbers, dtype=np.int32),
            None if lattices is None else np.ascontiguousarray(lattices, dtype=float),
            None if periodic is None else np.ascontiguousarray(periodic, dtype=bool),
            np.ascontiguousarray(params, dtype=float),
            None if cutoffs is None else np.ascontiguousarray(cutoffs, dtype=float),
            prec, positions.data_ptr(), out["energy"].data_ptr(),
            out["gradient"].data_ptr() if grad else None, out["sigma"].data_ptr() if grad else None, stream)
    return out
