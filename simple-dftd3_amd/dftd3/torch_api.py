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

_KIND = {"rational": 0, "bj": 0, "mrational": 0, "bjm": 0, "zero": 1, "mzero": 2, "zerom": 2,
         "optimizedpower": 3, "op": 3, "cso": 4, "z": 5}
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
    """[kind, s6, s8, s9, a1|rs6, a2|rs8, alp, bet, a3, a4] of the C ABI, from
    the keyword names of the damping parameter classes."""
    k = _KIND[kind]
    s6, s9, alp = param.get("s6", 1.0), param.get("s9", 1.0), param.get("alp", 14.0)
    if k in (0, 3):  # rational, optimized power
        row = [param["s8"], s9, param["a1"], param["a2"], alp, param.get("bet", 0.0), 0.0, 0.0]
    elif k in (1, 2):  # zero, modified zero
        row = [param["s8"], s9, param["rs6"], param.get("rs8", 1.0), alp, param.get("bet", 0.0), 0.0, 0.0]
    elif k == 4:  # CSO
        row = [0.0, s9, param["a1"], param.get("a2", 2.5), alp, 0.0, param.get("a3", 0.0), param.get("a4", 6.25)]
    else:  # Z
        row = [param.get("s8", 1.0), s9, param["a1"], 0.0, alp, 0.0, 0.0, 0.0]
    return [float(k), s6] + [float(v) for v in row]


class BatchSpec:
    """Host metadata of a batch (everything except the positions)."""

    def __init__(self, natoms, numbers, lattices, periodic, params, cutoffs):
        self.natoms = np.ascontiguousarray(natoms, dtype=np.int32)
        self.numbers = np.ascontiguousarray(numbers, dtype=np.int32)
        self.lattices = None if lattices is None else np.ascontiguousarray(lattices, dtype=float)
        self.periodic = None if periodic is None else np.ascontiguousarray(periodic, dtype=bool)
        self.params = np.ascontiguousarray(params, dtype=float)
        self.cutoffs = None if cutoffs is None else np.ascontiguousarray(cutoffs, dtype=float)

    def __len__(self):
        return len(self.natoms)


def batch_spec(systems: Sequence[dict]):
    """BatchSpec and the concatenated host positions of a list of systems."""
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
        p = p.detach().cpu().numpy() if isinstance(p, torch.Tensor) else np.asarray(p, dtype=float)
        pos.append(np.asarray(p, dtype=float).reshape(-1, 3))
    return BatchSpec(natoms, numbers, lattices, periodic, params, cutoffs), np.concatenate(pos)


def get_dispersion_batch(systems: Sequence[dict], precision: str = "fp64", grad: bool = True,
                         device: Optional[torch.device] = None) -> dict:
    """Evaluate a batch of systems.

    Each system is a dict with ``numbers`` (n,), ``positions`` (n, 3) in bohr
    (numpy array or tensor), optional ``lattice`` (3, 3) and ``periodic``
    (3,), ``kind`` ('rational', 'zero', 'mzero', 'optimizedpower', 'cso' or
    'z'), ``param`` (dict of damping parameters) and optional ``cutoff`` (dict with disp2, disp3, cn,
    width2, width3).  Positions are used as given (not wrapped).

    Returns ``energy`` (nsys,), and with ``grad`` also ``gradient`` (N, 3)
    over the concatenated atoms and ``sigma`` (nsys, 3, 3) laid out like the
    reference's ``virial``.
    """
    dev = torch.device(device or "cuda")
    spec, pos = batch_spec(systems)
    positions = torch.as_tensor(pos, dtype=torch.float64, device=dev).contiguous()
    return run_batch(spec, positions, precision=precision, grad=grad)


def _check_positions(positions: torch.Tensor):
    if positions.dtype != torch.float64 or not positions.is_contiguous() or not positions.is_cuda:
        raise ValueError("positions must be a contiguous float64 tensor on a GPU")


def run_batch(spec: BatchSpec, positions: torch.Tensor, precision: str = "fp64", grad: bool = True,
              out: Optional[dict] = None, ws: Optional[library.ModelHandle] = None) -> dict:
    """Batched call on prepared host metadata and a device tensor of the
    concatenated positions (float64, contiguous, bohr).  ``out`` may supply
    preallocated result tensors (energy, gradient, sigma); the call runs on
    the current stream."""
    _check_positions(positions)
    nsys = len(spec)
    dev = positions.device
    prec = {"fp64": library.PRECISION_FP64, "mixed": library.PRECISION_MIXED}[precision]
    if out is None:
        out = dict(energy=torch.empty(nsys, dtype=torch.float64, device=dev))
        if grad:
            out["gradient"] = torch.empty(int(np.sum(spec.natoms)), 3, dtype=torch.float64, device=dev)
            out["sigma"] = torch.empty(nsys, 3, 3, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    with torch.cuda.device(dev):
        library.get_dispersion_batch(
            ws or workspace(dev), spec.natoms, spec.numbers, spec.lattices, spec.periodic, spec.params, spec.cutoffs,
            prec, positions.data_ptr(), out["energy"].data_ptr(),
            out["gradient"].data_ptr() if grad else None, out["sigma"].data_ptr() if grad else None, stream)
    return out


def run_pairwise(spec: BatchSpec, positions: torch.Tensor, precision: str = "fp64",
                 ws: Optional[library.ModelHandle] = None) -> dict:
    """Pairwise-resolved energies of a batch (disp.f90:403-457) on the
    current stream: ``pair2`` (two-body) and ``pair3`` (three-body) are flat
    float64 tensors holding per system an (n, n) block in atom order, the
    blocks one after another; ``blocks(out, key)`` splits them into (n, n)
    views."""
    _check_positions(positions)
    dev = positions.device
    prec = {"fp64": library.PRECISION_FP64, "mixed": library.PRECISION_MIXED}[precision]
    total = int(np.sum(spec.natoms.astype(np.int64) ** 2))
    out = dict(pair2=torch.empty(total, dtype=torch.float64, device=dev),
               pair3=torch.empty(total, dtype=torch.float64, device=dev))
    stream = torch.cuda.current_stream(dev).cuda_stream
    with torch.cuda.device(dev):
        library.get_pairwise_dispersion_batch(
            ws or workspace(dev), spec.natoms, spec.numbers, spec.lattices, spec.periodic, spec.params, spec.cutoffs,
            prec, positions.data_ptr(), out["pair2"].data_ptr(), out["pair3"].data_ptr(), stream)
    out["natoms"] = spec.natoms
    return out


def blocks(out: dict, key: str):
    """(n, n) views of the per-system blocks of ``run_pairwise``'s ``key``."""
    views, off = [], 0
    for n in out["natoms"]:
        n = int(n)
        views.append(out[key][off:off + n * n].view(n, n))
        off += n * n
    return views


def run_hessian(spec: BatchSpec, positions: torch.Tensor, ws: Optional[library.ModelHandle] = None) -> dict:
    """Energies and analytic Hessians of a batch (disp.f90:236-308), FP64, on
    the current stream: ``energy`` (nsys,) and ``hessian``, a flat float64
    tensor holding per system a (3n, 3n) block, the blocks one after another;
    ``hessian_blocks(out)`` splits it into views."""
    _check_positions(positions)
    dev = positions.device
    n = spec.natoms.astype(np.int64)
    out = dict(energy=torch.empty(len(spec), dtype=torch.float64, device=dev),
               hessian=torch.empty(int(np.sum(9 * n * n)), dtype=torch.float64, device=dev), natoms=spec.natoms)
    stream = torch.cuda.current_stream(dev).cuda_stream
    with torch.cuda.device(dev):
        library.get_dispersion_hessian_batch(
            ws or workspace(dev), spec.natoms, spec.numbers, spec.lattices, spec.periodic, spec.params, spec.cutoffs,
            positions.data_ptr(), out["energy"].data_ptr(), out["hessian"].data_ptr(), stream)
    return out


def hessian_blocks(out: dict):
    """(3n, 3n) views of the per-system blocks of ``run_hessian``."""
    views, off = [], 0
    for n in out["natoms"]:
        m = 3 * int(n)
        views.append(out["hessian"][off:off + m * m].view(m, m))
        off += m * m
    return views


def count_terms(spec: BatchSpec, positions: torch.Tensor) -> np.ndarray:
    """(nsys, 3) int64 array of CN terms, two-body pair terms and ATM triple
    terms as the reference's loops visit them (counted on the GPU)."""
    _check_positions(positions)
    dev = positions.device
    stream = torch.cuda.current_stream(dev).cuda_stream
    with torch.cuda.device(dev):
        return library.count_terms_batch(workspace(dev), spec.natoms, spec.numbers, spec.lattices, spec.periodic,
                                         spec.params, spec.cutoffs, positions.data_ptr(), stream)


class BatchPlan:
    """A batch planned once and evaluated many times with new coordinates of
    the same systems (include/d3hip.h, plans).

    The plan owns a positions tensor and the output tensors; ``capture()``
    records one evaluation of them into a HIP graph and ``replay()`` launches
    it, so a caller writes new coordinates into ``plan.positions`` and
    replays::

        plan = BatchPlan(spec, positions, precision="mixed")
        plan.autotune()            # three-body launch shape for this batch
        plan.capture()
        plan.positions.copy_(new_positions)
        out = plan.replay()        # dict of energy, gradient, sigma tensors

    ``execute(positions)`` runs one evaluation without a graph.  Every call
    runs on the current stream of the plan's device; the capture uses a
    stream of its own.
    """

    def __init__(self, spec: BatchSpec, positions: torch.Tensor, precision: str = "mixed", grad: bool = True):
        _check_positions(positions)
        self.spec = spec
        self.grad = grad
        self.precision = precision
        self.device = positions.device
        self.natoms = int(np.sum(spec.natoms))
        prec = {"fp64": library.PRECISION_FP64, "mixed": library.PRECISION_MIXED}[precision]
        self.positions = positions.clone()
        nsys = len(spec)
        self.out = dict(energy=torch.zeros(nsys, dtype=torch.float64, device=self.device))
        if grad:
            self.out["gradient"] = torch.zeros(self.natoms, 3, dtype=torch.float64, device=self.device)
            self.out["sigma"] = torch.zeros(nsys, 3, 3, dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            self.handle = library.new_plan(spec.natoms, spec.numbers, spec.lattices, spec.periodic, spec.params,
                                           spec.cutoffs, prec, grad, self.positions.data_ptr(), self._stream())
        self._capture_stream = None

    def _stream(self):
        return torch.cuda.current_stream(self.device).cuda_stream

    def _ptrs(self, out):
        return (out["energy"].data_ptr(), out["gradient"].data_ptr() if self.grad else None,
                out["sigma"].data_ptr() if self.grad else None)

    def execute(self, positions: Optional[torch.Tensor] = None, out: Optional[dict] = None) -> dict:
        """One evaluation (default: the plan's own positions and outputs)."""
        pos = self.positions if positions is None else positions
        _check_positions(pos)
        if pos.shape[0] != self.natoms:
            raise ValueError("positions do not match the planned batch")
        out = self.out if out is None else out
        with torch.cuda.device(self.device):
            library.plan_execute(self.handle, pos.data_ptr(), *self._ptrs(out), self._stream())
        return out

    @property
    def shape(self):
        """(wavefronts per block, persistent blocks per CU) of the three-body kernel."""
        return library.plan_get_shape(self.handle)

    def set_shape(self, waves: int, blocks_per_cu: int) -> None:
        library.plan_set_shape(self.handle, waves, blocks_per_cu)

    def autotune(self, reps: int = 3):
        """Time the candidate shapes on the plan's positions and keep the
        fastest; returns (shape, {shape: mean ms per evaluation})."""
        with torch.cuda.device(self.device):
            best, times = library.plan_autotune(self.handle, self.positions.data_ptr(), *self._ptrs(self.out),
                                                reps, self._stream())
        shapes = library.plan_candidate_shapes()
        return shapes[best], dict(zip(shapes, times))

    def capture(self) -> None:
        """Record one evaluation of ``positions`` -> ``out`` into a HIP graph."""
        if self._capture_stream is None:
            self._capture_stream = torch.cuda.Stream(device=self.device)
        cur = torch.cuda.current_stream(self.device)
        self._capture_stream.wait_stream(cur)
        with torch.cuda.device(self.device):
            library.plan_capture(self.handle, self.positions.data_ptr(), *self._ptrs(self.out),
                                 self._capture_stream.cuda_stream)
        cur.wait_stream(self._capture_stream)

    def replay(self) -> dict:
        """Launch the captured graph on the current stream; returns ``out``."""
        with torch.cuda.device(self.device):
            library.plan_replay(self.handle, self._stream())
        return self.out

    def overflowed(self) -> bool:
        """True when an evaluation had more three-body slots than the plan
        holds.  Such an evaluation skips the three-body terms and returns NaN
        energies, gradients and virials for every system with a three-body
        term (systems without one are unaffected).  Synchronizes."""
        with torch.cuda.device(self.device):
            return library.plan_overflowed(self.handle, self._stream())
