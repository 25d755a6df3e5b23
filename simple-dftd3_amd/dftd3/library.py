"""ctypes binding of libd3hip.so, the C ABI declared in include/d3hip.h.

Mirrors python/dftd3/library.py of simple-dftd3 v1.6.0 (which binds the same
entry points through CFFI): handle classes whose lifetime is tied to the
Python object, an ``error_check`` wrapper that raises ``RuntimeError`` with
the library's message (library.py:114-128 of the reference), and one function
per C entry point with the reference's names.

The library is required: importing this module fails with ``ImportError`` if
libd3hip.so is missing or cannot be loaded.  There is no CPU fallback.
"""
import ctypes
import functools
import os
from typing import Optional

import numpy as np

_LIBNAME = "libd3hip.so"
_PATH = os.environ.get("D3HIP_LIBRARY", os.path.join(os.path.dirname(os.path.abspath(__file__)), _LIBNAME))

try:
    lib = ctypes.CDLL(_PATH)
except OSError as exc:  # pragma: no cover - exercised by deployments without a build
    raise ImportError(f"DFT-D3 HIP library unimportable ({_PATH}): {exc}") from exc

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_bp = ctypes.POINTER(ctypes.c_bool)
_d = ctypes.c_double

# Signatures of the exported entry points (include/d3hip.h).
_SIGNATURES = {
    "dftd3_get_version": (ctypes.c_int, []),
    "dftd3_new_error": (_vp, []),
    "dftd3_check_error": (ctypes.c_int, [_vp]),
    "dftd3_get_error": (None, [_vp, ctypes.c_char_p, _ip]),
    "dftd3_delete_error": (None, [ctypes.POINTER(_vp)]),
    "dftd3_new_structure": (_vp, [_vp, ctypes.c_int, _ip, _dp, _dp, _bp]),
    "dftd3_delete_structure": (None, [ctypes.POINTER(_vp)]),
    "dftd3_update_structure": (None, [_vp, _vp, _dp, _dp]),
    "dftd3_new_d3_model": (_vp, [_vp, _vp]),
    "dftd3_set_model_realspace_cutoff": (None, [_vp, _vp, _d, _d, _d]),
    "dftd3_set_model_realspace_cutoff_smooth": (None, [_vp, _vp, _d, _d, _d, _d, _d]),
    "dftd3_set_model_ghost_index": (None, [_vp, _vp, _ip, ctypes.c_int]),
    "dftd3_set_model_ewald": (None, [_vp, _vp, ctypes.c_int, _d, _d, ctypes.c_int]),
    "dftd3_set_model_work_partition": (None, [_vp, _vp, ctypes.c_int, ctypes.c_int]),
    "dftd3_set_model_mpi_comm": (None, [_vp, _vp, ctypes.c_int]),
    "dftd3_has_feature": (ctypes.c_bool, [ctypes.c_char_p]),
    "d3hip_get_model_lowrank": (None, [_vp, _vp, _ip, _dp, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "dftd3_delete_model": (None, [ctypes.POINTER(_vp)]),
    "dftd3_new_zero_damping": (_vp, [_vp, _d, _d, _d, _d, _d, _d]),
    "dftd3_load_zero_damping": (_vp, [_vp, ctypes.c_char_p, ctypes.c_bool]),
    "dftd3_new_rational_damping": (_vp, [_vp, _d, _d, _d, _d, _d, _d]),
    "dftd3_load_rational_damping": (_vp, [_vp, ctypes.c_char_p, ctypes.c_bool]),
    "dftd3_new_mzero_damping": (_vp, [_vp, _d, _d, _d, _d, _d, _d, _d]),
    "dftd3_load_mzero_damping": (_vp, [_vp, ctypes.c_char_p, ctypes.c_bool]),
    "dftd3_new_mrational_damping": (_vp, [_vp, _d, _d, _d, _d, _d, _d]),
    "dftd3_load_mrational_damping": (_vp, [_vp, ctypes.c_char_p, ctypes.c_bool]),
    "dftd3_new_optimizedpower_damping": (_vp, [_vp, _d, _d, _d, _d, _d, _d, _d]),
    "dftd3_load_optimizedpower_damping": (_vp, [_vp, ctypes.c_char_p, ctypes.c_bool]),
    "dftd3_new_cso_damping": (_vp, [_vp, _d, _d, _d, _d, _d, _d, _d]),
    "dftd3_load_cso_damping": (_vp, [_vp, ctypes.c_char_p, ctypes.c_bool]),
    "dftd3_new_z_damping": (_vp, [_vp, _d, _d, _d, _d, _d]),
    "dftd3_load_z_damping": (_vp, [_vp, ctypes.c_char_p, ctypes.c_bool]),
    "dftd3_delete_param": (None, [ctypes.POINTER(_vp)]),
    "dftd3_get_dispersion": (None, [_vp, _vp, _vp, _vp, _dp, _dp, _dp]),
    "dftd3_get_pairwise_dispersion": (None, [_vp, _vp, _vp, _vp, _dp, _dp]),
    "dftd3_get_dispersion_hessian": (None, [_vp, _vp, _vp, _vp, _dp, _dp]),
    "d3hip_get_dispersion_hessian_batch": (None, [_vp, _vp, ctypes.c_int, _ip, _ip, _dp, _bp, _dp, _dp, _vp, _vp,
                                                  _vp, _vp]),
    "d3hip_set_model_precision": (None, [_vp, _vp, ctypes.c_int]),
    "d3hip_new_workspace": (_vp, [_vp]),
    "d3hip_get_dispersion_batch": (None, [_vp, _vp, ctypes.c_int, _ip, _ip, _dp, _bp, _dp, _dp, ctypes.c_int,
                                          _vp, _vp, _vp, _vp, _vp]),
    "d3hip_get_dispersion_batch_noalloc": (None, [_vp, ctypes.c_int, _ip, _ip, _dp, _bp, _dp, _dp, ctypes.c_int,
                                                  _vp, _vp, _vp, _vp, _vp]),
    "d3hip_get_pairwise_dispersion_batch": (None, [_vp, _vp, ctypes.c_int, _ip, _ip, _dp, _bp, _dp, _dp,
                                                   ctypes.c_int, _vp, _vp, _vp, _vp]),
    "d3hip_count_terms_batch": (None, [_vp, _vp, ctypes.c_int, _ip, _ip, _dp, _bp, _dp, _dp, _vp,
                                       ctypes.POINTER(ctypes.c_longlong), _vp]),
    "d3hip_set_kernel_timing": (None, [_vp, _vp, ctypes.c_int]),
    "d3hip_get_kernel_time": (None, [_vp, _vp, ctypes.c_char_p, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "d3hip_new_plan": (_vp, [_vp, ctypes.c_int, _ip, _ip, _dp, _bp, _dp, _dp, ctypes.c_int, ctypes.c_int, _vp, _vp]),
    "d3hip_delete_plan": (None, [ctypes.POINTER(_vp)]),
    "d3hip_plan_execute": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "d3hip_plan_num_shapes": (ctypes.c_int, []),
    "d3hip_plan_candidate_shape": (None, [ctypes.c_int, _ip, _ip]),
    "d3hip_plan_get_shape": (None, [_vp, _vp, _ip, _ip]),
    "d3hip_plan_set_shape": (None, [_vp, _vp, ctypes.c_int, ctypes.c_int]),
    "d3hip_plan_autotune": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_float), _vp]),
    "d3hip_plan_capture": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "d3hip_plan_replay": (None, [_vp, _vp, _vp]),
    "d3hip_plan_overflowed": (ctypes.c_int, [_vp, _vp, _vp]),
}

for _name, (_res, _args) in _SIGNATURES.items():
    _fn = getattr(lib, _name)  # AttributeError: library does not export the symbol
    _fn.restype = _res
    _fn.argtypes = _args

PRECISION_FP64 = 0
PRECISION_MIXED = 1


def get_api_version() -> str:
    """API version of the C interface as 'major.minor.patch'."""
    v = lib.dftd3_get_version()
    return "{}.{}.{}".format(v // 10000, v % 10000 // 100, v % 100)


class Handle:
    """Owner of an opaque library handle, released when garbage collected."""

    _deleter = None

    def __init__(self, handle=None):
        self.handle = handle

    @classmethod
    def with_gc(cls, handle):
        return cls(handle)

    @classmethod
    def null(cls):
        return cls(None)

    def __del__(self):
        if self.handle and self._deleter is not None:
            ptr = _vp(self.handle)
            getattr(lib, self._deleter)(ctypes.byref(ptr))
            self.handle = None


class StructureHandle(Handle):
    _deleter = "dftd3_delete_structure"


class ModelHandle(Handle):
    _deleter = "dftd3_delete_model"


class ParamHandle(Handle):
    _deleter = "dftd3_delete_param"


class PlanHandle(Handle):
    _deleter = "d3hip_delete_plan"


class _Error:
    def __init__(self):
        self.handle = lib.dftd3_new_error()

    def __del__(self):
        if self.handle:
            ptr = _vp(self.handle)
            lib.dftd3_delete_error(ctypes.byref(ptr))
            self.handle = None


def error_check(func):
    """Run a library function with a fresh error handle, raise on error."""

    @functools.wraps(func)
    def handle_error(*args, **kwargs):
        err = _Error()
        value = func(err.handle, *args, **kwargs)
        if lib.dftd3_check_error(err.handle):
            buf = ctypes.create_string_buffer(512)
            size = ctypes.c_int(512)
            lib.dftd3_get_error(err.handle, buf, ctypes.byref(size))
            raise RuntimeError(buf.value.decode())
        return value

    return handle_error


def _cast(ctype, array: Optional[np.ndarray]):
    return None if array is None else array.ctypes.data_as(ctype)


def _char(value: Optional[str]):
    return None if value is None else value.encode()


def new_structure(natoms: int, numbers: np.ndarray, positions: np.ndarray, lattice: Optional[np.ndarray],
                  periodic: Optional[np.ndarray]) -> StructureHandle:
    """Create new molecular structure data"""
    return StructureHandle.with_gc(
        error_check(lib.dftd3_new_structure)(natoms, _cast(_ip, numbers), _cast(_dp, positions),
                                             _cast(_dp, lattice), _cast(_bp, periodic)))


def update_structure(mol: StructureHandle, positions: np.ndarray, lattice: Optional[np.ndarray]) -> None:
    """Update the molecular structure data"""
    return error_check(lib.dftd3_update_structure)(mol.handle, _cast(_dp, positions), _cast(_dp, lattice))


def new_d3_model(mol: StructureHandle) -> ModelHandle:
    """Create new D3 dispersion model"""
    return ModelHandle.with_gc(error_check(lib.dftd3_new_d3_model)(mol.handle))


def new_workspace() -> ModelHandle:
    """Create a device workspace for batched evaluations"""
    return ModelHandle.with_gc(error_check(lib.d3hip_new_workspace)())


def set_model_realspace_cutoff(disp: ModelHandle, disp2: float, disp3: float, cn: float, width2: float = 0.0,
                               width3: float = 0.0) -> None:
    """Set the realspace cutoffs of the model"""
    return error_check(lib.dftd3_set_model_realspace_cutoff_smooth)(disp.handle, disp2, disp3, cn, width2, width3)


def set_model_ghost_index(disp: ModelHandle, ghost: np.ndarray) -> None:
    """Disable dispersion contributions from selected atoms (0-based indices)"""
    idx = np.ascontiguousarray(np.asarray(ghost, dtype="i4") + 1)
    return error_check(lib.dftd3_set_model_ghost_index)(disp.handle, _cast(_ip, idx), idx.size)


def set_model_ewald(disp: ModelHandle, rank: int, tolerance: float, kcut: float, mesh: int) -> None:
    """Use a low-rank C6 expansion and the reciprocal-space two-body sum"""
    return error_check(lib.dftd3_set_model_ewald)(disp.handle, int(rank), float(tolerance), float(kcut), int(mesh))


def has_feature(feature: str) -> bool:
    """Whether an optional feature ("mpi") is available in this library"""
    return bool(lib.dftd3_has_feature(feature.encode()))


def set_model_work_partition(disp: ModelHandle, part: int, nparts: int) -> None:
    """Assign an externally managed part of the interaction loops to the model"""
    return error_check(lib.dftd3_set_model_work_partition)(disp.handle, int(part), int(nparts))


def get_model_lowrank(disp: ModelHandle, last_evaluation: bool = False) -> tuple:
    """(rank, largest relative error) of the model's low-rank expansion; with
    last_evaluation also the reciprocal cutoff and number of vectors of the
    last reciprocal-space evaluation"""
    rank = ctypes.c_int(0)
    err = ctypes.c_double(0.0)
    kcut = ctypes.c_double(0.0)
    nvec = ctypes.c_longlong(0)
    error_check(lib.d3hip_get_model_lowrank)(disp.handle, ctypes.byref(rank), ctypes.byref(err), ctypes.byref(kcut),
                                             ctypes.byref(nvec))
    if last_evaluation:
        return rank.value, err.value, kcut.value, nvec.value
    return rank.value, err.value


def set_model_precision(disp: ModelHandle, precision: int) -> None:
    """Select FP64 (0) or FP32-arithmetic / FP64-accumulation (1) pair terms"""
    return error_check(lib.d3hip_set_model_precision)(disp.handle, precision)


def new_zero_damping(s6: float, s8: float, s9: float, rs6: float, rs8: float, alp: float) -> ParamHandle:
    """Create new zero damping parameters"""
    return ParamHandle.with_gc(error_check(lib.dftd3_new_zero_damping)(s6, s8, s9, rs6, rs8, alp))


def load_zero_damping(method: str, atm: bool) -> ParamHandle:
    """Load zero damping parameters from internal storage"""
    return ParamHandle.with_gc(error_check(lib.dftd3_load_zero_damping)(_char(method), atm))


def new_rational_damping(s6: float, s8: float, s9: float, a1: float, a2: float, alp: float) -> ParamHandle:
    """Create new rational damping parameters"""
    return ParamHandle.with_gc(error_check(lib.dftd3_new_rational_damping)(s6, s8, s9, a1, a2, alp))


def load_rational_damping(method: str, atm: bool) -> ParamHandle:
    """Load rational damping parameters from internal storage"""
    return ParamHandle.with_gc(error_check(lib.dftd3_load_rational_damping)(_char(method), atm))


def new_mzero_damping(s6: float, s8: float, s9: float, rs6: float, rs8: float, alp: float,
                      bet: float) -> ParamHandle:
    """Create new modified zero damping parameters"""
    return ParamHandle.with_gc(error_check(lib.dftd3_new_mzero_damping)(s6, s8, s9, rs6, rs8, alp, bet))


def load_mzero_damping(method: str, atm: bool) -> ParamHandle:
    """Load modified zero damping parameters from internal storage"""
    return ParamHandle.with_gc(error_check(lib.dftd3_load_mzero_damping)(_char(method), atm))


def new_mrational_damping(s6: float, s8: float, s9: float, a1: float, a2: float, alp: float) -> ParamHandle:
    """Create new modified rational damping parameters"""
    return ParamHandle.with_gc(error_check(lib.dftd3_new_mrational_damping)(s6, s8, s9, a1, a2, alp))


def load_mrational_damping(method: str, atm: bool) -> ParamHandle:
    """Load modified rational damping parameters from internal storage"""
    return ParamHandle.with_gc(error_check(lib.dftd3_load_mrational_damping)(_char(method), atm))


def new_optimizedpower_damping(s6: float, s8: float, s9: float, a1: float, a2: float, alp: float,
                               bet: float) -> ParamHandle:
    """Create new optimized power damping parameters"""
    return ParamHandle.with_gc(error_check(lib.dftd3_new_optimizedpower_damping)(s6, s8, s9, a1, a2, alp, bet))


def load_optimizedpower_damping(method: str, atm: bool) -> ParamHandle:
    """Load optimized power damping parameters from internal storage"""
    return ParamHandle.with_gc(error_check(lib.dftd3_load_optimizedpower_damping)(_char(method), atm))


def new_cso_damping(s6: float, s9: float, a1: float, a2: float, a3: float, a4: float, alp: float) -> ParamHandle:
    """Create new CSO damping parameters"""
    return ParamHandle.with_gc(error_check(lib.dftd3_new_cso_damping)(s6, s9, a1, a2, a3, a4, alp))


def load_cso_damping(method: str, atm: bool) -> ParamHandle:
    """Load CSO damping parameters from internal storage"""
    return ParamHandle.with_gc(error_check(lib.dftd3_load_cso_damping)(_char(method), atm))


def new_z_damping(s6: float, s8: float, s9: float, a1: float, alp: float) -> ParamHandle:
    """Create new Z damping parameters"""
    return ParamHandle.with_gc(error_check(lib.dftd3_new_z_damping)(s6, s8, s9, a1, alp))


def load_z_damping(method: str, atm: bool) -> ParamHandle:
    """Load Z damping parameters from internal storage"""
    return ParamHandle.with_gc(error_check(lib.dftd3_load_z_damping)(_char(method), atm))


def get_dispersion(mol: StructureHandle, disp: ModelHandle, param: ParamHandle, energy: np.ndarray,
                   gradient: Optional[np.ndarray], sigma: Optional[np.ndarray]) -> None:
    """Evaluate the dispersion energy and its derivatives"""
    return error_check(lib.dftd3_get_dispersion)(mol.handle, disp.handle, param.handle, _cast(_dp, energy),
                                                 _cast(_dp, gradient), _cast(_dp, sigma))


def get_pairwise_dispersion(mol: StructureHandle, disp: ModelHandle, param: ParamHandle, energy2: np.ndarray,
                            energy3: np.ndarray) -> None:
    """Evaluate the pairwise representation of the dispersion energy"""
    return error_check(lib.dftd3_get_pairwise_dispersion)(mol.handle, disp.handle, param.handle,
                                                          _cast(_dp, energy2), _cast(_dp, energy3))


def get_dispersion_hessian(mol: StructureHandle, disp: ModelHandle, param: ParamHandle, energy: np.ndarray,
                           hessian: np.ndarray) -> None:
    """Evaluate the energy and the analytic Hessian of the dispersion energy"""
    return error_check(lib.dftd3_get_dispersion_hessian)(mol.handle, disp.handle, param.handle,
                                                         _cast(_dp, energy), _cast(_dp, hessian))


def get_dispersion_hessian_batch(workspace: ModelHandle, natoms, numbers, lattices, periodic, params, cutoffs,
                                 positions_ptr: int, energy_ptr: Optional[int], hessian_ptr: int,
                                 stream: Optional[int]) -> None:
    """Batched energies and Hessians on device pointers (include/d3hip.h, Part 2)"""
    return error_check(lib.d3hip_get_dispersion_hessian_batch)(
        workspace.handle, len(natoms), _cast(_ip, natoms), _cast(_ip, numbers), _cast(_dp, lattices),
        _cast(_bp, periodic), _cast(_dp, params), _cast(_dp, cutoffs), positions_ptr, energy_ptr, hessian_ptr,
        stream)


def get_pairwise_dispersion_batch(workspace: ModelHandle, natoms, numbers, lattices, periodic, params, cutoffs,
                                  precision: int, positions_ptr: int, pair2_ptr: int, pair3_ptr: int,
                                  stream: Optional[int]) -> None:
    """Batched pair energies on device pointers (include/d3hip.h, Part 2)"""
    return error_check(lib.d3hip_get_pairwise_dispersion_batch)(
        workspace.handle, len(natoms), _cast(_ip, natoms), _cast(_ip, numbers), _cast(_dp, lattices),
        _cast(_bp, periodic), _cast(_dp, params), _cast(_dp, cutoffs), precision, positions_ptr, pair2_ptr,
        pair3_ptr, stream)


def count_terms_batch(workspace: ModelHandle, natoms, numbers, lattices, periodic, params, cutoffs,
                      positions_ptr: int, stream: Optional[int]) -> np.ndarray:
    """Work units per system: (nsys, 3) array of CN, pair and triple terms"""
    counts = np.zeros((len(natoms), 3), dtype=np.int64)
    error_check(lib.d3hip_count_terms_batch)(
        workspace.handle, len(natoms), _cast(_ip, natoms), _cast(_ip, numbers), _cast(_dp, lattices),
        _cast(_bp, periodic), _cast(_dp, params), _cast(_dp, cutoffs), positions_ptr,
        counts.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), stream)
    return counts


def set_kernel_timing(disp: ModelHandle, enable: bool) -> None:
    """Time kernel launches with HIP events (resets the totals)"""
    return error_check(lib.d3hip_set_kernel_timing)(disp.handle, 1 if enable else 0)


def get_kernel_time(disp: ModelHandle, name: str):
    """(total milliseconds, launches) of a kernel group since timing was enabled"""
    ms = ctypes.c_double()
    calls = ctypes.c_longlong()
    error_check(lib.d3hip_get_kernel_time)(disp.handle, name.encode(), ctypes.byref(ms), ctypes.byref(calls))
    return ms.value, calls.value


def get_dispersion_batch(workspace: ModelHandle, natoms, numbers, lattices, periodic, params, cutoffs,
                         precision: int, positions_ptr: int, energy_ptr: int, gradient_ptr: Optional[int],
                         sigma_ptr: Optional[int], stream: Optional[int]) -> None:
    """Batched evaluation on device pointers (include/d3hip.h, Part 2)"""
    return error_check(lib.d3hip_get_dispersion_batch)(
        workspace.handle, len(natoms), _cast(_ip, natoms), _cast(_ip, numbers), _cast(_dp, lattices),
        _cast(_bp, periodic), _cast(_dp, params), _cast(_dp, cutoffs), precision, positions_ptr, energy_ptr,
        gradient_ptr, sigma_ptr, stream)


def get_dispersion_batch_noalloc(natoms, numbers, lattices, periodic, params, cutoffs, precision: int,
                                 positions_ptr: int, energy_ptr: int, gradient_ptr: Optional[int],
                                 sigma_ptr: Optional[int], stream: Optional[int]) -> None:
    """Batched evaluation on device pointers without any device allocation
    (include/d3hip.h, d3hip_get_dispersion_batch_noalloc)"""
    return error_check(lib.d3hip_get_dispersion_batch_noalloc)(
        len(natoms), _cast(_ip, natoms), _cast(_ip, numbers), _cast(_dp, lattices), _cast(_bp, periodic),
        _cast(_dp, params), _cast(_dp, cutoffs), precision, positions_ptr, energy_ptr, gradient_ptr, sigma_ptr, stream)


def new_plan(natoms, numbers, lattices, periodic, params, cutoffs, precision: int, grad: bool,
             positions_ptr: int, stream: Optional[int]) -> PlanHandle:
    """Plan a batch once for repeated evaluations (include/d3hip.h, plans)"""
    return PlanHandle.with_gc(error_check(lib.d3hip_new_plan)(
        len(natoms), _cast(_ip, natoms), _cast(_ip, numbers), _cast(_dp, lattices), _cast(_bp, periodic),
        _cast(_dp, params), _cast(_dp, cutoffs), precision, 1 if grad else 0, positions_ptr, stream))


def plan_execute(plan: PlanHandle, positions_ptr: int, energy_ptr: int, gradient_ptr: Optional[int],
                 sigma_ptr: Optional[int], stream: Optional[int]) -> None:
    """One evaluation of a plan on device pointers"""
    return error_check(lib.d3hip_plan_execute)(plan.handle, positions_ptr, energy_ptr, gradient_ptr, sigma_ptr,
                                               stream)


def plan_candidate_shapes():
    """[(wavefronts per block, persistent blocks per CU)] of the three-body kernel"""
    out = []
    for k in range(lib.d3hip_plan_num_shapes()):
        w, b = ctypes.c_int(), ctypes.c_int()
        lib.d3hip_plan_candidate_shape(k, ctypes.byref(w), ctypes.byref(b))
        out.append((w.value, b.value))
    return out


def plan_get_shape(plan: PlanHandle):
    w, b = ctypes.c_int(), ctypes.c_int()
    error_check(lib.d3hip_plan_get_shape)(plan.handle, ctypes.byref(w), ctypes.byref(b))
    return w.value, b.value


def plan_set_shape(plan: PlanHandle, waves: int, blocks_per_cu: int) -> None:
    return error_check(lib.d3hip_plan_set_shape)(plan.handle, waves, blocks_per_cu)


def plan_autotune(plan: PlanHandle, positions_ptr: int, energy_ptr: int, gradient_ptr: Optional[int],
                  sigma_ptr: Optional[int], reps: int, stream: Optional[int]):
    """(index of the fastest shape, mean milliseconds of every candidate)"""
    times = (ctypes.c_float * lib.d3hip_plan_num_shapes())()
    best = error_check(lib.d3hip_plan_autotune)(plan.handle, positions_ptr, energy_ptr, gradient_ptr, sigma_ptr,
                                                reps, times, stream)
    return best, list(times)


def plan_capture(plan: PlanHandle, positions_ptr: int, energy_ptr: int, gradient_ptr: Optional[int],
                 sigma_ptr: Optional[int], stream: Optional[int]) -> None:
    """Record one evaluation of the plan into a HIP graph"""
    return error_check(lib.d3hip_plan_capture)(plan.handle, positions_ptr, energy_ptr, gradient_ptr, sigma_ptr,
                                               stream)


def plan_replay(plan: PlanHandle, stream: Optional[int]) -> None:
    """Launch the captured graph of a plan"""
    return error_check(lib.d3hip_plan_replay)(plan.handle, stream)


def plan_overflowed(plan: PlanHandle, stream: Optional[int]) -> bool:
    """True when an evaluation had more three-body slots than the plan holds"""
    return bool(error_check(lib.d3hip_plan_overflowed)(plan.handle, stream))
