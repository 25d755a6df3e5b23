"""Drop-in Python interface of the D3 energy / gradient / virial path.

Same module path, class names, keyword arguments, defaults and error
messages as python/dftd3/interface.py of simple-dftd3 v1.6.0 for the path
this library implements:

* ``Structure(numbers, positions, lattice=None, periodic=None)`` and
  ``Structure.update(positions, lattice=None)`` (interface.py:63-152),
* the damping parameter classes ``RationalDampingParam``,
  ``ZeroDampingParam``, ``ModifiedRationalDampingParam``,
  ``ModifiedZeroDampingParam``, ``OptimizedPowerDampingParam``,
  ``CSODampingParam`` and ``ZDampingParam``, built from explicit parameters
  or loaded by method name (interface.py:155-502),
* ``DispersionModel(numbers, positions, lattice, periodic)`` with
  ``set_realspace_cutoff``, ``set_ghost_atoms`` and
  ``get_dispersion(param, grad)`` returning ``energy`` (0-d array),
  ``gradient`` (n, 3) and ``virial`` (3, 3) in Hartree and bohr
  (interface.py:505-616), and ``get_pairwise_dispersion(param)`` returning
  the ``additive pairwise energy`` and ``non-additive pairwise energy``
  (n, n) matrices (interface.py:618-636), and ``get_hessian(param)`` returning
  the ``energy`` and the analytic (3n, 3n) ``hessian`` (interface.py:637-658).

All evaluation happens on the GPU through libd3hip.so.  Additions beyond the
reference: ``DispersionModel.set_precision("fp64" | "mixed")`` and
``get_lowrank()``.  The counterpoise correction (GCP) of the reference is not
part of this library; its class ``GeometricCounterpoise`` exists so that
imports keep working, and raises ``NotImplementedError`` when constructed.
"""
from typing import Optional

import numpy as np

from . import library


class Structure:
    """Molecular structure data with immutable species and periodicity.

    Raises ValueError on invalid input (wrong shapes), RuntimeError with the
    library message otherwise (e.g. "Too close interatomic distances found").
    """

    _mol = library.StructureHandle.null()

    def __init__(self, numbers: np.ndarray, positions: np.ndarray, lattice: Optional[np.ndarray] = None,
                 periodic: Optional[np.ndarray] = None):
        numbers = np.asarray(numbers)
        positions = np.asarray(positions)
        if positions.size % 3 != 0:
            raise ValueError("Expected tripels of cartesian coordinates")
        if 3 * numbers.size != positions.size:
            raise ValueError("Dimension missmatch between numbers and positions")
        self._natoms = len(numbers)
        _numbers = np.ascontiguousarray(numbers, dtype="i4")
        _positions = np.ascontiguousarray(positions, dtype=float)
        _lattice = None
        if lattice is not None:
            lattice = np.asarray(lattice)
            if lattice.size != 9:
                raise ValueError("Invalid lattice provided")
            _lattice = np.ascontiguousarray(lattice, dtype="float")
        _periodic = None
        if periodic is not None:
            periodic = np.asarray(periodic)
            if periodic.size != 3:
                raise ValueError("Invalid periodicity provided")
            _periodic = np.ascontiguousarray(periodic, dtype="bool")
        self._mol = library.new_structure(self._natoms, _numbers, _positions, _lattice, _periodic)

    def __len__(self):
        return self._natoms

    def update(self, positions: np.ndarray, lattice: Optional[np.ndarray] = None) -> None:
        """Update coordinates and (optionally) the lattice, both in bohr."""
        positions = np.asarray(positions)
        if 3 * len(self) != positions.size:
            raise ValueError("Dimension missmatch for positions")
        _positions = np.ascontiguousarray(positions, dtype="float")
        _lattice = None
        if lattice is not None:
            lattice = np.asarray(lattice)
            if lattice.size != 9:
                raise ValueError("Invalid lattice provided")
            _lattice = np.ascontiguousarray(lattice, dtype="float")
        library.update_structure(self._mol, _positions, _lattice)


class DampingParam:
    """Damping parameters, created from a method name or explicit values.

    Mixing both ways is not allowed: with ``method`` the parameters are
    loaded, otherwise all required values must be given as keywords.
    """

    _param = library.ParamHandle.null()

    def __init__(self, **kwargs):
        if "method" in kwargs and kwargs["method"] is None:
            del kwargs["method"]
        if "method" in kwargs:
            self._param = self.load_param(**kwargs)
        else:
            self._param = self.new_param(**kwargs)

    @staticmethod
    def load_param(method, **kwargs) -> library.ParamHandle:
        raise NotImplementedError("Child class has to define parameter loading")

    @staticmethod
    def new_param(**kwargs) -> library.ParamHandle:
        raise NotImplementedError("Child class has to define parameter construction")


class RationalDampingParam(DampingParam):
    """Rational (Becke-Johnson) damping; the ATM term uses zero damping."""

    def __init__(self, **kwargs):
        _rename_kwargs(kwargs, "alpha6", "alp")
        DampingParam.__init__(self, **kwargs)

    @staticmethod
    def load_param(method: str, atm: bool = False) -> library.ParamHandle:
        return library.load_rational_damping(method, atm)

    @staticmethod
    def new_param(*, s6: float = 1.0, s8: float, s9: float = 1.0, a1: float, a2: float,
                  alp: float = 14.0) -> library.ParamHandle:
        return library.new_rational_damping(s6, s8, s9, a1, a2, alp)


class ZeroDampingParam(DampingParam):
    """Zero (Chai-Head-Gordon) damping of the original DFT-D3."""

    def __init__(self, **kwargs):
        _rename_kwargs(kwargs, "sr6", "rs6")
        _rename_kwargs(kwargs, "sr8", "rs8")
        _rename_kwargs(kwargs, "alpha6", "alp")
        DampingParam.__init__(self, **kwargs)

    @staticmethod
    def load_param(method: str, atm: bool = False) -> library.ParamHandle:
        return library.load_zero_damping(method, atm)

    @staticmethod
    def new_param(*, s6: float = 1.0, s8: float, s9: float = 1.0, rs6: float, rs8: float = 1.0,
                  alp: float = 14.0) -> library.ParamHandle:
        return library.new_zero_damping(s6, s8, s9, rs6, rs8, alp)


class ModifiedRationalDampingParam(DampingParam):
    """Rational damping function with the modified parameters of Smith et al.
    (interface.py:290-329); a full parameter set is the same as
    ``RationalDampingParam``."""

    def __init__(self, **kwargs):
        _rename_kwargs(kwargs, "alpha6", "alp")
        DampingParam.__init__(self, **kwargs)

    @staticmethod
    def load_param(method: str, atm: bool = False) -> library.ParamHandle:
        return library.load_mrational_damping(method, atm)

    @staticmethod
    def new_param(*, s6: float = 1.0, s8: float, s9: float = 1.0, a1: float, a2: float,
                  alp: float = 14.0) -> library.ParamHandle:
        return library.new_mrational_damping(s6, s8, s9, a1, a2, alp)


class ModifiedZeroDampingParam(DampingParam):
    """Modified zero damping with the offset ``bet`` (interface.py:332-376);
    identical to zero damping for ``bet=0``."""

    def __init__(self, **kwargs):
        _rename_kwargs(kwargs, "sr6", "rs6")
        _rename_kwargs(kwargs, "sr8", "rs8")
        _rename_kwargs(kwargs, "alpha6", "alp")
        _rename_kwargs(kwargs, "beta", "bet")
        DampingParam.__init__(self, **kwargs)

    @staticmethod
    def load_param(method: str, atm: bool = False) -> library.ParamHandle:
        return library.load_mzero_damping(method, atm)

    @staticmethod
    def new_param(*, s6: float = 1.0, s8: float, s9: float = 1.0, rs6: float, rs8: float = 1.0,
                  alp: float = 14.0, bet: float) -> library.ParamHandle:
        return library.new_mzero_damping(s6, s8, s9, rs6, rs8, alp, bet)


class OptimizedPowerDampingParam(DampingParam):
    """Optimized power damping (interface.py:379-421); ``bet=0`` is the
    rational damping."""

    def __init__(self, **kwargs):
        _rename_kwargs(kwargs, "alpha6", "alp")
        _rename_kwargs(kwargs, "beta", "bet")
        DampingParam.__init__(self, **kwargs)

    @staticmethod
    def load_param(method: str, atm: bool = False) -> library.ParamHandle:
        return library.load_optimizedpower_damping(method, atm)

    @staticmethod
    def new_param(*, s6: float = 1.0, s8: float, s9: float = 1.0, a1: float, a2: float, alp: float = 14.0,
                  bet) -> library.ParamHandle:
        return library.new_optimizedpower_damping(s6, s8, s9, a1, a2, alp, bet)


class CSODampingParam(DampingParam):
    """CSO (C6-scaled only) damping (interface.py:424-464)."""

    def __init__(self, **kwargs):
        _rename_kwargs(kwargs, "alpha6", "alp")
        DampingParam.__init__(self, **kwargs)

    @staticmethod
    def load_param(method: str, atm: bool = False) -> library.ParamHandle:
        return library.load_cso_damping(method, atm)

    @staticmethod
    def new_param(*, s6: float = 1.0, s9: float = 1.0, a1: float, a2: float = 2.5, a3: float = 0.0,
                  a4: float = 6.25, alp: float = 14.0) -> library.ParamHandle:
        return library.new_cso_damping(s6, s9, a1, a2, a3, a4, alp)


class ZDampingParam(DampingParam):
    """Z damping (interface.py:467-502); the parameter loader has no
    entries."""

    def __init__(self, **kwargs):
        _rename_kwargs(kwargs, "alpha6", "alp")
        DampingParam.__init__(self, **kwargs)

    @staticmethod
    def load_param(method: str, atm: bool = False) -> library.ParamHandle:
        return library.load_z_damping(method, atm)

    @staticmethod
    def new_param(*, s6: float = 1.0, s8: float = 1.0, s9: float = 1.0, a1: float,
                  alp: float = 14.0) -> library.ParamHandle:
        return library.new_z_damping(s6, s8, s9, a1, alp)


class DispersionModel(Structure):
    """Dispersion model bound to its structure (species and periodicity)."""

    _disp = library.ModelHandle.null()

    def __init__(self, numbers: np.ndarray, positions: np.ndarray, lattice: Optional[np.ndarray] = None,
                 periodic: Optional[np.ndarray] = None):
        Structure.__init__(self, numbers, positions, lattice, periodic)
        self._disp = library.new_d3_model(self._mol)

    def set_realspace_cutoff(self, disp2: float, disp3: float, cn: float, width2: float = 0.0,
                             width3: float = 0.0):
        """Set realspace cutoffs for evaluation of interactions"""
        library.set_model_realspace_cutoff(self._disp, disp2, disp3, cn, width2, width3)

    def set_ghost_atoms(self, ghost_atoms) -> None:
        """Disable dispersion contributions from selected atoms (0-based)."""
        library.set_model_ghost_index(self._disp, np.asarray(ghost_atoms, dtype="i4"))

    def set_precision(self, precision: str) -> None:
        """'fp64' (default) or 'mixed' (FP32 arithmetic, FP64 accumulation)."""
        modes = {"fp64": library.PRECISION_FP64, "mixed": library.PRECISION_MIXED}
        if precision not in modes:
            raise ValueError(f"Unknown precision '{precision}', expected one of {sorted(modes)}")
        library.set_model_precision(self._disp, modes[precision])

    def get_dispersion(self, param: DampingParam, grad: bool) -> dict:
        """Perform actual evaluation of the dispersion correction"""
        _energy = np.array(0.0)
        if grad:
            _gradient = np.zeros((len(self), 3))
            _sigma = np.zeros((3, 3))
        else:
            _gradient = None
            _sigma = None
        library.get_dispersion(self._mol, self._disp, param._param, _energy, _gradient, _sigma)
        results = dict(energy=_energy)
        if _gradient is not None:
            results.update(gradient=_gradient)
        if _sigma is not None:
            results.update(virial=_sigma)
        return results

    def set_ewald_summation(self, rank: int = 0, tolerance: float = 0.0, kcut: float = 0.0, mesh: int = 0) -> None:
        """Evaluate the two-body dispersion energy by summation over the
        reciprocal lattice (interface.py:540-568): every C6 of the model comes
        from a low-rank expansion of the reference C6 (rank > 0 fixed, else
        the smallest within ``tolerance``, default 1e-4); under 3D periodic
        boundary conditions the two-body term is summed over the reciprocal
        lattice within ``kcut`` (default from the damping radii) and the
        two-body real-space cutoff is ignored.  Rational and zero damping
        only.  ``mesh < 0`` sums the vectors directly, ``mesh = 0`` uses a
        particle mesh derived from the cutoff and ``mesh > 0`` that many mesh
        points per direction (rounded up to a power of two)."""
        library.set_model_ewald(self._disp, rank, tolerance, kcut, mesh)

    def get_lowrank(self) -> tuple:
        """(rank, largest relative error of the reference C6) of the low-rank
        expansion set by set_ewald_summation, (0, 0.0) without one."""
        return library.get_model_lowrank(self._disp)

    def set_work_partition(self, part: int, nparts: int) -> None:
        """
        Assign an externally managed part of the interaction loops to this
        model (interface.py:575-589).

        Parts are zero based, summing the results of all parts reproduces the
        complete calculation.  Structure dependent quantities are evaluated
        for the full system on every part, only the pairwise, three-body and
        reciprocal space loops are partitioned.  Applies to get_dispersion and
        get_hessian.
        """
        library.set_model_work_partition(self._disp, part, nparts)

    def get_pairwise_dispersion(self, param: DampingParam) -> dict:
        """Evaluate pairwise representation of the dispersion energy
        (interface.py:618-636): the additive (two-body) and non-additive
        (three-body) energy of every pair of atoms, (n, n) arrays."""
        _pair_disp2 = np.zeros((len(self), len(self)))
        _pair_disp3 = np.zeros((len(self), len(self)))
        library.get_pairwise_dispersion(self._mol, self._disp, param._param, _pair_disp2, _pair_disp3)
        return {
            "additive pairwise energy": _pair_disp2,
            "non-additive pairwise energy": _pair_disp3,
        }

    def get_hessian(self, param: DampingParam) -> dict:
        """Analytical second derivatives of the dispersion energy with respect
        to the nuclear coordinates (interface.py:637-658): ``energy`` and the
        symmetric ``3*n`` by ``3*n`` ``hessian`` in Hartree per Bohr squared,
        index ``3*i + c`` for component ``c`` of atom ``i`` (always FP64)."""
        _energy = np.array(0.0)
        _hessian = np.zeros((3 * len(self), 3 * len(self)))
        library.get_dispersion_hessian(self._mol, self._disp, param._param, _energy, _hessian)
        return {"energy": _energy, "hessian": _hessian}


def _rename_kwargs(kwargs, old_name, new_name):
    """Accept the legacy spelling of a damping parameter."""
    if old_name in kwargs and new_name not in kwargs:
        kwargs[new_name] = kwargs[old_name]
        del kwargs[old_name]


class GeometricCounterpoise(Structure):
    """Counterpoise correction (interface.py:661-724 of the reference).

    The geometric counterpoise correction (gCP) is not part of this library:
    the class exists with the reference's constructor signature and methods
    so that code importing it keeps working, and constructing it raises
    ``NotImplementedError``.
    """

    def __init__(self, numbers: np.ndarray, positions: np.ndarray, lattice: Optional[np.ndarray] = None,
                 periodic: Optional[np.ndarray] = None, method: Optional[str] = None,
                 basis: Optional[str] = None):
        raise NotImplementedError("The geometric counterpoise correction is not implemented by this library")

    def set_realspace_cutoff(self, bas: float, srb: float):
        raise NotImplementedError("The geometric counterpoise correction is not implemented by this library")

    def set_work_partition(self, part: int, nparts: int) -> None:
        raise NotImplementedError("The geometric counterpoise correction is not implemented by this library")

    def get_counterpoise(self, grad: bool) -> dict:
        raise NotImplementedError("The geometric counterpoise correction is not implemented by this library")
