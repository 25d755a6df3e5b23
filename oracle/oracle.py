"""ctypes front end of the CPU oracle (oracle/d3_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module as the checker.  The
library under test never imports it.

The oracle restates simple-dftd3 v1.6.0 in FP64; the C file cites the
reference file:line of every function.  ``ORACLE_AATOAU`` is the
Angstrom -> bohr factor of mctc-lib: the Bohr radius hbar/(m_e c alpha) from
the CODATA 2018 constants (h and c exact, m_e, alpha measured).  Both tight
2D goldens of test/unit/test_periodic_2d.f90 (gh185 BJ, BP86 zero damping)
agree with it to 100 epsilon, while the rounded value 0.529177210903 A misses
the BP86 one by 2.4e-13 Eh (tests/test_oracle_golden.py).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libd3oracle.so")
TABLES = os.path.join(HERE, "..", "simple-dftd3_amd", "data", "d3_tables.npz")


def _aatoau():
    import math
    hbar = 6.62607015e-34 / (2.0 * math.pi)  # J s
    bohr = hbar / (9.1093837015e-31 * 299792458.0 * 7.2973525693e-3)  # m
    return 1.0 / (bohr * 1.0e10)


ORACLE_AATOAU = _aatoau()

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    """Compile the oracle in place (make)."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)


def _ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def lib(aatoau=ORACLE_AATOAU):
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.d3o_rcov.restype = ctypes.c_double
        _lib.d3o_rcov.argtypes = [ctypes.c_int]
        _lib.d3o_r4r2.restype = ctypes.c_double
        _lib.d3o_r4r2.argtypes = [ctypes.c_int]
        _lib.d3o_rvdw.restype = ctypes.c_double
        _lib.d3o_rvdw.argtypes = [ctypes.c_int, ctypes.c_int]
        _lib.d3o_get_c6.restype = ctypes.c_double
        _lib.d3o_get_c6.argtypes = [ctypes.c_int] * 4
        _lib.d3o_num_threads.restype = ctypes.c_int
        _lib.d3o_fourier_term.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double] + [ctypes.c_void_p] * 4
        _lib.d3o_lowrank_info.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.d3o_lowrank_atomic_c6.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        set_tables(aatoau)
    return _lib


def set_tables(aatoau=ORACLE_AATOAU):
    t = np.load(TABLES)
    keep = [np.ascontiguousarray(t["nref"], dtype=np.int32),
            np.ascontiguousarray(t["refcn"], dtype=np.float64),
            np.ascontiguousarray(t["c6ref"], dtype=np.float64),
            np.ascontiguousarray(t["r4_over_r2"], dtype=np.float64),
            np.ascontiguousarray(t["vdwrad_aa"], dtype=np.float64),
            np.ascontiguousarray(t["covrad_2009_aa"], dtype=np.float64)]
    _lib.d3o_set_tables(_ptr(keep[0], _ip), *[_ptr(a) for a in keep[1:]], ctypes.c_double(aatoau))


def _geom(numbers, positions, lattice, periodic):
    num = np.ascontiguousarray(numbers, dtype=np.int32)
    xyz = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    lat = np.zeros((3, 3)) if lattice is None else np.ascontiguousarray(lattice, dtype=np.float64).reshape(3, 3)
    if periodic is None:
        periodic = [lattice is not None] * 3
    pbc = np.ascontiguousarray(np.array(periodic, dtype=bool).astype(np.int32).reshape(-1))
    if pbc.size == 1:  # a single flag applies to all directions (mctc-lib)
        pbc = np.repeat(pbc, 3)
    return num, xyz, lat, pbc


def lattice_points(periodic, lattice, cutoff):
    L = lib()
    lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(3, 3)
    pbc = np.ascontiguousarray(np.array(periodic, dtype=np.int32).reshape(3))
    n = L.d3o_lattice_points(_ptr(pbc, _ip), _ptr(lat), ctypes.c_double(cutoff), None, 0)
    out = np.zeros((n, 3))
    L.d3o_lattice_points(_ptr(pbc, _ip), _ptr(lat), ctypes.c_double(cutoff), _ptr(out), n)
    return out


def wrap(positions, lattice, periodic):
    L = lib()
    xyz = np.array(positions, dtype=np.float64).reshape(-1, 3).copy()
    lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(3, 3)
    pbc = np.ascontiguousarray(np.array(periodic, dtype=np.int32).reshape(3))
    L.d3o_wrap(len(xyz), _ptr(xyz), _ptr(lat), _ptr(pbc, _ip))
    return xyz


def too_close(positions):
    xyz = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    return bool(lib().d3o_verify(len(xyz), _ptr(xyz)))


def coordination_number(numbers, positions, lattice=None, periodic=None, cutoff=40.0):
    L = lib()
    num, xyz, lat, pbc = _geom(numbers, positions, lattice, periodic)
    tr = lattice_points(pbc, lat, cutoff)
    cn = np.zeros(len(num))
    L.d3o_cn(len(num), _ptr(num, _ip), _ptr(xyz), len(tr), _ptr(tr), ctypes.c_double(cutoff), _ptr(cn))
    return cn


def weights(numbers, cn, grad=True):
    L = lib()
    num = np.ascontiguousarray(numbers, dtype=np.int32)
    cn = np.ascontiguousarray(cn, dtype=np.float64)
    t = np.load(TABLES)
    mref = int(max(t["nref"][z - 1] for z in num))
    gw = np.zeros((len(num), mref))
    gwd = np.zeros((len(num), mref)) if grad else None
    L.d3o_weights(len(num), _ptr(num, _ip), _ptr(cn), mref, _ptr(gw), _ptr(gwd))
    return gw, gwd


def atomic_c6(numbers, gw, gwd=None, ghost=None):
    L = lib()
    num = np.ascontiguousarray(numbers, dtype=np.int32)
    n = len(num)
    gw = np.ascontiguousarray(gw, dtype=np.float64)
    gwd = None if gwd is None else np.ascontiguousarray(gwd, dtype=np.float64)
    gh = None if ghost is None else np.ascontiguousarray(ghost, dtype=np.int32)
    c6 = np.zeros((n, n))
    dc6 = np.zeros((n, n)) if gwd is not None else None
    L.d3o_c6(n, _ptr(num, _ip), gw.shape[1], _ptr(gw), _ptr(gwd), _ptr(gh, _ip), _ptr(c6), _ptr(dc6))
    # the oracle stores Fortran (column-major) matrices; return [i, j] = (i, j)
    return c6.T.copy(), (None if dc6 is None else dc6.T.copy())


RATIONAL, ZERO, MZERO, OPTIMIZEDPOWER, CSO, ZDAMP = 0, 1, 2, 3, 4, 5


def param_vector(kind, param):
    """{s6, s8, s9, a1|rs6, a2|rs8, alp, bet, a3, a4} of the C oracle from a
    parameter dict with the names of the reference's constructors."""
    p = dict(param)
    s6, s9, alp = p.get("s6", 1.0), p.get("s9", 1.0), p.get("alp", 14.0)
    if kind in (RATIONAL, OPTIMIZEDPOWER):
        v = [s6, p["s8"], s9, p["a1"], p["a2"], alp, p.get("bet", 0.0), 0.0, 0.0]
    elif kind in (ZERO, MZERO):
        v = [s6, p["s8"], s9, p["rs6"], p.get("rs8", 1.0), alp, p.get("bet", 0.0), 0.0, 0.0]
    elif kind == CSO:
        v = [s6, 0.0, s9, p["a1"], p.get("a2", 2.5), alp, 0.0, p.get("a3", 0.0), p.get("a4", 6.25)]
    elif kind == ZDAMP:
        v = [s6, p.get("s8", 1.0), s9, p["a1"], 0.0, alp, 0.0, 0.0, 0.0]
    else:
        raise ValueError(f"unknown damping kind {kind}")
    return np.array(v, dtype=np.float64)
DEFAULT_CUTOFF = (40.0, 60.0, 40.0, 0.0, 0.0)  # cn, disp2, disp3, width2, width3


def dispersion(numbers, positions, lattice=None, periodic=None, kind=RATIONAL,
               param=None, cutoff=DEFAULT_CUTOFF, grad=True, ghost=None, wrap_cell=False):
    """FP64 oracle of get_dispersion.

    param: dict of damping parameters named as the reference's constructors
    (s6, s8, s9, alp and a1, a2 / rs6, rs8 / bet / a3, a4 by family).
    Returns a dict with energy, gradient (n,3), sigma (3,3 Fortran order as
    returned by the C API, i.e. sigma[b][a] = sigma(a,b)), virial (= the
    Python interface's array), cn, dEdcn and energies (atomic).
    """
    L = lib()
    num, xyz, lat, pbc = _geom(numbers, positions, lattice, periodic)
    if wrap_cell:
        xyz = wrap(xyz, lat, pbc)
    par = param_vector(kind, param)
    cut = np.array(cutoff, dtype=np.float64)
    n = len(num)
    gh = None if ghost is None else np.ascontiguousarray(ghost, dtype=np.int32)
    energy = np.zeros(1)
    g = np.zeros((n, 3)) if grad else None
    s = np.zeros((3, 3)) if grad else None
    cn = np.zeros(n)
    dedcn = np.zeros(n)
    en = np.zeros(n)
    L.d3o_get_dispersion(n, _ptr(num, _ip), _ptr(xyz), _ptr(lat), _ptr(pbc, _ip), kind, _ptr(par),
                         _ptr(cut), _ptr(gh, _ip), _ptr(energy), _ptr(g), _ptr(s), _ptr(cn),
                         _ptr(dedcn), _ptr(en))
    out = dict(energy=float(energy[0]), cn=cn, energies=en)
    if grad:
        out.update(gradient=g, sigma=s, virial=s.copy(), dEdcn=dedcn)
    return out


def pairwise(numbers, positions, lattice=None, periodic=None, kind=RATIONAL, param=None,
             cutoff=DEFAULT_CUTOFF, ghost=None):
    """FP64 oracle of get_pairwise_dispersion (disp.f90:403-457): returns the
    additive (two-body) and non-additive (three-body) pair energy matrices
    (n, n), [i, j] = energy2(i, j)."""
    L = lib()
    num, xyz, lat, pbc = _geom(numbers, positions, lattice, periodic)
    par = param_vector(kind, param)
    cut = np.array(cutoff, dtype=np.float64)
    n = len(num)
    gh = None if ghost is None else np.ascontiguousarray(ghost, dtype=np.int32)
    e2 = np.zeros((n, n))
    e3 = np.zeros((n, n))
    L.d3o_get_pairwise_dispersion(n, _ptr(num, _ip), _ptr(xyz), _ptr(lat), _ptr(pbc, _ip), kind,
                                  _ptr(par), _ptr(cut), _ptr(gh, _ip), _ptr(e2), _ptr(e3))
    return e2.T.copy(), e3.T.copy()


def hessian(numbers, positions, lattice=None, periodic=None, kind=RATIONAL, param=None, cutoff=DEFAULT_CUTOFF,
            ghost=None):
    """FP64 oracle of get_dispersion_hessian (disp.f90:236-308): energy and
    the (3n, 3n) analytic Hessian, [3 i + a, 3 j + b] = d2E / dR_ia dR_jb."""
    L = lib()
    num, xyz, lat, pbc = _geom(numbers, positions, lattice, periodic)
    par = param_vector(kind, param)
    cut = np.array(cutoff, dtype=np.float64)
    n = len(num)
    gh = None if ghost is None else np.ascontiguousarray(ghost, dtype=np.int32)
    energy = np.zeros(1)
    h = np.zeros((3 * n, 3 * n))
    L.d3o_get_dispersion_hessian(n, _ptr(num, _ip), _ptr(xyz), _ptr(lat), _ptr(pbc, _ip), kind, _ptr(par),
                                 _ptr(cut), _ptr(gh, _ip), _ptr(energy), _ptr(h))
    return dict(energy=float(energy[0]), hessian=h.T.copy())


def atm_hessian_parts(numbers, positions, lattice=None, periodic=None, s9=1.0, alp=14.0, cutoff=DEFAULT_CUTOFF):
    """The three-body part of the Hessian before the CN chain rule
    (atm.f90:38-429, as called from disp.f90:288-298), under the partition
    of set_partition.  Row-major: c6 etc. [i, j] with the derivatives by
    CN_i, hessian [3 i + a, 3 j + b], dEdcndr [3 i + a, k], dEdcndcn [k, l]."""
    L = lib()
    num, xyz, lat, pbc = _geom(numbers, positions, lattice, periodic)
    cut = np.array(cutoff, dtype=np.float64)
    n = len(num)
    mats = [np.zeros((n, n)) for _ in range(4)]
    h = np.zeros((3 * n, 3 * n))
    dedcn = np.zeros(n)
    dedcndr = np.zeros((3 * n, n))
    dedcndcn = np.zeros((n, n))
    L.d3o_atm_hessian_parts(n, _ptr(num, _ip), _ptr(xyz), _ptr(lat), _ptr(pbc, _ip), _ptr(cut),
                            ctypes.c_double(s9), ctypes.c_double(alp), *[_ptr(m) for m in mats], _ptr(h),
                            _ptr(dedcn), _ptr(dedcndr), _ptr(dedcndcn))
    c6, dc6, d2c6, d2c6ij = (m.T.copy() for m in mats)
    return dict(c6=c6, dc6=dc6, d2c6=d2c6, d2c6ij=d2c6ij, hessian=h.T.copy(), dedcn=dedcn, dedcndr=dedcndr,
                dedcndcn=dedcndcn)


class EwaldUnsupported(ValueError):
    """The damping function has no reciprocal-space representation."""


def ewald(numbers, positions, lattice=None, periodic=None, kind=RATIONAL, param=None, cutoff=DEFAULT_CUTOFF,
          grad=True, ghost=None, rank=0, tolerance=0.0, kcut=0.0, mesh=-1):
    """FP64 oracle of get_dispersion with a low-rank model (model.f90:185-196):
    all C6 from the low-rank expansion, the two-body term summed over the
    reciprocal lattice (fourier/ewald.f90) under 3D periodicity and in real
    space otherwise.  rank / tolerance / kcut as set_model_ewald
    (api.f90:447-451; non-positive values select the defaults); mesh < 0 (the
    default here) sums the reciprocal vectors directly (fourier/ewald.f90),
    mesh >= 0 on the particle mesh (fourier/spme.f90; 0 derives the mesh from
    the cutoff).  Returns the dict of dispersion(); raises EwaldUnsupported
    for damping functions without a Fourier representation."""
    L = lib()
    num, xyz, lat, pbc = _geom(numbers, positions, lattice, periodic)
    par = param_vector(kind, param)
    cut = np.array(cutoff, dtype=np.float64)
    n = len(num)
    gh = None if ghost is None else np.ascontiguousarray(ghost, dtype=np.int32)
    ew = np.array([rank, tolerance, kcut, mesh], dtype=np.float64)
    energy = np.zeros(1)
    g = np.zeros((n, 3)) if grad else None
    s = np.zeros((3, 3)) if grad else None
    en = np.zeros(n)
    rc = L.d3o_get_dispersion_ewald(n, _ptr(num, _ip), _ptr(xyz), _ptr(lat), _ptr(pbc, _ip), kind, _ptr(par),
                                    _ptr(cut), _ptr(gh, _ip), _ptr(ew), _ptr(energy), _ptr(g), _ptr(s), _ptr(en))
    if rc == 1:
        raise EwaldUnsupported("Damping function does not support Ewald summation")
    if rc == 2:
        raise EwaldUnsupported("Damped pair potential has no known Fourier transform")
    out = dict(energy=float(energy[0]), energies=en)
    if grad:
        out.update(gradient=g, sigma=s, virial=s.copy())
    return out


def lowrank_info(numbers, rank=0, tolerance=0.0):
    """(rank, max relative error, eigenvalues kept, size of the reference C6
    matrix) of the low-rank expansion of a structure's species."""
    L = lib()
    num = np.ascontiguousarray(numbers, dtype=np.int32)
    err = ctypes.c_double()
    nd = ctypes.c_int()
    lam = np.zeros(7 * len(set(num.tolist())))
    r = L.d3o_lowrank_info(len(num), _ptr(num, _ip), int(rank), float(tolerance), ctypes.byref(err), _ptr(lam),
                           ctypes.byref(nd))
    return r, err.value, lam[:r].copy(), nd.value


def lowrank_atomic_c6(numbers, gw, rank=0, tolerance=0.0, ghost=None):
    """Pair C6 (n, n) of the low-rank expansion from reference weights gw."""
    L = lib()
    num = np.ascontiguousarray(numbers, dtype=np.int32)
    gw = np.ascontiguousarray(gw, dtype=np.float64)
    n = len(num)
    gh = None if ghost is None else np.ascontiguousarray(ghost, dtype=np.int32)
    c6 = np.zeros((n, n))
    L.d3o_lowrank_atomic_c6(n, _ptr(num, _ip), int(rank), float(tolerance), gw.shape[1], _ptr(gw), _ptr(gh, _ip),
                            _ptr(c6))
    return c6.T.copy()


def jacobi(amat):
    """Eigenvalues (decreasing magnitude) and eigenvectors (columns) by the
    reference's cyclic Jacobi method."""
    a = np.asfortranarray(amat, dtype=np.float64)
    n = a.shape[0]
    ev = np.zeros(n)
    vec = np.zeros((n, n), order="F")
    lib().d3o_jacobi(n, a.ctypes.data_as(_dp), _ptr(ev), vec.ctypes.data_as(_dp))
    return ev, vec


def fft3d(grid, sign):
    """The oracle's radix-2 transform of a complex grid [m3][m2][m1] with
    exponent sign `sign` (unnormalised)."""
    g = np.ascontiguousarray(grid, dtype=np.complex128)
    re = np.ascontiguousarray(g.real)
    im = np.ascontiguousarray(g.imag)
    m = np.array(g.shape[::-1], dtype=np.int32)
    lib().d3o_fft3d(_ptr(re), _ptr(im), _ptr(m, _ip), int(sign))
    return re + 1j * im


def bspline(w):
    """(theta[6], dtheta[6]) of the order-6 cardinal B-spline at offset w."""
    L = lib()
    L.d3o_bspline.argtypes = [ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
    t = np.zeros(6)
    d = np.zeros(6)
    L.d3o_bspline(float(w), _ptr(t), _ptr(d))
    return t, d


def fourier_term(pref, m, alpha, radius, k):
    """(phi, dphi/dk, value at r = 0, reciprocal cutoff at 1e-8) of the term
    pref r^m / (r^alpha + radius^alpha)."""
    out = [ctypes.c_double() for _ in range(4)]
    lib().d3o_fourier_term(float(pref), int(m), int(alpha), float(radius), float(k), *[ctypes.byref(o) for o in out])
    return tuple(o.value for o in out)


def count_terms(positions, lattice=None, periodic=None, cutoff=DEFAULT_CUTOFF, triples=True):
    """Exact counts of CN, two-body and ATM terms the reference loops visit."""
    L = lib()
    num, xyz, lat, pbc = _geom(np.zeros(len(np.reshape(positions, (-1, 3))), dtype=np.int32) + 1,
                               positions, lattice, periodic)
    cut = np.array(cutoff, dtype=np.float64)
    a, b, c = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
    L.d3o_count_terms(len(xyz), _ptr(xyz), _ptr(lat), _ptr(pbc, _ip), _ptr(cut), ctypes.byref(a),
                      ctypes.byref(b), ctypes.byref(c) if triples else None)
    return a.value, b.value, (c.value if triples else None)


def num_threads():
    return lib().d3o_num_threads()


def set_num_threads(n: int):
    """Threads of the oracle's parallel loops from now on."""
    lib().d3o_set_num_threads(int(n))


def set_sample_stride(stride: int):
    """Visit only outer atoms iat % stride == 0 (bounded timing samples)."""
    lib().d3o_set_sample_stride(int(stride))


def set_partition(part: int, nparts: int) -> bool:
    """Work partition of the pair, triple and reciprocal-space loops from now
    on (partition.f90:57-148; nparts = 1: every term).  An invalid partition
    resets to the serial one and returns False."""
    return lib().d3o_set_partition(int(part), int(nparts)) == 0


class partition:
    """``with partition(p, n):`` evaluates part p of n and restores the
    serial loops afterwards."""

    def __init__(self, part: int, nparts: int):
        self.part, self.nparts = part, nparts

    def __enter__(self):
        if not set_partition(self.part, self.nparts):
            raise ValueError("Invalid dispersion work partition")
        return self

    def __exit__(self, *exc):
        set_partition(0, 1)
        return False


def time_sampled_dispersion(stride, *args, **kwargs):
    """Wall time of one sampled E+grad+sigma evaluation (seconds).  The
    results of a sampled run are partial sums and are not returned."""
    import time
    set_sample_stride(stride)
    try:
        t0 = time.perf_counter()
        dispersion(*args, **kwargs)
        return time.perf_counter() - t0
    finally:
        set_sample_stride(1)


def time_sampled_pairwise(stride, *args, **kwargs):
    """Wall time of one sampled pairwise evaluation (seconds); the results of
    a sampled run are partial sums and are not returned."""
    import time
    set_sample_stride(stride)
    try:
        t0 = time.perf_counter()
        pairwise(*args, **kwargs)
        return time.perf_counter() - t0
    finally:
        set_sample_stride(1)
