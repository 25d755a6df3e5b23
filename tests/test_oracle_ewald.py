"""Pins the oracle's reciprocal-space (Ewald) two-body term and the low-rank
C6 expansion it rests on (fourier/*.f90) with what fixes them independently:
the closed-form transform of the leading Becke-Johnson term, numpy's
eigensolver, the full-rank limit (identical to the plain C6), a converged
real-space lattice sum, invariances (translation, supercell) and finite
differences.  The checks follow test/unit/test_fourier.f90 and
python/dftd3/test_interface.py:252-318 of the reference, with its
thresholds.
"""
import numpy as np
import pytest

import oracle as O

from helpers import cell

PBE_BJ = dict(s6=1.0, s9=0.0, alp=14.0, a1=0.4289, s8=0.7875, a2=4.4407)   # test_fourier.f90:57-59
PBE_ZERO = dict(s6=1.0, s9=0.0, alp=14.0, rs6=1.217, s8=0.722, rs8=1.0)    # test_fourier.f90:61-63
PBE_MZERO = dict(s6=1.0, s9=0.0, alp=14.0, rs6=1.0, s8=1.0, rs8=1.0, bet=0.1)
CUTOFF = (30.0, 60.0, 15.0, 0.0, 0.0)   # test_fourier.f90:38-39 (cn, disp2, disp3)
LONG = (30.0, 400.0, 15.0, 0.0, 0.0)    # test_fourier.f90:71-72
PBC = [True, True, True]


def cubic(alat, rep):
    """Two-species cubic cell of test_fourier.f90:122-166 get_cubic."""
    frac = np.array([[0.05, 0.10, 0.15], [0.55, 0.35, 0.45]])
    num, xyz = [], []
    for ix in range(rep):
        for iy in range(rep):
            for iz in range(rep):
                shift = alat * np.array([ix, iy, iz], dtype=float)
                num += [6, 8]
                xyz += [alat * frac[0] + shift, alat * frac[1] + shift]
    return np.array(num), np.array(xyz), np.eye(3) * alat * rep


# ---------------------------------------------------------------- transform
@pytest.mark.parametrize("ik", range(21))
def test_fourier_transform_closed_form(ik):
    """test_fourier.f90:690-716: r^0 / (r^6 + b^6) has a closed transform."""
    k, b = 0.05 * ik, 4.5
    phi, _, _, _ = O.fourier_term(1.0, 0, 6, b, k)
    kb = k * b
    if k > 0:
        ref = 2 * np.pi**2 / (3 * k * b**4) * (2 * np.sin(np.sqrt(3) * kb / 2 - np.pi / 6) * np.exp(-kb / 2)
                                              + np.exp(-kb))
    else:
        ref = 2 * np.pi**2 / (3 * b**3)
    assert abs(phi - ref) <= 1e-10 * abs(ref)


@pytest.mark.parametrize("m,alpha", [(0, 6), (0, 8), (8, 14), (8, 16)])
def test_fourier_transform_quadrature(m, alpha):
    """phi(k) = 4 pi / k int r sin(kr) f(r) dr for the four terms the rational
    and zero damping use, by numerical quadrature; and dphi/dk by central
    differences (test_fourier.f90:718-724)."""
    from scipy.integrate import quad
    rad = 3.7
    f = lambda r: r**m / (r**alpha + rad**alpha)
    for k in (0.3, 0.9, 1.7):
        phi, dphi, _, _ = O.fourier_term(1.3, m, alpha, rad, k)
        val, _ = quad(lambda r: r * f(r), 0, np.inf, weight="sin", wvar=k, limit=400, epsabs=1e-15, epsrel=1e-12)
        assert phi == pytest.approx(1.3 * 4 * np.pi / k * val, rel=1e-8, abs=1e-12)
        h = 1e-6
        p1 = O.fourier_term(1.3, m, alpha, rad, k + h)[0]
        p0 = O.fourier_term(1.3, m, alpha, rad, k - h)[0]
        assert dphi == pytest.approx((p1 - p0) / (2 * h), abs=1e-6)
    # k = 0: the integral of the term over space
    phi0 = O.fourier_term(1.3, m, alpha, rad, 0.0)[0]
    val, _ = quad(lambda r: 4 * np.pi * r * r * f(r), 0, np.inf, limit=400, epsabs=1e-15, epsrel=1e-12)
    assert phi0 == pytest.approx(1.3 * val, rel=1e-9)


def test_potential_zero_and_cutoff():
    """test_fourier.f90:726-731 and kernel.f90:151-164."""
    _, _, p0, kc = O.fourier_term(1.0, 0, 6, 4.5, 0.0)
    assert p0 == pytest.approx(1.0 / 4.5**6, rel=1e-14)
    assert kc == pytest.approx(-np.log(1e-8) / (4.5 * np.sin(np.pi / 6)), rel=1e-14)
    assert O.fourier_term(1.0, 8, 14, 4.5, 0.0)[2] == 0.0


# ------------------------------------------------------------- eigensolver
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_jacobi_against_numpy(seed):
    """test_fourier.f90:171-245: eigenpairs reproduce the matrix, vectors
    are orthonormal, eigenvalues agree with LAPACK's, sorted by magnitude."""
    rng = np.random.default_rng(seed)
    n = 12 + 5 * seed
    a = rng.normal(size=(n, n))
    a = a + a.T
    ev, vec = O.jacobi(a)
    assert np.all(np.diff(np.abs(ev)) <= 1e-12)
    ref = np.linalg.eigvalsh(a)
    ref = ref[np.argsort(-np.abs(ref), kind="stable")]
    np.testing.assert_allclose(ev, ref, atol=1e-11 * np.abs(ref).max())
    np.testing.assert_allclose(vec @ np.diag(ev) @ vec.T, a, atol=1e-11 * np.abs(a).max())
    np.testing.assert_allclose(vec.T @ vec, np.eye(n), atol=1e-12)


def test_jacobi_diagonal():
    d = np.array([3.0, -7.0, 0.5, 11.0])
    ev, vec = O.jacobi(np.diag(d))
    np.testing.assert_array_equal(ev, [11.0, -7.0, 3.0, 0.5])


# ------------------------------------------------------------------ low rank
def _tyr(goldens):
    g = goldens["validation_01-energy-d3-bj"]
    return np.array(g["numbers"]), np.array(g["positions"])


def test_lowrank_rank_accuracy(goldens):
    """test_fourier.f90:402-441: error decreases with rank, full rank exact."""
    num, _ = _tyr(goldens)
    _, _, _, ndim = O.lowrank_info(num, 1)
    r1, e1, _, _ = O.lowrank_info(num, 1)
    rh, eh, _, _ = O.lowrank_info(num, ndim // 2)
    rf, ef, lam, _ = O.lowrank_info(num, 10000)
    assert (r1, rh, rf) == (1, ndim // 2, ndim)
    assert e1 > eh > ef
    assert ef <= np.sqrt(np.finfo(float).eps)
    assert np.all(np.diff(np.abs(lam)) <= 0)


def test_lowrank_truncation(goldens):
    """test_fourier.f90:445-486: tolerance 1e-4 truncates the rank and keeps
    every pair C6 within ten times the tolerance."""
    num, xyz = _tyr(goldens)
    r, err, _, ndim = O.lowrank_info(num, 0, 1e-4)
    assert r < ndim and err <= 1e-4
    cn = O.coordination_number(num, xyz, cutoff=30.0)
    gw, _ = O.weights(num, cn)
    c6ref, _ = O.atomic_c6(num, gw)
    c6 = O.lowrank_atomic_c6(num, gw, 0, 1e-4)
    assert np.all(np.abs(c6 - c6ref) <= 10 * 1e-4 * np.abs(c6ref))
    c6full = O.lowrank_atomic_c6(num, gw, 10000)
    np.testing.assert_allclose(c6full, c6ref, rtol=1e-12)


def test_lowrank_ghost(goldens):
    num, xyz = _tyr(goldens)
    cn = O.coordination_number(num, xyz, cutoff=30.0)
    gw, _ = O.weights(num, cn)
    ghost = np.zeros(len(num), dtype=np.int32)
    ghost[[0, 5]] = 1
    c6 = O.lowrank_atomic_c6(num, gw, 10000, ghost=ghost)
    assert np.all(c6[[0, 5], :] == 0) and np.all(c6[:, [0, 5]] == 0)


def test_lowrank_molecular_is_real_space(goldens):
    """test_fourier.f90 ewald-molecular: without 3D periodicity the low-rank
    model sums in real space; at full rank it equals the plain model."""
    num, xyz = _tyr(goldens)
    ref = O.dispersion(num, xyz, kind=O.RATIONAL, param=PBE_BJ, cutoff=CUTOFF)
    res = O.ewald(num, xyz, kind=O.RATIONAL, param=PBE_BJ, cutoff=CUTOFF, rank=10000)
    assert res["energy"] == pytest.approx(ref["energy"], abs=1e-10)
    np.testing.assert_allclose(res["gradient"], ref["gradient"], atol=1e-11)


# -------------------------------------------------------------------- Ewald
def test_ewald_converged_bj():
    """test_fourier.f90 ewald-converged-bj and test_interface.py:252-297."""
    num, xyz, lat = cubic(8.0, 1)
    ew = O.ewald(num, xyz, lat, PBC, O.RATIONAL, PBE_BJ, CUTOFF, grad=False, rank=10000, kcut=10.0)["energy"]
    conv = O.dispersion(num, xyz, lat, PBC, O.RATIONAL, PBE_BJ, LONG, grad=False)["energy"]
    trunc = O.dispersion(num, xyz, lat, PBC, O.RATIONAL, PBE_BJ, CUTOFF, grad=False)["energy"]
    assert abs(ew - conv) <= 5e-8
    assert abs(trunc - conv) >= 1e-6
    # default rank (tolerance 1e-4) and default cutoff from the damping radii
    ewd = O.ewald(num, xyz, lat, PBC, O.RATIONAL, PBE_BJ, CUTOFF, grad=False)["energy"]
    assert abs(ewd - conv) <= 1e-6


def test_ewald_converged_zero():
    num, xyz, lat = cubic(8.0, 1)
    ew = O.ewald(num, xyz, lat, PBC, O.ZERO, PBE_ZERO, LONG, grad=False, rank=10000, kcut=20.0)["energy"]
    conv = O.dispersion(num, xyz, lat, PBC, O.ZERO, PBE_ZERO, LONG, grad=False)["energy"]
    assert abs(ew - conv) <= 1e-8


def test_ewald_kcut_convergence():
    """test_fourier.f90 ewald-kcut-convergence."""
    num, xyz, lat = cubic(8.0, 1)
    eref = O.ewald(num, xyz, lat, PBC, O.RATIONAL, PBE_BJ, CUTOFF, grad=False, rank=10000, kcut=14.0)["energy"]
    last = np.inf
    for step in range(1, 5):
        e = O.ewald(num, xyz, lat, PBC, O.RATIONAL, PBE_BJ, CUTOFF, grad=False, rank=10000, kcut=2.0 * step)
        diff = abs(e["energy"] - eref)
        assert diff <= last
        last = diff
    assert last <= 1e-9


def test_ewald_supercell_and_translation():
    """test_fourier.f90 ewald-supercell, ewald-translation."""
    num, xyz, lat = cubic(8.0, 1)
    e1 = O.ewald(num, xyz, lat, PBC, O.RATIONAL, PBE_BJ, CUTOFF, grad=False, rank=10000, kcut=10.0)["energy"]
    n2, x2, l2 = cubic(8.0, 2)
    e2 = O.ewald(n2, x2, l2, PBC, O.RATIONAL, PBE_BJ, CUTOFF, grad=False, rank=10000, kcut=10.0)["energy"]
    assert e2 / len(n2) == pytest.approx(e1 / len(num), abs=1e-10)
    es = O.ewald(num, xyz + np.array([1.3, -2.7, 4.1]), lat, PBC, O.RATIONAL, PBE_BJ, CUTOFF, grad=False,
                 rank=10000, kcut=10.0)["energy"]
    assert es == pytest.approx(e1, abs=1e-12)


def _energy(num, xyz, lat, kind, par, **kw):
    return O.ewald(num, xyz, lat, PBC, kind, par, CUTOFF, grad=False, **kw)["energy"]


@pytest.mark.parametrize("kind,par,kw", [
    (O.RATIONAL, PBE_BJ, dict(rank=10000, kcut=10.0)),
    (O.ZERO, PBE_ZERO, dict(kcut=12.0)),
    (O.RATIONAL, dict(PBE_BJ, s9=1.0), dict(tolerance=1e-3)),
])
def test_ewald_gradient_fd(kind, par, kw):
    """test_fourier.f90 ewald-gradient: central differences, step 1e-4,
    within 1e-8; forces sum to zero.  The third case adds the three-body
    term and CN chain with a truncated expansion."""
    num, xyz, lat = cubic(8.0, 1)
    res = O.ewald(num, xyz, lat, PBC, kind, par, CUTOFF, **kw)
    step = 1e-4
    num_g = np.zeros_like(xyz)
    for i in range(len(num)):
        for c in range(3):
            xp = xyz.copy(); xp[i, c] += step
            xm = xyz.copy(); xm[i, c] -= step
            num_g[i, c] = (_energy(num, xp, lat, kind, par, **kw) - _energy(num, xm, lat, kind, par, **kw)) / (2 * step)
    assert np.abs(res["gradient"] - num_g).max() <= 1e-8
    assert np.abs(res["gradient"].sum(axis=0)).max() <= 1e-10


@pytest.mark.parametrize("kind,par,kw", [
    (O.RATIONAL, PBE_BJ, dict(rank=10000, kcut=10.0)),
    (O.ZERO, PBE_ZERO, dict(kcut=12.0)),
])
def test_ewald_sigma_fd(kind, par, kw):
    """test_fourier.f90 ewald-sigma: strain derivatives, step 1e-5, 1e-7."""
    num, xyz, lat = cubic(8.0, 1)
    res = O.ewald(num, xyz, lat, PBC, kind, par, CUTOFF, **kw)
    step = 1e-5
    num_s = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            ep = np.eye(3); ep[b, a] += step
            em = np.eye(3); em[b, a] -= step
            er = _energy(num, xyz @ ep.T, lat @ ep.T, kind, par, **kw)
            el = _energy(num, xyz @ em.T, lat @ em.T, kind, par, **kw)
            num_s[a, b] = (er - el) / (2 * step)
    # res["sigma"][b][a] = sigma(a, b)
    assert np.abs(res["sigma"].T - num_s).max() <= 1e-7


def test_ewald_random_cell_matches_converged():
    """A strained triclinic cell of four species, BJ with the three-body term:
    Ewald at full rank equals the real-space lattice sum extrapolated in the
    cutoff R with its -C6/r^6 tail, E(R) = E + A / R^3 (fit at 250 and 300
    bohr; each alone is 5e-7 and 3e-7 Eh short)."""
    num, xyz, lat = cell(7, 10, 11.0)
    p = dict(PBE_BJ, s9=1.0)
    ew = O.ewald(num, xyz, lat, PBC, O.RATIONAL, p, CUTOFF, grad=False, rank=10000, kcut=12.0)["energy"]
    e = {r: O.dispersion(num, xyz, lat, PBC, O.RATIONAL, p, (30.0, r, 15.0, 0.0, 0.0), grad=False)["energy"]
         for r in (250.0, 300.0)}
    einf = (e[300.0] * 300.0**3 - e[250.0] * 250.0**3) / (300.0**3 - 250.0**3)
    assert abs(ew - einf) <= 1e-9
    assert abs(ew - e[300.0]) >= 1e-7


def test_ewald_unsupported():
    """test_interface.py:300-318; zero damping needs an integer exponent."""
    num, xyz, lat = cubic(8.0, 1)
    with pytest.raises(O.EwaldUnsupported, match="does not support Ewald summation"):
        O.ewald(num, xyz, lat, PBC, O.MZERO, PBE_MZERO, CUTOFF, grad=False)
    with pytest.raises(O.EwaldUnsupported, match="does not support Ewald summation"):
        O.ewald(num, xyz, lat, PBC, O.ZERO, dict(PBE_ZERO, alp=13.5), CUTOFF, grad=False)
    # not 3D periodic: the real-space path serves every family
    O.ewald(num, xyz, lat, [True, True, False], O.MZERO, PBE_MZERO, CUTOFF, grad=False)


# ------------------------------------------------------------- particle mesh
def test_fft_against_numpy():
    """fft.f90: unnormalised radix-2 transform, exp(sign 2 pi i ...)."""
    rng = np.random.default_rng(3)
    g = rng.normal(size=(4, 8, 16)) + 1j * rng.normal(size=(4, 8, 16))
    np.testing.assert_allclose(O.fft3d(g, -1), np.fft.fftn(g), atol=1e-12)
    np.testing.assert_allclose(O.fft3d(g, +1), np.fft.ifftn(g) * g.size, atol=1e-12)


@pytest.mark.parametrize("w", [0.0, 0.13, 0.5, 0.77, 0.999])
def test_bspline_weights(w):
    """spme.f90:457-499: the order-6 weights are the cardinal B-spline M6 at
    w + 5 - j (partition of unity, closed form by the recursion on M_n) and
    the derivative weights are dM6/dw."""
    t, d = O.bspline(w)

    def m_n(x, n):
        if n == 2:
            return max(0.0, 1.0 - abs(x - 1.0))
        return (x * m_n(x, n - 1) + (n - x) * m_n(x - 1, n - 1)) / (n - 1)
    ref = [m_n(w + 5 - j, 6) for j in range(6)]
    np.testing.assert_allclose(t, ref, atol=1e-15)
    assert t.sum() == pytest.approx(1.0, abs=1e-15) and abs(d.sum()) <= 1e-15
    h = 1e-6
    if 0.0 < w - h and w + h < 1.0:
        np.testing.assert_allclose(d, (O.bspline(w + h)[0] - O.bspline(w - h)[0]) / (2 * h), atol=1e-8)


def test_spme_matches_direct_sum():
    """test_fourier.f90 spme-energy / spme-gradient: on the 2x2x2 supercell
    the particle mesh (64 points, and the mesh derived from the cutoff)
    reproduces the direct reciprocal sum: energy and gradient within 1e-8,
    virial within 1e-7."""
    num, xyz, lat = cubic(8.0, 2)
    ew = O.ewald(num, xyz, lat, PBC, O.RATIONAL, PBE_BJ, CUTOFF, rank=10000, kcut=10.0)
    for mesh in (64, 0):
        pm = O.ewald(num, xyz, lat, PBC, O.RATIONAL, PBE_BJ, CUTOFF, rank=10000, kcut=10.0, mesh=mesh)
        assert abs(pm["energy"] - ew["energy"]) <= 1e-8
        assert np.abs(pm["gradient"] - ew["gradient"]).max() <= 1e-8
        assert np.abs(pm["sigma"] - ew["sigma"]).max() <= 1e-7


def test_spme_default_mesh_python_case():
    """test_interface.py:283-296: the default (mesh from the cutoff) equals the
    direct sum within 1e-8 on the one-cell C/O case, kcut 10."""
    num, xyz, lat = cubic(8.0, 1)
    ew = O.ewald(num, xyz, lat, PBC, O.RATIONAL, PBE_BJ, CUTOFF, grad=False, kcut=10.0)["energy"]
    pm = O.ewald(num, xyz, lat, PBC, O.RATIONAL, PBE_BJ, CUTOFF, grad=False, kcut=10.0, mesh=0)["energy"]
    assert abs(pm - ew) <= 1e-8


@pytest.mark.parametrize("kind,par", [(O.RATIONAL, dict(PBE_BJ, s9=1.0)), (O.ZERO, PBE_ZERO)])
def test_spme_gradient_and_virial_fd(kind, par):
    """The mesh energy is a smooth function of the positions and the cell:
    its analytic gradient and virial against central differences (a
    strained 10-atom cell, mesh 16)."""
    num, xyz, lat = cell(7, 10, 11.0)
    kw = dict(rank=4, kcut=6.0, mesh=16)
    res = O.ewald(num, xyz, lat, PBC, kind, par, CUTOFF, **kw)
    e = lambda x, l: O.ewald(num, x, l, PBC, kind, par, CUTOFF, grad=False, **kw)["energy"]  # noqa: E731
    h = 1e-4
    g = np.zeros_like(xyz)
    for i in range(len(num)):
        for c in range(3):
            xp = xyz.copy(); xp[i, c] += h
            xm = xyz.copy(); xm[i, c] -= h
            g[i, c] = (e(xp, lat) - e(xm, lat)) / (2 * h)
    assert np.abs(res["gradient"] - g).max() <= 1e-8
    hs = 1e-5
    sg = np.zeros((3, 3))
    for a in range(3):
        for b in range(3):
            ep = np.eye(3); ep[b, a] += hs
            em = np.eye(3); em[b, a] -= hs
            sg[a, b] = (e(xyz @ ep.T, lat @ ep.T) - e(xyz @ em.T, lat @ em.T)) / (2 * hs)
    assert np.abs(res["sigma"].T - sg).max() <= 1e-7
