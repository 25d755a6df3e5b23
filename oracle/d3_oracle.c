/*
 * d3_oracle.c -- CPU oracle for the DFT-D3 energy / gradient / virial path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against and the CPU baseline timed beside it; the library never
 * links, loads or calls it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it.
 *
 * It is a plain FP64 C restatement of the loops of simple-dftd3 v1.6.0
 * (the reference), each function citing the reference file:line it follows.
 * Loop nests, the triangular index sets, the distance formulas and their
 * operand order follow the reference so that the set of pair / triple terms
 * is the same; OpenMP parallel loops with thread-private accumulators merged
 * in a critical section mirror the reference's shared-memory strategy.
 *
 * Pinning (see tests/test_oracle_golden.py): reference golden vectors
 * (test/unit/test_regression.f90 Ta90 3D, test/unit/test_periodic_2d.f90
 * gh185 and BP86 2D, test/validation/0{1,2,6,7}-*.json,
 * python/dftd3/test_interface.py energies) plus finite differences of the
 * energy for gradient and virial, closed forms (dimer CN, isolated atom C6,
 * weight normalisation) and brute-force enumeration of the lattice images.
 * The reciprocal-space part (low-rank C6, direct Ewald sum, particle mesh;
 * tests/test_oracle_ewald.py): closed-form and quadrature transforms,
 * LAPACK eigenpairs, the full-rank limit, converged real-space lattice sums,
 * numpy's FFT, the B-spline recursion, invariances and finite differences.
 *
 * Conventions: positions xyz[n][3] in bohr (Fortran xyz(3,n)), lattice
 * lat[3][3] with rows = lattice vectors (Fortran lattice(3,3) columns),
 * gradient[n][3], sigma[3][3] stored in Fortran order sigma(a,b) at
 * sigma[(b)*3 + (a)] -- i.e. the array the C API of the reference returns.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAX_ELEM 103
#define MAX_REF 7
#define NPAIR (MAX_ELEM * (MAX_ELEM + 1) / 2)
#define EPS DBL_EPSILON
/* Element (r, c) of an nat x nat matrix in Fortran (column-major) order, the
 * storage of the reference's c6(:,:) and dc6dcn(:,:) (disp.f90:135-136). */
#define FIDX(r, c) ((size_t)(r) + (size_t)(c) * (size_t)nat)

/* ------------------------------------------------------------------ tables */
static int g_nref[MAX_ELEM];
static double g_refcn[MAX_ELEM][MAX_REF];
static double *g_c6ref; /* [NPAIR][MAX_REF][MAX_REF] as [ic][jref][iref] */
static double g_r4r2raw[118];
static double g_vdwrad_aa[NPAIR];
static double g_covrad_aa[118];
static double g_aatoau;

/* Bounded timing samples: only outer atoms iat with iat % stride == 0 are
 * visited by the CN, two-body, ATM and chain-rule loops and the pairwise
 * two-body and ATM loops (1 = all).  The C6
 * matrix is always complete: the ATM loop skips triples with a vanishing
 * C6 (atm.f90:725-726), so a partial matrix would skip real work. */
static int g_sample_stride = 1;
void d3o_set_sample_stride(int stride) { g_sample_stride = stride > 0 ? stride : 1; }

/* partition.f90:57-148: the externally managed work partition of the pair,
 * triple and reciprocal-space loops (part of nparts; nparts = 1 keeps every
 * term).  owns_pair takes 0-based atom indices jat <= iat: the pair's index
 * (iat - 1) iat / 2 + jat - 1 of the 1-based reference, modulo nparts;
 * owns_index a 0-based loop index (the reference's idx - 1). */
static int g_part = 0, g_nparts = 1;
int d3o_set_partition(int part, int nparts)
{
    if (nparts <= 0 || part < 0 || part >= nparts) { g_part = 0; g_nparts = 1; return 1; }
    g_part = part;
    g_nparts = nparts;
    return 0;
}
static int owns_pair(int iat, int jat)
{
    if (g_nparts == 1) return 1;
    long long idx = (long long)iat * (iat + 1) / 2 + jat;
    return idx % g_nparts == g_part;
}
static int owns_index(long long idx) { return g_nparts == 1 || idx % g_nparts == g_part; }

/* Copy the element tables (data/d3_tables.npz) into the oracle. */
void d3o_set_tables(const int *nref, const double *refcn, const double *c6ref,
                    const double *r4_over_r2, const double *vdwrad_aa,
                    const double *covrad_aa, double aatoau)
{
    memcpy(g_nref, nref, sizeof g_nref);
    memcpy(g_refcn, refcn, sizeof g_refcn);
    free(g_c6ref);
    g_c6ref = (double *)malloc(sizeof(double) * NPAIR * MAX_REF * MAX_REF);
    memcpy(g_c6ref, c6ref, sizeof(double) * NPAIR * MAX_REF * MAX_REF);
    memcpy(g_r4r2raw, r4_over_r2, sizeof g_r4r2raw);
    memcpy(g_vdwrad_aa, vdwrad_aa, sizeof g_vdwrad_aa);
    memcpy(g_covrad_aa, covrad_aa, sizeof g_covrad_aa);
    g_aatoau = aatoau;
}

/* mctc-lib covalent_rad_d3 = 4/3 * covalent_rad_2009 (model.f90:125). */
double d3o_rcov(int z) { return 4.0 / 3.0 * (g_aatoau * g_covrad_aa[z - 1]); }

/* data/r4r2.f90:64-66,91-105: sqrt(0.5 * r4r2(Z) * sqrt(Z)). */
double d3o_r4r2(int z)
{
    if (z <= 0 || z > 118) return 0.0;
    return sqrt(0.5 * (g_r4r2raw[z - 1] * sqrt((double)z)));
}

/* data/vdwrad.f90:38,1163-1197 (only the in-table branch: Z <= 103). */
double d3o_rvdw(int z1, int z2)
{
    int a = z1 > z2 ? z1 : z2, b = z1 > z2 ? z2 : z1;
    return g_aatoau * g_vdwrad_aa[b + a * (a - 1) / 2 - 1];
}

/* reference.f90:160-171 get_c6(iref, jref, ati, atj), 1-based refs. */
double d3o_get_c6(int iref, int jref, int ati, int atj)
{
    int ic;
    if (ati > atj) {
        ic = atj + ati * (ati - 1) / 2;
        return g_c6ref[((size_t)(ic - 1) * MAX_REF + (jref - 1)) * MAX_REF + (iref - 1)];
    }
    ic = ati + atj * (atj - 1) / 2;
    return g_c6ref[((size_t)(ic - 1) * MAX_REF + (iref - 1)) * MAX_REF + (jref - 1)];
}

/* ------------------------------------------------------------ geometry */
static void cross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - b[1] * a[2];
    c[1] = a[2] * b[0] - b[2] * a[0];
    c[2] = a[0] * b[1] - b[0] * a[1];
}

static double norm2v(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* cutoff.f90:204-227 get_translations: rep from plane normals. */
static void get_translations(const double lat[3][3], double rthr, int rep[3])
{
    double n[3][3];
    cross(lat[1], lat[2], n[0]);
    cross(lat[2], lat[0], n[1]);
    cross(lat[0], lat[1], n[2]);
    for (int k = 0; k < 3; ++k) {
        double l = norm2v(n[k]);
        for (int c = 0; c < 3; ++c) n[k][c] = n[k][c] / l;
    }
    for (int k = 0; k < 3; ++k) {
        double cs = n[k][0] * lat[k][0] + n[k][1] * lat[k][1] + n[k][2] * lat[k][2];
        rep[k] = (int)ceil(fabs(rthr / cs));
    }
}

/*
 * cutoff.f90:140-186 get_lattice_points.  lat[k] is lattice vector k.
 * Writes at most maxtr translations [ntr][3], returns ntr (or the needed
 * count if it exceeds maxtr).  Loop order ix outer, iy, iz inner.
 */
int d3o_lattice_points(const int *periodic, const double *latflat, double rthr,
                       double *trans, int maxtr)
{
    double lat[3][3];
    int rep[3] = {0, 0, 0};
    int np = periodic[0] + periodic[1] + periodic[2];
    memcpy(lat, latflat, sizeof lat);
    if (np == 0) {
        if (maxtr >= 1) trans[0] = trans[1] = trans[2] = 0.0;
        return 1;
    }
    if (np == 3) {
        get_translations(lat, rthr, rep);
    } else if (np == 1) {
        for (int k = 0; k < 3; ++k)
            rep[k] = periodic[k] ? (int)ceil(rthr / norm2v(lat[k])) : 0;
    } else {
        /* get_normals, then where(spread(periodic, 2, 3)) norms = lat:
         * the mask selects Cartesian component c of every vector. */
        double nrm[3][3];
        for (int k = 0; k < 3; ++k) cross(lat[(k + 1) % 3], lat[(k + 2) % 3], nrm[k]);
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c)
                if (periodic[c]) nrm[k][c] = lat[k][c];
        get_translations(nrm, rthr, rep);
        for (int k = 0; k < 3; ++k)
            if (!periodic[k]) rep[k] = 0;
    }
    int ntr = (2 * rep[0] + 1) * (2 * rep[1] + 1) * (2 * rep[2] + 1);
    if (ntr > maxtr) return ntr;
    int itr = 0;
    for (int ix = -rep[0]; ix <= rep[0]; ++ix)
        for (int iy = -rep[1]; iy <= rep[1]; ++iy)
            for (int iz = -rep[2]; iz <= rep[2]; ++iz) {
                for (int c = 0; c < 3; ++c)
                    trans[3 * itr + c] = lat[0][c] * ix + lat[1][c] * iy + lat[2][c] * iz;
                ++itr;
            }
    return ntr;
}

/* mctc-lib matinv_3x3 (adjugate over determinant), used by utils.f90:38. */
static void matinv3(const double a[3][3], double b[3][3])
{
    /* a is Fortran lattice(3,3): a[col][row] in our row-vector storage */
    double m[3][3]; /* m[r][c] = Fortran a(r+1,c+1) */
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m[r][c] = a[c][r];
    double det = m[0][0] * m[1][1] * m[2][2] - m[0][0] * m[1][2] * m[2][1]
               - m[0][1] * m[1][0] * m[2][2] + m[0][1] * m[1][2] * m[2][0]
               + m[0][2] * m[1][0] * m[2][1] - m[0][2] * m[1][1] * m[2][0];
    double di = 1.0 / det;
    double inv[3][3];
    inv[0][0] = +di * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
    inv[1][0] = -di * (m[1][0] * m[2][2] - m[1][2] * m[2][0]);
    inv[2][0] = +di * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    inv[0][1] = -di * (m[0][1] * m[2][2] - m[0][2] * m[2][1]);
    inv[1][1] = +di * (m[0][0] * m[2][2] - m[0][2] * m[2][0]);
    inv[2][1] = -di * (m[0][0] * m[2][1] - m[0][1] * m[2][0]);
    inv[0][2] = +di * (m[0][1] * m[1][2] - m[0][2] * m[1][1]);
    inv[1][2] = -di * (m[0][0] * m[1][2] - m[0][2] * m[1][0]);
    inv[2][2] = +di * (m[0][0] * m[1][1] - m[0][1] * m[1][0]);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) b[r][c] = inv[r][c]; /* b[r][c] = inv(r,c) */
}

/* utils.f90:48-62 shift_back_abc. */
static double shift_back(double in)
{
    const double eps = 1.0e-14;
    double out = in;
    if (in < (0.0 - eps)) out = in + ceil(-in);
    if (in > (1.0 + eps)) out = in - floor(in);
    if (fabs(in - 1.0) < eps) out = in - 1.0;
    return out;
}

/* utils.f90:29-45 wrap_to_central_cell. */
void d3o_wrap(int nat, double *xyz, const double *latflat, const int *periodic)
{
    if (!(periodic[0] || periodic[1] || periodic[2])) return;
    double lat[3][3], inv[3][3];
    memcpy(lat, latflat, sizeof lat);
    matinv3(lat, inv);
    for (int i = 0; i < nat; ++i) {
        double *x = xyz + 3 * i, f[3];
        for (int r = 0; r < 3; ++r) f[r] = inv[r][0] * x[0] + inv[r][1] * x[1] + inv[r][2] * x[2];
        for (int r = 0; r < 3; ++r) f[r] = shift_back(f[r]);
        /* matmul(lattice, vec): x(r) = sum_c lattice(r,c) f(c) = sum_c lat[c][r] f(c) */
        for (int r = 0; r < 3; ++r) x[r] = lat[0][r] * f[0] + lat[1][r] * f[1] + lat[2][r] * f[2];
    }
}

/* api.f90:1517-1530 verify_structure: 1 if any pair is closer than 1e-9. */
int d3o_verify(int nat, const double *xyz)
{
    int stat = 0;
    for (int i = 0; i < nat; ++i)
        for (int j = 0; j < i; ++j) {
            double d[3] = {xyz[3 * j] - xyz[3 * i], xyz[3 * j + 1] - xyz[3 * i + 1],
                           xyz[3 * j + 2] - xyz[3 * i + 2]};
            if (norm2v(d) < 1.0e-9) ++stat;
        }
    return stat > 0;
}

/* cutoff.f90:73-99 smooth_cutoff. */
static void smooth_cutoff(double r, double cutoff, double width, double *sw, double *dswdr)
{
    if (width <= 0.0 || width >= cutoff) {
        *sw = 1.0; *dswdr = 0.0;
    } else {
        double inner = cutoff - width;
        if (r <= inner) { *sw = 1.0; *dswdr = 0.0; }
        else if (r >= cutoff) { *sw = 0.0; *dswdr = 0.0; }
        else {
            double x = (cutoff - r) / width;
            *sw = x * x * x * (10.0 + x * (-15.0 + 6.0 * x));
            *dswdr = -30.0 * x * x * (1.0 - x) * (1.0 - x) / width;
        }
    }
}

/* ------------------------------------------------------ coordination */
#define KCN 16.0

/* ncoord.f90:86-144 get_partitioned_coordination_number (no partition). */
void d3o_cn(int nat, const int *num, const double *xyz, int ntr, const double *trans,
            double cutoff, double *cn)
{
    double cutoff2 = cutoff * cutoff;
    memset(cn, 0, sizeof(double) * nat);
#pragma omp parallel
    {
        double *loc = (double *)calloc(nat, sizeof(double));
#pragma omp for schedule(static)
        for (int iat = 0; iat < nat; ++iat) {
            if (iat % g_sample_stride) continue; /* bounded timing sample */
            for (int jat = 0; jat <= iat; ++jat) {
                double rc = d3o_rcov(num[iat]) + d3o_rcov(num[jat]);
                for (int itr = 0; itr < ntr; ++itr) {
                    const double *t = trans + 3 * itr;
                    double v0 = xyz[3 * iat] - (xyz[3 * jat] + t[0]);
                    double v1 = xyz[3 * iat + 1] - (xyz[3 * jat + 1] + t[1]);
                    double v2 = xyz[3 * iat + 2] - (xyz[3 * jat + 2] + t[2]);
                    double r2 = v0 * v0 + v1 * v1 + v2 * v2;
                    if (r2 > cutoff2 || r2 < 1.0e-12) continue;
                    double r1 = sqrt(r2);
                    double countf = 1.0 / (1.0 + exp(-KCN * (rc / r1 - 1.0)));
                    loc[iat] += countf;
                    if (iat != jat) loc[jat] += countf;
                }
            }
        }
#pragma omp critical(d3o_cn_)
        for (int i = 0; i < nat; ++i) cn[i] += loc[i];
        free(loc);
    }
}

/* ncoord.f90:152-228 add_coordination_number_derivs. */
void d3o_cn_derivs(int nat, const int *num, const double *xyz, int ntr, const double *trans,
                   double cutoff, const double *dEdcn, double *gradient, double *sigma)
{
    double cutoff2 = cutoff * cutoff;
#pragma omp parallel
    {
        double *gl = (double *)calloc(3 * (size_t)nat, sizeof(double));
        double sl[9] = {0};
#pragma omp for schedule(static)
        for (int iat = 0; iat < nat; ++iat) {
            if (iat % g_sample_stride) continue; /* bounded timing sample */
            for (int jat = 0; jat <= iat; ++jat) {
                double rc = d3o_rcov(num[iat]) + d3o_rcov(num[jat]);
                for (int itr = 0; itr < ntr; ++itr) {
                    const double *t = trans + 3 * itr;
                    double v[3];
                    for (int c = 0; c < 3; ++c) v[c] = xyz[3 * iat + c] - (xyz[3 * jat + c] + t[c]);
                    double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                    if (r2 > cutoff2 || r2 < 1.0e-12) continue;
                    double r1 = sqrt(r2);
                    double expterm = exp(-KCN * (rc / r1 - 1.0));
                    double cf = 1.0 / (1.0 + expterm);
                    double dcf = -cf * (1.0 - cf) * KCN * rc / r2;
                    double cd[3];
                    for (int c = 0; c < 3; ++c) cd[c] = dcf * v[c] / r1;
                    double de = dEdcn[iat] + dEdcn[jat];
                    for (int c = 0; c < 3; ++c) {
                        gl[3 * iat + c] += cd[c] * de;
                        gl[3 * jat + c] -= cd[c] * de;
                    }
                    double ds = dEdcn[iat] + (jat != iat ? dEdcn[jat] : 0.0);
                    /* ds(a,b) = countd(b)*vec(a); sigma(a,b) at [b*3+a] */
                    for (int b = 0; b < 3; ++b)
                        for (int a = 0; a < 3; ++a) sl[b * 3 + a] += cd[b] * v[a] * ds;
                }
            }
        }
#pragma omp critical(d3o_cnd_)
        {
            for (int i = 0; i < 3 * nat; ++i) gradient[i] += gl[i];
            for (int k = 0; k < 9; ++k) sigma[k] += sl[k];
        }
        free(gl);
    }
}

/* ---------------------------------------------------------- weights */
static int is_exceptional(double v) { return isnan(v) || fabs(v) > DBL_MAX; }

/*
 * model.f90:201-324 weight_references, wf = 4 (model.f90:80).
 * gwvec/gwdcn are [nat][mref]; gwdcn may be NULL (energy-only branch).
 */
void d3o_weights(int nat, const int *num, const double *cn, int mref, double *gwvec, double *gwdcn)
{
    const double wf = 4.0, wf2 = 2.0 * wf;
    memset(gwvec, 0, sizeof(double) * nat * mref);
    if (gwdcn) memset(gwdcn, 0, sizeof(double) * nat * mref);
#pragma omp parallel for schedule(static)
    for (int iat = 0; iat < nat; ++iat) {
        int z = num[iat], nr = g_nref[z - 1];
        double gwt[MAX_REF], norm = 0.0, dnorm = 0.0, maxcn = -1.0e300;
        for (int r = 0; r < nr; ++r) {
            double d = cn[iat] - g_refcn[z - 1][r];
            double gw = exp(-wf * (d * d));
            gwt[r] = gw;
            double dcn = g_refcn[z - 1][r] - cn[iat];
            norm += gw;
            dnorm += wf2 * dcn * gw;
            if (g_refcn[z - 1][r] > maxcn) maxcn = g_refcn[z - 1][r];
        }
        norm = 1.0 / norm;
        for (int r = 0; r < nr; ++r) {
            double expw = gwt[r];
            double dcn = g_refcn[z - 1][r] - cn[iat];
            double expd = wf2 * dcn * expw;
            double gwk = expw * norm;
            if (is_exceptional(gwk)) gwk = (maxcn == g_refcn[z - 1][r]) ? 1.0 : 0.0;
            gwvec[iat * mref + r] = gwk;
            if (gwdcn) {
                double dgwk = expd * norm - expw * dnorm * (norm * norm);
                if (is_exceptional(dgwk)) dgwk = 0.0;
                gwdcn[iat * mref + r] = dgwk;
            }
        }
    }
}

/*
 * model.f90:340-491 get_atomic_c6, first-derivative and plain branches.
 * c6 and dc6dcn are nat x nat in Fortran order: c6[FIDX(i,j)] = c6(i,j)
 * and dc6dcn[FIDX(i,j)] = dc6dcn(i,j) = dC6_ij / dCN_i.
 */
void d3o_c6(int nat, const int *num, int mref, const double *gwvec, const double *gwdcn,
            const int *ghost, double *c6, double *dc6dcn)
{
    memset(c6, 0, sizeof(double) * nat * nat);
    if (dc6dcn) memset(dc6dcn, 0, sizeof(double) * nat * nat);
#pragma omp parallel for schedule(static)
    for (int iat = 0; iat < nat; ++iat) {
        if (ghost && ghost[iat]) continue;
        int zi = num[iat];
        for (int jat = 0; jat <= iat; ++jat) {
            if (ghost && ghost[jat]) continue;
            int zj = num[jat];
            double dc6 = 0.0, dci = 0.0, dcj = 0.0;
            for (int ir = 0; ir < g_nref[zi - 1]; ++ir)
                for (int jr = 0; jr < g_nref[zj - 1]; ++jr) {
                    double refc6 = d3o_get_c6(ir + 1, jr + 1, zi, zj);
                    dc6 += gwvec[iat * mref + ir] * gwvec[jat * mref + jr] * refc6;
                    if (dc6dcn) {
                        dci += gwdcn[iat * mref + ir] * gwvec[jat * mref + jr] * refc6;
                        dcj += gwvec[iat * mref + ir] * gwdcn[jat * mref + jr] * refc6;
                    }
                }
            c6[FIDX(iat, jat)] = dc6;
            c6[FIDX(jat, iat)] = dc6;
            if (dc6dcn) {
                dc6dcn[FIDX(iat, jat)] = dci;
                dc6dcn[FIDX(jat, iat)] = dcj;
            }
        }
    }
}

/* --------------------------------------------------------- two-body */
/* Damping families (kind): 0 rational (BJ, also the modified rational
 * parameters), 1 zero, 2 modified zero, 3 optimized power, 4 CSO, 5 Z. */
enum { K_RATIONAL = 0, K_ZERO = 1, K_MZERO = 2, K_OP = 3, K_CSO = 4, K_Z = 5 };
typedef struct {
    int kind;
    double s6, s8, s9; /* scaling */
    double a1, a2;     /* rational, optimized power, CSO, Z: a1, a2; zero, modified zero: rs6, rs8 */
    double alp;
    double bet;        /* modified zero, optimized power */
    double a3, a4;     /* CSO */
} d3o_param;

/*
 * Species identifiers of the structure (mctc-lib structure_type%id): the
 * elements numbered 1, 2, ... in the order of their first appearance.  Only
 * the Z damping uses them (z.f90: r0ij = a1 / (izp + jzp)).
 */
static void species_ids(int nat, const int *num, int *id)
{
    int nid = 0;
    for (int i = 0; i < nat; ++i) {
        id[i] = 0;
        for (int k = 0; k < i; ++k)
            if (num[k] == num[i]) { id[i] = id[k]; break; }
        if (!id[i]) id[i] = ++nid;
    }
}

/* The per-pair constants of one damping family (everything the image loop
 * of the reference hoists out of it). */
typedef struct {
    double rrij, c6ij;
    double r0ij, r0ij6, r0ij8;     /* rational, optimized power (with ab), CSO (d6^6 in r0ij6) */
    double rs6r0, rs8r0;           /* zero: rs6 R0, rs8 R0; modified zero: 1/(rs6 R0), 1/(rs8 R0) */
    double betr0, betr02rs6, betr02rs8;
    double ab;                     /* optimized power: r0ij**bet */
    double a2r0ij;                 /* CSO */
    double zr0;                    /* Z: a1 / (id_i + id_j) */
} pair_const;

static pair_const pair_setup(const d3o_param *p, int zi, int zj, int idi, int idj, double c6ij)
{
    pair_const c;
    memset(&c, 0, sizeof c);
    c.rrij = 3 * d3o_r4r2(zi) * d3o_r4r2(zj);
    c.c6ij = c6ij;
    switch (p->kind) {
    case K_RATIONAL: case K_OP: {  /* rational.f90:452-458, optimizedpower.f90 */
        c.r0ij = p->a1 * sqrt(c.rrij) + p->a2;
        double r0ij2 = c.r0ij * c.r0ij;
        c.r0ij6 = r0ij2 * r0ij2 * r0ij2;
        c.r0ij8 = c.r0ij6 * r0ij2;
        if (p->kind == K_OP) c.ab = pow(c.r0ij, p->bet);
        break;
    }
    case K_ZERO: {  /* zero.f90:484-488 */
        double r0 = d3o_rvdw(zj, zi);
        c.rs6r0 = p->a1 * r0;
        c.rs8r0 = p->a2 * r0;
        break;
    }
    case K_MZERO: {  /* mzero.f90:441-447 */
        double r0 = d3o_rvdw(zj, zi);
        c.r0ij = r0;
        c.rs6r0 = 1.0 / (p->a1 * r0);
        c.rs8r0 = 1.0 / (p->a2 * r0);
        c.betr0 = p->bet * r0;
        c.betr02rs6 = p->a1 * p->bet * r0 * r0;
        c.betr02rs8 = p->a2 * p->bet * r0 * r0;
        break;
    }
    case K_CSO: {  /* cso.f90: d6 = a3 sqrt(rrij) + a4 */
        c.r0ij = sqrt(c.rrij);
        double d6 = p->a3 * c.r0ij + p->a4;
        c.r0ij6 = pow(d6, 6);
        c.a2r0ij = p->a2 * c.r0ij;
        break;
    }
    case K_Z:  /* z.f90: r0ij = a1 / (izp + jzp) */
        c.zr0 = p->a1 / (double)(idi + idj);
        break;
    }
    return c;
}

/*
 * One pair term of the damped two-body energy without the switching
 * function: returns edisp0 (E = -c6 sw edisp0 / 2 per image); with grad
 * also gdisp0 (dE/dr / r, dG = -c6 gdisp vec) and, for the Z damping whose
 * damping depends on C6, dedc6 = d(c6 edisp0)/dc6 (z.f90 dEdc6 / -sw).
 * rational.f90:466-481, zero.f90:494-512, mzero.f90:456-470,
 * optimizedpower.f90 (t6/t8/d6/d8), cso.f90 (ef/sf/sig/t6), z.f90.
 */
static double pair_term(const d3o_param *p, const pair_const *c, double r2, double r1, int grad,
                        double *gdisp0, double *dedc6)
{
    double e0 = 0.0, g0 = 0.0, r4 = r2 * r2;
    *dedc6 = 0.0;
    switch (p->kind) {
    case K_RATIONAL: {
        double t6 = 1.0 / (r4 * r2 + c->r0ij6);
        double t8 = 1.0 / (r4 * r4 + c->r0ij8);
        e0 = p->s6 * t6 + p->s8 * c->rrij * t8;
        if (grad) {
            double d6 = -6 * r4 * t6 * t6;
            double d8 = -8 * r4 * r2 * t8 * t8;
            g0 = p->s6 * d6 + p->s8 * c->rrij * d8;
        }
        break;
    }
    case K_ZERO: {
        double alp6 = p->alp, alp8 = p->alp + 2.0;
        double r6 = r2 * r2 * r2, r8 = r6 * r2;
        double t6 = pow(c->rs6r0 / r1, alp6);
        double t8 = pow(c->rs8r0 / r1, alp8);
        double f6 = 1.0 / (1.0 + 6.0 * t6);
        double f8 = 1.0 / (1.0 + 6.0 * t8);
        e0 = p->s6 * f6 / r6 + p->s8 * c->rrij * f8 / r8;
        if (grad) {
            double d6 = -6.0 * f6 / r2 + 6.0 * alp6 * t6 * f6 * f6 / r2;
            double d8 = -8.0 * f8 / r2 + 6.0 * alp8 * t8 * f8 * f8 / r2;
            g0 = p->s6 * d6 / r6 + p->s8 * c->rrij * d8 / r8;
        }
        break;
    }
    case K_MZERO: {
        double alp6 = p->alp, alp8 = p->alp + 2.0;
        double r6 = r2 * r2 * r2, r8 = r6 * r2;
        double t6 = pow(r1 * c->rs6r0 + c->betr0, -alp6);
        double t8 = pow(r1 * c->rs8r0 + c->betr0, -alp8);
        double f6 = 1.0 / (1.0 + 6.0 * t6);
        double f8 = 1.0 / (1.0 + 6.0 * t8);
        e0 = p->s6 * f6 / r6 + p->s8 * c->rrij * f8 / r8;
        if (grad) {
            double d6 = -6.0 * f6 / r2 + 6.0 * alp6 * t6 * f6 * f6 / (r2 + c->betr02rs6 * r1);
            double d8 = -8.0 * f8 / r2 + 6.0 * alp8 * t8 * f8 * f8 / (r2 + c->betr02rs8 * r1);
            g0 = p->s6 * d6 / r6 + p->s8 * c->rrij * d8 / r8;
        }
        break;
    }
    case K_OP: {
        double abr0ij6 = c->ab * c->r0ij6, abr0ij8 = c->ab * c->r0ij8;
        double rb = pow(r2, p->bet * 0.5);
        double t6 = rb / (rb * r4 * r2 + abr0ij6);
        double t8 = rb / (rb * r4 * r4 + abr0ij8);
        e0 = p->s6 * t6 + p->s8 * c->rrij * t8;
        if (grad) {
            double d6 = -6 * r4 * t6 * t6 + p->bet * abr0ij6 * t6 * t6 / (rb * r2);
            double d8 = -8 * r4 * r2 * t8 * t8 + p->bet * abr0ij8 * t8 * t8 / (rb * r2);
            g0 = p->s6 * d6 + p->s8 * c->rrij * d8;
        }
        break;
    }
    case K_CSO: {
        double ef = exp(r1 - c->a2r0ij);
        double sf = 1.0 / (1.0 + ef);
        double sig = p->s6 + p->a1 * sf;
        double t6 = 1.0 / (r4 * r2 + c->r0ij6);
        e0 = sig * t6;
        if (grad) {
            double dsig = -p->a1 * sf * (1.0 - sf) / r1;
            double dt6 = -6 * r4 * t6 * t6;
            g0 = dsig * t6 + sig * dt6;
        }
        break;
    }
    case K_Z: {
        double t6 = 1.0 / (r4 * r2 + c->zr0 * c->c6ij);
        double t8 = 1.0 / (r4 * r4 + c->zr0 * c->c6ij * c->rrij);
        e0 = p->s6 * t6 + p->s8 * c->rrij * t8;
        if (grad) {
            double d6 = -6 * r4 * t6 * t6;
            double d8 = -8 * r4 * r2 * t8 * t8;
            g0 = p->s6 * d6 + p->s8 * c->rrij * d8;
            *dedc6 = p->s6 * r4 * r2 * t6 * t6 + p->s8 * c->rrij * r4 * r4 * t8 * t8;
        }
        break;
    }
    }
    *gdisp0 = g0;
    return e0;
}

/* get_dispersion2 returns early without any term (rational.f90:267 and the
 * same line of every family; CSO tests s6 and a1). */
static int two_body_active(const d3o_param *p)
{
    if (p->kind == K_CSO) return !(fabs(p->s6) < EPS && fabs(p->a1) < EPS);
    return !(fabs(p->s6) < EPS && fabs(p->s8) < EPS);
}

/*
 * get_dispersion2 of every family: rational.f90:220-514 (BJ), zero.f90:243-553,
 * mzero.f90:194-508, optimizedpower.f90, cso.f90, z.f90.  grad != 0 selects
 * the E+grad+sigma branch (get_dispersion_derivs), otherwise the energy
 * branch (get_dispersion_energy).  dE/dCN is -dC6/dCN edisp, for the Z
 * damping dC6/dCN * dEdc6 with dEdc6 = -sw d(c6 edisp0)/dc6 (z.f90).
 */
static void disp2(const d3o_param *p, int nat, const int *num, const double *xyz, int ntr,
                  const double *trans, double cutoff, double width, const double *c6,
                  const double *dc6dcn, double *energy, double *dEdcn, double *gradient,
                  double *sigma, int grad)
{
    if (!two_body_active(p)) return;
    double cutoff2 = cutoff * cutoff;
    int *id = (int *)malloc(sizeof(int) * (nat > 0 ? nat : 1));
    species_ids(nat, num, id);
#pragma omp parallel
    {
        double *el = (double *)calloc(nat, sizeof(double));
        double *dl = grad ? (double *)calloc(nat, sizeof(double)) : NULL;
        double *gl = grad ? (double *)calloc(3 * (size_t)nat, sizeof(double)) : NULL;
        double sl[9] = {0};
#pragma omp for schedule(static)
        for (int iat = 0; iat < nat; ++iat) {
            if (iat % g_sample_stride) continue; /* bounded timing sample */
            for (int jat = 0; jat <= iat; ++jat) {
                if (!owns_pair(iat, jat)) continue; /* rational.f90:453 */
                pair_const pc = pair_setup(p, num[iat], num[jat], id[iat], id[jat], c6[FIDX(jat, iat)]);
                double c6ij = pc.c6ij;
                for (int jtr = 0; jtr < ntr; ++jtr) {
                    const double *t = trans + 3 * jtr;
                    double v[3];
                    for (int c = 0; c < 3; ++c) v[c] = xyz[3 * iat + c] - (xyz[3 * jat + c] + t[c]);
                    double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                    if (r2 > cutoff2 || r2 < EPS) continue;
                    double r1 = sqrt(r2), sw, dswdr, gdisp0, dedc6;
                    smooth_cutoff(r1, cutoff, width, &sw, &dswdr);
                    double edisp0 = pair_term(p, &pc, r2, r1, grad, &gdisp0, &dedc6);
                    double edisp = sw * edisp0;
                    double dE = -c6ij * edisp * 0.5;
                    el[iat] += dE;
                    if (iat != jat) el[jat] += dE;
                    if (!grad) continue;
                    double gdisp = sw * gdisp0 + dswdr * edisp0 / r1;
                    double dG[3], dS[9];
                    for (int c = 0; c < 3; ++c) dG[c] = -c6ij * gdisp * v[c];
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) dS[b * 3 + a] = dG[a] * v[b] * 0.5;
                    /* dE/dC6 of the pair (z.f90 dEdc6; -edisp for the other families) */
                    double dEdc6 = p->kind == K_Z ? -sw * dedc6 : -edisp;
                    dl[iat] += dc6dcn[FIDX(iat, jat)] * dEdc6;
                    for (int k = 0; k < 9; ++k) sl[k] += dS[k];
                    if (iat != jat) {
                        dl[jat] += dc6dcn[FIDX(jat, iat)] * dEdc6;
                        for (int c = 0; c < 3; ++c) {
                            gl[3 * iat + c] += dG[c];
                            gl[3 * jat + c] -= dG[c];
                        }
                        for (int k = 0; k < 9; ++k) sl[k] += dS[k];
                    }
                }
            }
        }
#pragma omp critical(d3o_disp2_)
        {
            for (int i = 0; i < nat; ++i) energy[i] += el[i];
            if (grad) {
                for (int i = 0; i < nat; ++i) dEdcn[i] += dl[i];
                for (int i = 0; i < 3 * nat; ++i) gradient[i] += gl[i];
                for (int k = 0; k < 9; ++k) sigma[k] += sl[k];
            }
        }
        free(el); free(dl); free(gl);
    }
    free(id);
}

/* atm.f90:954-980 triple_scale (0-based indices). */
static double triple_scale(int ii, int jj, int kk)
{
    if (ii == jj) return (ii == kk) ? 1.0 / 6.0 : 0.5;
    if (ii != kk && jj != kk) return 1.0;
    return 0.5;
}

/*
 * atm.f90:433-825 get_atm_dispersion with rs9 = 4/3 and alp + 2
 * (rational.f90:518-566, zero.f90:557-604).  grad selects
 * get_atm_dispersion_derivs (atm.f90:625-825) or the energy branch
 * (atm.f90:500-621).
 */
static void atm(int nat, const int *num, const double *xyz, int ntr, const double *trans,
                double cutoff, double width, double s9, double rs9, double alp,
                const double *c6, const double *dc6dcn, double *energy, double *dEdcn,
                double *gradient, double *sigma, int grad)
{
    if (fabs(s9) < EPS) return;
    double cutoff2 = cutoff * cutoff, alp3 = alp / 3.0;
#pragma omp parallel
    {
        double *el = (double *)calloc(nat, sizeof(double));
        double *dl = grad ? (double *)calloc(nat, sizeof(double)) : NULL;
        double *gl = grad ? (double *)calloc(3 * (size_t)nat, sizeof(double)) : NULL;
        double sl[9] = {0};
#pragma omp for schedule(static)
        for (int iat = 0; iat < nat; ++iat) {
            if (iat % g_sample_stride) continue; /* bounded timing sample */
            int zi = num[iat];
            for (int jat = 0; jat <= iat; ++jat) {
                if (!owns_pair(iat, jat)) continue; /* atm.f90:709 */
                int zj = num[jat];
                double c6ij = c6[FIDX(jat, iat)];
                double r0ij = rs9 * d3o_rvdw(zj, zi);
                for (int jtr = 0; jtr < ntr; ++jtr) {
                    const double *tj = trans + 3 * jtr;
                    double vij[3];
                    for (int c = 0; c < 3; ++c) vij[c] = xyz[3 * jat + c] + tj[c] - xyz[3 * iat + c];
                    double r2ij = vij[0] * vij[0] + vij[1] * vij[1] + vij[2] * vij[2];
                    if (r2ij > cutoff2 || r2ij < EPS) continue;
                    double rij = sqrt(r2ij), swij, dswijdr;
                    smooth_cutoff(rij, cutoff, width, &swij, &dswijdr);
                    for (int kat = 0; kat <= jat; ++kat) {
                        int zk = num[kat];
                        double c6ik = c6[FIDX(kat, iat)];
                        double c6jk = c6[FIDX(kat, jat)];
                        if (grad && (fabs(c6ij) < EPS || fabs(c6ik) < EPS || fabs(c6jk) < EPS))
                            continue;
                        double c9 = -s9 * sqrt(fabs(c6ij * c6ik * c6jk));
                        double r0ik = rs9 * d3o_rvdw(zk, zi);
                        double r0jk = rs9 * d3o_rvdw(zk, zj);
                        double r0 = r0ij * r0ik * r0jk;
                        double triple = triple_scale(iat, jat, kat);
                        for (int ktr = 0; ktr < ntr; ++ktr) {
                            const double *tk = trans + 3 * ktr;
                            double vik[3], vjk[3];
                            for (int c = 0; c < 3; ++c) vik[c] = xyz[3 * kat + c] + tk[c] - xyz[3 * iat + c];
                            double r2ik = vik[0] * vik[0] + vik[1] * vik[1] + vik[2] * vik[2];
                            if (r2ik > cutoff2 || r2ik < EPS) continue;
                            double rik = sqrt(r2ik), swik, dswikdr;
                            smooth_cutoff(rik, cutoff, width, &swik, &dswikdr);
                            for (int c = 0; c < 3; ++c) vjk[c] = vik[c] - vij[c];
                            double r2jk = vjk[0] * vjk[0] + vjk[1] * vjk[1] + vjk[2] * vjk[2];
                            if (r2jk > cutoff2 || r2jk < EPS) continue;
                            double rjk = sqrt(r2jk), swjk, dswjkdr;
                            smooth_cutoff(rjk, cutoff, width, &swjk, &dswjkdr);
                            double sw = swij * swik * swjk;
                            double r2 = r2ij * r2ik * r2jk;
                            double r1 = sqrt(r2);
                            double r3 = r2 * r1;
                            double r5 = r3 * r2;
                            double fdmp = 1.0 / (1.0 + 6.0 * pow(r0 / r1, alp3));
                            double ang = 0.375 * (r2ij + r2jk - r2ik) * (r2ij - r2jk + r2ik)
                                       * (-r2ij + r2jk + r2ik) / r5 + 1.0 / r3;
                            double rr = ang * fdmp;
                            if (!grad) {
                                double dE = rr * c9 * triple * sw / 3.0;
                                el[iat] -= dE;
                                el[jat] -= dE;
                                el[kat] -= dE;
                                continue;
                            }
                            double r0r1alp3 = pow(r0 / r1, alp3);
                            double dfdmp = -2.0 * alp * r0r1alp3 * fdmp * fdmp;
                            double dang, dGij[3], dGik[3], dGjk[3];
                            dang = -0.375 * (r2ij * r2ij * r2ij + r2ij * r2ij * (r2jk + r2ik)
                                   + r2ij * (3.0 * r2jk * r2jk + 2.0 * r2jk * r2ik + 3.0 * r2ik * r2ik)
                                   - 5.0 * (r2jk - r2ik) * (r2jk - r2ik) * (r2jk + r2ik)) / r5;
                            double dE0 = rr * c9;
                            for (int c = 0; c < 3; ++c)
                                dGij[c] = sw * c9 * (-dang * fdmp + ang * dfdmp) / r2ij * vij[c]
                                        - dE0 * dswijdr / rij * swik * swjk * vij[c];
                            dang = -0.375 * (r2ik * r2ik * r2ik + r2ik * r2ik * (r2jk + r2ij)
                                   + r2ik * (3.0 * r2jk * r2jk + 2.0 * r2jk * r2ij + 3.0 * r2ij * r2ij)
                                   - 5.0 * (r2jk - r2ij) * (r2jk - r2ij) * (r2jk + r2ij)) / r5;
                            for (int c = 0; c < 3; ++c)
                                dGik[c] = sw * c9 * (-dang * fdmp + ang * dfdmp) / r2ik * vik[c]
                                        - dE0 * dswikdr / rik * swij * swjk * vik[c];
                            dang = -0.375 * (r2jk * r2jk * r2jk + r2jk * r2jk * (r2ik + r2ij)
                                   + r2jk * (3.0 * r2ik * r2ik + 2.0 * r2ik * r2ij + 3.0 * r2ij * r2ij)
                                   - 5.0 * (r2ik - r2ij) * (r2ik - r2ij) * (r2ik + r2ij)) / r5;
                            for (int c = 0; c < 3; ++c)
                                dGjk[c] = sw * c9 * (-dang * fdmp + ang * dfdmp) / r2jk * vjk[c]
                                        - dE0 * dswjkdr / rjk * swij * swik * vjk[c];
                            double dE = dE0 * triple * sw;
                            el[iat] -= dE / 3.0;
                            el[jat] -= dE / 3.0;
                            el[kat] -= dE / 3.0;
                            for (int c = 0; c < 3; ++c) {
                                gl[3 * iat + c] -= (dGij[c] + dGik[c]) * triple;
                                gl[3 * jat + c] += (dGij[c] - dGjk[c]) * triple;
                                gl[3 * kat + c] += (dGik[c] + dGjk[c]) * triple;
                            }
                            for (int a = 0; a < 3; ++a)
                                for (int b = 0; b < 3; ++b)
                                    sl[b * 3 + a] += (dGij[a] * vij[b] + dGik[a] * vik[b]
                                                      + dGjk[a] * vjk[b]) * triple;
                            dl[iat] -= dE * 0.5 * (dc6dcn[FIDX(iat, jat)] / c6ij
                                                   + dc6dcn[FIDX(iat, kat)] / c6ik);
                            dl[jat] -= dE * 0.5 * (dc6dcn[FIDX(jat, iat)] / c6ij
                                                   + dc6dcn[FIDX(jat, kat)] / c6jk);
                            dl[kat] -= dE * 0.5 * (dc6dcn[FIDX(kat, iat)] / c6ik
                                                   + dc6dcn[FIDX(kat, jat)] / c6jk);
                        }
                    }
                }
            }
        }
#pragma omp critical(d3o_atm_)
        {
            for (int i = 0; i < nat; ++i) energy[i] += el[i];
            if (grad) {
                for (int i = 0; i < nat; ++i) dEdcn[i] += dl[i];
                for (int i = 0; i < 3 * nat; ++i) gradient[i] += gl[i];
                for (int k = 0; k < 9; ++k) sigma[k] += sl[k];
            }
        }
        free(el); free(dl); free(gl);
    }
}

/* -------------------------------------------------------- pairwise */
/*
 * get_pairwise_dispersion2 of every family (rational.f90:623-705,
 * zero.f90:662-755, mzero.f90:617-707, optimizedpower.f90, cso.f90, z.f90):
 * the two-body energy of every pair (iat, jat <= iat) summed over the images,
 * half of each term on (jat, iat) and half on (iat, jat) (once on the
 * diagonal).  energy is nat x nat in Fortran order (FIDX).  Unlike
 * get_dispersion2 there is no early return for vanishing scalings.
 */
static void disp2_pairwise(const d3o_param *p, int nat, const int *num, const double *xyz, int ntr,
                           const double *trans, double cutoff, double width, const double *c6,
                           double *energy)
{
    double cutoff2 = cutoff * cutoff;
    int *id = (int *)malloc(sizeof(int) * (nat > 0 ? nat : 1));
    species_ids(nat, num, id);
#pragma omp parallel
    {
        double *el = (double *)calloc((size_t)nat * nat, sizeof(double));
#pragma omp for schedule(static)
        for (int iat = 0; iat < nat; ++iat) {
            if (iat % g_sample_stride) continue; /* bounded timing sample */
            for (int jat = 0; jat <= iat; ++jat) {
                pair_const pc = pair_setup(p, num[iat], num[jat], id[iat], id[jat], c6[FIDX(jat, iat)]);
                for (int jtr = 0; jtr < ntr; ++jtr) {
                    const double *t = trans + 3 * jtr;
                    double v[3];
                    for (int c = 0; c < 3; ++c) v[c] = xyz[3 * iat + c] - (xyz[3 * jat + c] + t[c]);
                    double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                    if (r2 > cutoff2 || r2 < EPS) continue;
                    double r1 = sqrt(r2), sw, dswdr, g0, d0;
                    smooth_cutoff(r1, cutoff, width, &sw, &dswdr);
                    double edisp = sw * pair_term(p, &pc, r2, r1, 0, &g0, &d0);
                    double dE = -pc.c6ij * edisp * 0.5;
                    el[FIDX(jat, iat)] += dE;
                    if (iat != jat) el[FIDX(iat, jat)] += dE;
                }
            }
        }
#pragma omp critical(d3o_disp2_pairwise_)
        for (size_t k = 0; k < (size_t)nat * nat; ++k) energy[k] += el[k];
        free(el);
    }
    free(id);
}

/*
 * atm.f90:829-950 get_atm_pairwise_dispersion: every triple term, divided
 * by six, on the six ordered pairs of its atoms.  No C6 screening (unlike
 * the gradient branch).
 */
static void atm_pairwise(int nat, const int *num, const double *xyz, int ntr, const double *trans,
                         double cutoff, double width, double s9, double rs9, double alp,
                         const double *c6, double *energy)
{
    if (fabs(s9) < EPS) return;
    double cutoff2 = cutoff * cutoff, alp3 = alp / 3.0;
#pragma omp parallel
    {
        double *el = (double *)calloc((size_t)nat * nat, sizeof(double));
#pragma omp for schedule(static)
        for (int iat = 0; iat < nat; ++iat) {
            if (iat % g_sample_stride) continue; /* bounded timing sample */
            int zi = num[iat];
            for (int jat = 0; jat <= iat; ++jat) {
                int zj = num[jat];
                double c6ij = c6[FIDX(jat, iat)];
                double r0ij = rs9 * d3o_rvdw(zj, zi);
                for (int jtr = 0; jtr < ntr; ++jtr) {
                    const double *tj = trans + 3 * jtr;
                    double vij[3];
                    for (int c = 0; c < 3; ++c) vij[c] = xyz[3 * jat + c] + tj[c] - xyz[3 * iat + c];
                    double r2ij = vij[0] * vij[0] + vij[1] * vij[1] + vij[2] * vij[2];
                    if (r2ij > cutoff2 || r2ij < EPS) continue;
                    double rij = sqrt(r2ij), swij, dsw;
                    smooth_cutoff(rij, cutoff, width, &swij, &dsw);
                    for (int kat = 0; kat <= jat; ++kat) {
                        int zk = num[kat];
                        double c6ik = c6[FIDX(kat, iat)];
                        double c6jk = c6[FIDX(kat, jat)];
                        double c9 = -s9 * sqrt(fabs(c6ij * c6ik * c6jk));
                        double r0 = r0ij * (rs9 * d3o_rvdw(zk, zi)) * (rs9 * d3o_rvdw(zk, zj));
                        double triple = triple_scale(iat, jat, kat);
                        for (int ktr = 0; ktr < ntr; ++ktr) {
                            const double *tk = trans + 3 * ktr;
                            double vik[3], vjk[3];
                            for (int c = 0; c < 3; ++c) vik[c] = xyz[3 * kat + c] + tk[c] - xyz[3 * iat + c];
                            double r2ik = vik[0] * vik[0] + vik[1] * vik[1] + vik[2] * vik[2];
                            if (r2ik > cutoff2 || r2ik < EPS) continue;
                            double rik = sqrt(r2ik), swik;
                            smooth_cutoff(rik, cutoff, width, &swik, &dsw);
                            for (int c = 0; c < 3; ++c) vjk[c] = vik[c] - vij[c];
                            double r2jk = vjk[0] * vjk[0] + vjk[1] * vjk[1] + vjk[2] * vjk[2];
                            if (r2jk > cutoff2 || r2jk < EPS) continue;
                            double rjk = sqrt(r2jk), swjk;
                            smooth_cutoff(rjk, cutoff, width, &swjk, &dsw);
                            double sw = swij * swik * swjk;
                            double r2 = r2ij * r2ik * r2jk;
                            double r1 = sqrt(r2), r3 = r2 * r1, r5 = r3 * r2;
                            double fdmp = 1.0 / (1.0 + 6.0 * pow(r0 / r1, alp3));
                            double ang = 0.375 * (r2ij + r2jk - r2ik) * (r2ij - r2jk + r2ik)
                                       * (-r2ij + r2jk + r2ik) / r5 + 1.0 / r3;
                            double dE = ang * fdmp * c9 * triple * sw / 6.0;
                            el[FIDX(jat, iat)] -= dE;
                            el[FIDX(kat, iat)] -= dE;
                            el[FIDX(iat, jat)] -= dE;
                            el[FIDX(kat, jat)] -= dE;
                            el[FIDX(iat, kat)] -= dE;
                            el[FIDX(jat, kat)] -= dE;
                        }
                    }
                }
            }
        }
#pragma omp critical(d3o_atm_pairwise_)
        for (size_t k = 0; k < (size_t)nat * nat; ++k) energy[k] += el[k];
        free(el);
    }
}

/* ------------------------------------------------------- orchestrator */
static int lattice_alloc(const int *periodic, const double *lat, double rthr, double **trans)
{
    int n = d3o_lattice_points(periodic, lat, rthr, NULL, 0);
    *trans = (double *)malloc(sizeof(double) * 3 * n);
    d3o_lattice_points(periodic, lat, rthr, *trans, n);
    return n;
}

/*
 * disp.f90:56-185 get_dispersion_atomic_v2 followed by disp.f90:352-399
 * (energy = sum of atomic energies).  par = {s6, s8, s9, a1|rs6, a2|rs8, alp,
 * bet, a3, a4} (see d3o_param),
 * cut = {cn, disp2, disp3, width2, width3}.  gradient / sigma may be NULL
 * (grad = present(gradient) .or. present(sigma), disp.f90:102).
 * Optional intermediates (may be NULL): cn_out[nat], dEdcn_out[nat],
 * energies_out[nat].  Returns 0 on success.
 */
int d3o_get_dispersion(int nat, const int *num, const double *xyz, const double *lattice,
                       const int *periodic, int kind, const double *par, const double *cut,
                       const int *ghost, double *energy, double *gradient, double *sigma,
                       double *cn_out, double *dEdcn_out, double *energies_out)
{
    d3o_param p = {kind, par[0], par[1], par[2], par[3], par[4], par[5], par[6], par[7], par[8]};
    int grad = (gradient != NULL) || (sigma != NULL);
    int mref = 0;
    for (int i = 0; i < nat; ++i)
        if (g_nref[num[i] - 1] > mref) mref = g_nref[num[i] - 1];
    double zero_lat[9] = {0};
    int zero_pbc[3] = {0, 0, 0};
    if (!lattice) lattice = zero_lat;
    if (!periodic) periodic = zero_pbc;

    double *cn = (double *)calloc(nat, sizeof(double));
    double *trans;
    int ntr = lattice_alloc(periodic, lattice, cut[0], &trans);
    d3o_cn(nat, num, xyz, ntr, trans, cut[0], cn);
    free(trans);

    double *gw = (double *)calloc((size_t)nat * mref, sizeof(double));
    double *gwd = grad ? (double *)calloc((size_t)nat * mref, sizeof(double)) : NULL;
    d3o_weights(nat, num, cn, mref, gw, gwd);
    double *c6 = (double *)calloc((size_t)nat * nat, sizeof(double));
    double *dc6 = grad ? (double *)calloc((size_t)nat * nat, sizeof(double)) : NULL;
    d3o_c6(nat, num, mref, gw, gwd, ghost, c6, dc6);

    double *en = (double *)calloc(nat, sizeof(double));
    double *dEdcn = grad ? (double *)calloc(nat, sizeof(double)) : NULL;
    double *gl = grad ? (double *)calloc(3 * (size_t)nat, sizeof(double)) : NULL;
    double sl[9] = {0};

    ntr = lattice_alloc(periodic, lattice, cut[1], &trans);
    disp2(&p, nat, num, xyz, ntr, trans, cut[1], cut[3], c6, dc6, en, dEdcn, gl, sl, grad);
    free(trans);
    ntr = lattice_alloc(periodic, lattice, cut[2], &trans);
    atm(nat, num, xyz, ntr, trans, cut[2], cut[4], p.s9, 4.0 / 3.0, p.alp + 2.0, c6, dc6, en,
        dEdcn, gl, sl, grad);
    free(trans);
    if (grad) {
        ntr = lattice_alloc(periodic, lattice, cut[0], &trans);
        d3o_cn_derivs(nat, num, xyz, ntr, trans, cut[0], dEdcn, gl, sl);
        free(trans);
        if (gradient) memcpy(gradient, gl, sizeof(double) * 3 * nat);
        if (sigma) memcpy(sigma, sl, sizeof sl);
    }
    double e = 0.0;
    for (int i = 0; i < nat; ++i) e += en[i];
    *energy = e;
    if (cn_out) memcpy(cn_out, cn, sizeof(double) * nat);
    if (dEdcn_out && grad) memcpy(dEdcn_out, dEdcn, sizeof(double) * nat);
    if (energies_out) memcpy(energies_out, en, sizeof(double) * nat);
    free(cn); free(gw); free(gwd); free(c6); free(dc6); free(en); free(dEdcn); free(gl);
    return 0;
}

/*
 * disp.f90:403-457 get_pairwise_dispersion_v2: CN, weights and C6 as for the
 * energy, then the pairwise two-body and three-body energies.  pair2 and
 * pair3 are nat x nat (Fortran order; both are symmetric).
 */
int d3o_get_pairwise_dispersion(int nat, const int *num, const double *xyz, const double *lattice,
                                const int *periodic, int kind, const double *par, const double *cut,
                                const int *ghost, double *pair2, double *pair3)
{
    d3o_param p = {kind, par[0], par[1], par[2], par[3], par[4], par[5], par[6], par[7], par[8]};
    int mref = 0;
    for (int i = 0; i < nat; ++i)
        if (g_nref[num[i] - 1] > mref) mref = g_nref[num[i] - 1];
    double zero_lat[9] = {0};
    int zero_pbc[3] = {0, 0, 0};
    if (!lattice) lattice = zero_lat;
    if (!periodic) periodic = zero_pbc;

    double *cn = (double *)calloc(nat, sizeof(double));
    double *trans;
    int ntr = lattice_alloc(periodic, lattice, cut[0], &trans);
    d3o_cn(nat, num, xyz, ntr, trans, cut[0], cn);
    free(trans);
    double *gw = (double *)calloc((size_t)nat * mref, sizeof(double));
    d3o_weights(nat, num, cn, mref, gw, NULL);
    double *c6 = (double *)calloc((size_t)nat * nat, sizeof(double));
    d3o_c6(nat, num, mref, gw, NULL, ghost, c6, NULL);

    memset(pair2, 0, sizeof(double) * nat * nat);
    memset(pair3, 0, sizeof(double) * nat * nat);
    ntr = lattice_alloc(periodic, lattice, cut[1], &trans);
    disp2_pairwise(&p, nat, num, xyz, ntr, trans, cut[1], cut[3], c6, pair2);
    free(trans);
    ntr = lattice_alloc(periodic, lattice, cut[2], &trans);
    atm_pairwise(nat, num, xyz, ntr, trans, cut[2], cut[4], p.s9, 4.0 / 3.0, p.alp + 2.0, c6, pair3);
    free(trans);
    free(cn); free(gw); free(c6);
    return 0;
}

/* --------------------------------------------------------- Hessian */
/* cutoff.f90:104-137 smooth_cutoff_r2: switching function and its first and
 * second derivative with respect to r^2. */
static void smooth_cutoff_r2(double r2, double cutoff, double width, double *sw, double *dsw, double *d2sw)
{
    double r1 = sqrt(r2), dswdr = 0.0, d2swdr2 = 0.0;
    if (width <= 0.0 || width >= cutoff) {
        *sw = 1.0;
    } else {
        double inner = cutoff - width;
        if (r1 <= inner) {
            *sw = 1.0;
        } else if (r1 >= cutoff) {
            *sw = 0.0;
        } else {
            double x = (cutoff - r1) / width;
            *sw = x * x * x * (10.0 + x * (-15.0 + 6.0 * x));
            dswdr = -30.0 * x * x * (1.0 - x) * (1.0 - x) / width;
            d2swdr2 = 60.0 * x * (1.0 - x) * (1.0 - 2.0 * x) / (width * width);
        }
    }
    *dsw = 0.5 * dswdr / r1;
    *d2sw = 0.25 * (d2swdr2 - dswdr / r1) / r2;
}

/*
 * get_damping_kernel of every family (rational.f90, zero.f90, mzero.f90,
 * optimizedpower.f90, cso.f90, z.f90): the pair energy e(u, C6) with u = r^2
 * and its derivatives e_u, e_c, e_uu, e_uc, e_cc.
 */
static void damping_kernel(const d3o_param *p, int zi, int zj, int idi, int idj, double u, double c6,
                           double *e, double *er, double *ec, double *err, double *erc, double *ecc)
{
    *e = *er = *ec = *err = *erc = *ecc = 0.0;
    if (p->kind != K_CSO && fabs(p->s6) < EPS && fabs(p->s8) < EPS) return;
    double rrij = 3 * d3o_r4r2(zi) * d3o_r4r2(zj);
    double pv = 0.0, pd = 0.0, pdd = 0.0;
    switch (p->kind) {
    case K_RATIONAL: {
        double r0ij = p->a1 * sqrt(rrij) + p->a2;
        double r0ij2 = r0ij * r0ij, r0ij6 = r0ij2 * r0ij2 * r0ij2, r0ij8 = r0ij6 * r0ij2;
        double t6 = 1.0 / (u * u * u + r0ij6);
        double t8 = 1.0 / (u * u * u * u + r0ij8);
        double t6d = -3.0 * u * u * t6 * t6;
        double t8d = -4.0 * u * u * u * t8 * t8;
        double t6dd = -6.0 * u * t6 * t6 + 18.0 * pow(u, 4) * pow(t6, 3);
        double t8dd = -12.0 * u * u * t8 * t8 + 32.0 * pow(u, 6) * pow(t8, 3);
        pv = p->s6 * t6 + p->s8 * rrij * t8;
        pd = p->s6 * t6d + p->s8 * rrij * t8d;
        pdd = p->s6 * t6dd + p->s8 * rrij * t8dd;
        break;
    }
    case K_ZERO: case K_MZERO: {
        double r0ij = d3o_rvdw(zj, zi);
        double alp6 = p->alp, alp8 = p->alp + 2.0;
        double t6, t8, t6d, t8d, t6dd, t8dd;
        if (p->kind == K_ZERO) {
            double a6 = 0.5 * alp6, a8 = 0.5 * alp8;
            t6 = pow(p->a1 * r0ij / sqrt(u), alp6);
            t8 = pow(p->a2 * r0ij / sqrt(u), alp8);
            t6d = -a6 * t6 / u;
            t8d = -a8 * t8 / u;
            t6dd = a6 * (a6 + 1.0) * t6 / (u * u);
            t8dd = a8 * (a8 + 1.0) * t8 / (u * u);
        } else {
            double r1 = sqrt(u), betr0 = p->bet * r0ij;
            double w6 = r1 / (p->a1 * r0ij) + betr0, w8 = r1 / (p->a2 * r0ij) + betr0;
            double w6d = 1.0 / (p->a1 * r0ij * 2.0 * r1), w8d = 1.0 / (p->a2 * r0ij * 2.0 * r1);
            double w6dd = -1.0 / (p->a1 * r0ij * 4.0 * u * r1), w8dd = -1.0 / (p->a2 * r0ij * 4.0 * u * r1);
            t6 = pow(w6, -alp6);
            t8 = pow(w8, -alp8);
            t6d = -alp6 * pow(w6, -alp6 - 1.0) * w6d;
            t8d = -alp8 * pow(w8, -alp8 - 1.0) * w8d;
            t6dd = alp6 * (alp6 + 1.0) * pow(w6, -alp6 - 2.0) * w6d * w6d - alp6 * pow(w6, -alp6 - 1.0) * w6dd;
            t8dd = alp8 * (alp8 + 1.0) * pow(w8, -alp8 - 2.0) * w8d * w8d - alp8 * pow(w8, -alp8 - 1.0) * w8dd;
        }
        double f6 = 1.0 / (1.0 + 6.0 * t6), f8 = 1.0 / (1.0 + 6.0 * t8);
        double f6d = -6.0 * t6d * f6 * f6, f8d = -6.0 * t8d * f8 * f8;
        double f6dd = -6.0 * t6dd * f6 * f6 + 72.0 * t6d * t6d * pow(f6, 3);
        double f8dd = -6.0 * t8dd * f8 * f8 + 72.0 * t8d * t8d * pow(f8, 3);
        double c6t = f6 / pow(u, 3);
        double c6td = f6d / pow(u, 3) - 3.0 * f6 / pow(u, 4);
        double c6tdd = f6dd / pow(u, 3) - 6.0 * f6d / pow(u, 4) + 12.0 * f6 / pow(u, 5);
        double c8t = f8 / pow(u, 4);
        double c8td = f8d / pow(u, 4) - 4.0 * f8 / pow(u, 5);
        double c8tdd = f8dd / pow(u, 4) - 8.0 * f8d / pow(u, 5) + 20.0 * f8 / pow(u, 6);
        pv = p->s6 * c6t + p->s8 * rrij * c8t;
        pd = p->s6 * c6td + p->s8 * rrij * c8td;
        pdd = p->s6 * c6tdd + p->s8 * rrij * c8tdd;
        break;
    }
    case K_OP: {
        double r0ij = p->a1 * sqrt(rrij) + p->a2;
        double r0ij2 = r0ij * r0ij, r0ij6 = r0ij2 * r0ij2 * r0ij2, r0ij8 = r0ij6 * r0ij2;
        double ab = pow(r0ij, p->bet), abr0ij6 = ab * r0ij6, abr0ij8 = ab * r0ij8;
        double q = 0.5 * p->bet;
        double d6 = pow(u, 3) + abr0ij6 * pow(u, -q);
        double d8 = pow(u, 4) + abr0ij8 * pow(u, -q);
        double d6d = 3.0 * u * u - q * abr0ij6 * pow(u, -q - 1.0);
        double d8d = 4.0 * pow(u, 3) - q * abr0ij8 * pow(u, -q - 1.0);
        double d6dd = 6.0 * u + q * (q + 1.0) * abr0ij6 * pow(u, -q - 2.0);
        double d8dd = 12.0 * u * u + q * (q + 1.0) * abr0ij8 * pow(u, -q - 2.0);
        double t6 = 1.0 / d6, t8 = 1.0 / d8;
        double t6d = -d6d * t6 * t6, t8d = -d8d * t8 * t8;
        double t6dd = -d6dd * t6 * t6 + 2.0 * d6d * d6d * pow(t6, 3);
        double t8dd = -d8dd * t8 * t8 + 2.0 * d8d * d8d * pow(t8, 3);
        pv = p->s6 * t6 + p->s8 * rrij * t8;
        pd = p->s6 * t6d + p->s8 * rrij * t8d;
        pdd = p->s6 * t6dd + p->s8 * rrij * t8dd;
        break;
    }
    case K_CSO: {
        double r1 = sqrt(u), r0ij = sqrt(rrij);
        double dd6 = p->a3 * r0ij + p->a4, dd6_6 = pow(dd6, 6);
        double ef = exp(r1 - p->a2 * r0ij);
        double sf = 1.0 / (1.0 + ef);
        double sfd = -sf * (1.0 - sf), sfdd = sf * (1.0 - sf) * (1.0 - 2.0 * sf);
        double sig = p->s6 + p->a1 * sf, sigd = p->a1 * sfd, sigdd = p->a1 * sfdd;
        double sigu = 0.5 * sigd / r1;
        double siguu = 0.25 * sigdd / u - 0.25 * sigd / (u * r1);
        double t6 = 1.0 / (u * u * u + dd6_6);
        double t6d = -3.0 * u * u * t6 * t6;
        double t6dd = -6.0 * u * t6 * t6 + 18.0 * pow(u, 4) * pow(t6, 3);
        pv = sig * t6;
        pd = sigu * t6 + sig * t6d;
        pdd = siguu * t6 + 2.0 * sigu * t6d + sig * t6dd;
        break;
    }
    case K_Z: {
        double r0ij = p->a1 / (double)(idi + idj), b8 = r0ij * rrij;
        double t6 = 1.0 / (u * u * u + r0ij * c6);
        double t8 = 1.0 / (u * u * u * u + b8 * c6);
        double t6u = -3.0 * u * u * t6 * t6, t8u = -4.0 * pow(u, 3) * t8 * t8;
        double t6uu = -6.0 * u * t6 * t6 + 18.0 * pow(u, 4) * pow(t6, 3);
        double t8uu = -12.0 * u * u * t8 * t8 + 32.0 * pow(u, 6) * pow(t8, 3);
        double t6c = -r0ij * t6 * t6, t8c = -b8 * t8 * t8;
        double t6uc = 6.0 * u * u * r0ij * pow(t6, 3), t8uc = 8.0 * pow(u, 3) * b8 * pow(t8, 3);
        double t6cc = 2.0 * r0ij * r0ij * pow(t6, 3), t8cc = 2.0 * b8 * b8 * pow(t8, 3);
        double pz = p->s6 * t6 + p->s8 * rrij * t8;
        double pu = p->s6 * t6u + p->s8 * rrij * t8u;
        double puu = p->s6 * t6uu + p->s8 * rrij * t8uu;
        double pc = p->s6 * t6c + p->s8 * rrij * t8c;
        double puc = p->s6 * t6uc + p->s8 * rrij * t8uc;
        double pcc = p->s6 * t6cc + p->s8 * rrij * t8cc;
        *e = -c6 * pz;
        *er = -c6 * pu;
        *ec = -pz - c6 * pc;
        *err = -c6 * puu;
        *erc = -pu - c6 * puc;
        *ecc = -2.0 * pc - c6 * pcc;
        return;
    }
    }
    *e = -c6 * pv;
    *er = -c6 * pd;
    *ec = -pv;
    *err = -c6 * pdd;
    *erc = -pd;
}

/* model.f90:201-287 weight_references with the second derivative (gwd2cn). */
static void weights2(int nat, const int *num, const double *cn, int mref, double *gw, double *gwd, double *gwd2)
{
    const double wf2 = 8.0;
    memset(gw, 0, sizeof(double) * nat * mref);
    memset(gwd, 0, sizeof(double) * nat * mref);
    memset(gwd2, 0, sizeof(double) * nat * mref);
    for (int iat = 0; iat < nat; ++iat) {
        int z = num[iat], nr = g_nref[z - 1];
        double gwt[MAX_REF], norm = 0.0, dnorm = 0.0, d2norm = 0.0, maxcn = -1.0e300;
        for (int r = 0; r < nr; ++r) {
            double d = cn[iat] - g_refcn[z - 1][r];
            double w = exp(-4.0 * (d * d));
            gwt[r] = w;
            double dcn = g_refcn[z - 1][r] - cn[iat];
            norm += w;
            dnorm += wf2 * dcn * w;
            d2norm += (wf2 * wf2 * dcn * dcn - wf2) * w;
            if (g_refcn[z - 1][r] > maxcn) maxcn = g_refcn[z - 1][r];
        }
        norm = 1.0 / norm;
        for (int r = 0; r < nr; ++r) {
            double expw = gwt[r], dcn = g_refcn[z - 1][r] - cn[iat];
            double expd = wf2 * dcn * expw, expdd = (wf2 * wf2 * dcn * dcn - wf2) * expw;
            double gwk = expw * norm;
            if (is_exceptional(gwk)) gwk = (maxcn == g_refcn[z - 1][r]) ? 1.0 : 0.0;
            gw[iat * mref + r] = gwk;
            double dgwk = expd * norm - expw * dnorm * norm * norm;
            if (is_exceptional(dgwk)) dgwk = 0.0;
            gwd[iat * mref + r] = dgwk;
            double d2gwk = expdd * norm - 2.0 * expd * dnorm * norm * norm - expw * d2norm * norm * norm
                         + 2.0 * expw * dnorm * dnorm * norm * norm * norm;
            if (is_exceptional(d2gwk)) d2gwk = 0.0;
            gwd2[iat * mref + r] = d2gwk;
        }
    }
}

/* model.f90:384-428 get_atomic_c6, second-derivative branch (Fortran order). */
static void c6_second(int nat, const int *num, int mref, const double *gw, const double *gwd, const double *gwd2,
                      const int *ghost, double *c6, double *dc6, double *d2c6, double *d2c6ij)
{
    size_t nn = (size_t)nat * nat;
    memset(c6, 0, sizeof(double) * nn); memset(dc6, 0, sizeof(double) * nn);
    memset(d2c6, 0, sizeof(double) * nn); memset(d2c6ij, 0, sizeof(double) * nn);
    for (int iat = 0; iat < nat; ++iat) {
        if (ghost && ghost[iat]) continue;
        int zi = num[iat];
        for (int jat = 0; jat <= iat; ++jat) {
            if (ghost && ghost[jat]) continue;
            int zj = num[jat];
            double a = 0, di = 0, dj = 0, ddi = 0, ddj = 0, mix = 0;
            for (int ir = 0; ir < g_nref[zi - 1]; ++ir)
                for (int jr = 0; jr < g_nref[zj - 1]; ++jr) {
                    double ref = d3o_get_c6(ir + 1, jr + 1, zi, zj);
                    double wi = gw[iat * mref + ir], wj = gw[jat * mref + jr];
                    double dwi = gwd[iat * mref + ir], dwj = gwd[jat * mref + jr];
                    a += wi * wj * ref;
                    di += dwi * wj * ref;
                    dj += wi * dwj * ref;
                    ddi += gwd2[iat * mref + ir] * wj * ref;
                    ddj += wi * gwd2[jat * mref + jr] * ref;
                    mix += dwi * dwj * ref;
                }
            c6[FIDX(iat, jat)] = a; c6[FIDX(jat, iat)] = a;
            dc6[FIDX(iat, jat)] = di; dc6[FIDX(jat, iat)] = dj;
            d2c6[FIDX(iat, jat)] = ddi; d2c6[FIDX(jat, iat)] = ddj;
            d2c6ij[FIDX(iat, jat)] = mix; d2c6ij[FIDX(jat, iat)] = mix;
        }
    }
}

/*
 * Derivatives of the coordination numbers (the mctc-lib exponential count of
 * the same pairs as ncoord.f90:86-144): dr[(3 i + c) nat + k] = dCN_k / dR_ic.
 * A pair term (i, j != i, T) with v = R_i - R_j - T adds dcf/dr v/r to the
 * rows of i and subtracts it from the rows of j, in the columns of i and j;
 * self images (j = i) do not depend on the positions.
 */
static void cn_dcndr(int nat, const int *num, const double *xyz, int ntr, const double *trans, double cutoff,
                     double *dr)
{
    double cutoff2 = cutoff * cutoff;
    memset(dr, 0, sizeof(double) * 3 * (size_t)nat * nat);
    for (int iat = 0; iat < nat; ++iat)
        for (int jat = 0; jat < iat; ++jat) {
            double rc = d3o_rcov(num[iat]) + d3o_rcov(num[jat]);
            for (int itr = 0; itr < ntr; ++itr) {
                const double *t = trans + 3 * itr;
                double v[3];
                for (int c = 0; c < 3; ++c) v[c] = xyz[3 * iat + c] - (xyz[3 * jat + c] + t[c]);
                double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                if (r2 > cutoff2 || r2 < 1.0e-12) continue;
                double r1 = sqrt(r2);
                double cf = 1.0 / (1.0 + exp(-KCN * (rc / r1 - 1.0)));
                double dcf = -cf * (1.0 - cf) * KCN * rc / r2;
                for (int c = 0; c < 3; ++c) {
                    double g = dcf * v[c] / r1;
                    dr[(size_t)(3 * iat + c) * nat + iat] += g;
                    dr[(size_t)(3 * iat + c) * nat + jat] += g;
                    dr[(size_t)(3 * jat + c) * nat + iat] -= g;
                    dr[(size_t)(3 * jat + c) * nat + jat] -= g;
                }
            }
        }
}

/* Hessian element (row, col) of the 3 nat x 3 nat matrix, Fortran order. */
#define HIDX(r, c) ((size_t)(r) + (size_t)(c) * (size_t)ndim)

/*
 * damping.f90:393-539 get_dispersion2_hessian: the two-body second
 * derivatives at fixed CN, and dE/dCN, d2E/(dR dCN), d2E/dCN2 through C6.
 * dEdcndr is [3 nat][nat] and dEdcndcn [nat][nat], both row-major.
 */
static void disp2_hessian(const d3o_param *p, int nat, const int *num, const double *xyz, int ntr,
                          const double *trans, double cutoff, double width, const double *c6, const double *dc6,
                          const double *d2c6, const double *d2c6ij, double *hess, double *dEdcn,
                          double *dEdcndr, double *dEdcndcn)
{
    int ndim = 3 * nat;
    double cutoff2 = cutoff * cutoff;
    int *id = (int *)malloc(sizeof(int) * (nat > 0 ? nat : 1));
    species_ids(nat, num, id);
    for (int iat = 0; iat < nat; ++iat) {
        for (int jat = 0; jat <= iat; ++jat) {
            if (!owns_pair(iat, jat)) continue; /* damping.f90:452 */
            double c6ij = c6[FIDX(jat, iat)];
            double fac, cnd[2], cnd2[2][2];
            int nc, cnat[2];
            if (iat != jat) {
                fac = 1.0; nc = 2; cnat[0] = iat; cnat[1] = jat;
                cnd[0] = dc6[FIDX(iat, jat)]; cnd[1] = dc6[FIDX(jat, iat)];
                cnd2[0][0] = d2c6[FIDX(iat, jat)]; cnd2[1][1] = d2c6[FIDX(jat, iat)];
                cnd2[0][1] = d2c6ij[FIDX(iat, jat)]; cnd2[1][0] = d2c6ij[FIDX(iat, jat)];
            } else {
                fac = 0.5; nc = 1; cnat[0] = iat;
                cnd[0] = 2.0 * dc6[FIDX(iat, iat)];
                cnd2[0][0] = 2.0 * d2c6[FIDX(iat, iat)] + 2.0 * d2c6ij[FIDX(iat, iat)];
            }
            for (int jtr = 0; jtr < ntr; ++jtr) {
                const double *t = trans + 3 * jtr;
                double v[3];
                for (int c = 0; c < 3; ++c) v[c] = xyz[3 * iat + c] - (xyz[3 * jat + c] + t[c]);
                double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                if (r2 > cutoff2 || r2 < EPS) continue;
                double e0, e0r, e0c, e0rr, e0rc, e0cc, sw, swr, swrr;
                damping_kernel(p, num[iat], num[jat], id[iat], id[jat], r2, c6ij, &e0, &e0r, &e0c, &e0rr, &e0rc,
                               &e0cc);
                smooth_cutoff_r2(r2, cutoff, width, &sw, &swr, &swrr);
                double fr = fac * (swr * e0 + sw * e0r);
                double frr = fac * (swrr * e0 + 2.0 * swr * e0r + sw * e0rr);
                double fc = fac * sw * e0c;
                double frc = fac * (swr * e0c + sw * e0rc);
                double fcc = fac * sw * e0cc;
                double dr2[3][2] = {{0}};
                if (iat != jat) {
                    double block[3][3];
                    for (int a = 0; a < 3; ++a) {
                        for (int b = 0; b < 3; ++b) block[a][b] = 4.0 * frr * v[a] * v[b];
                        block[a][a] += 2.0 * fr;
                    }
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            hess[HIDX(3 * iat + a, 3 * iat + b)] += block[a][b];
                            hess[HIDX(3 * jat + a, 3 * jat + b)] += block[a][b];
                            hess[HIDX(3 * iat + a, 3 * jat + b)] -= block[a][b];
                            hess[HIDX(3 * jat + a, 3 * iat + b)] -= block[a][b];
                        }
                    for (int c = 0; c < 3; ++c) { dr2[c][0] = 2.0 * v[c]; dr2[c][1] = -2.0 * v[c]; }
                }
                for (int ia = 0; ia < nc; ++ia) {
                    dEdcn[cnat[ia]] += fc * cnd[ia];
                    for (int ib = 0; ib < nc; ++ib)
                        dEdcndcn[(size_t)cnat[ia] * nat + cnat[ib]] += fcc * cnd[ia] * cnd[ib] + fc * cnd2[ia][ib];
                }
                if (iat != jat)
                    for (int ia = 0; ia < nc; ++ia)
                        for (int c = 0; c < 3; ++c) {
                            dEdcndr[(size_t)(3 * iat + c) * nat + cnat[ia]] += frc * dr2[c][0] * cnd[ia];
                            dEdcndr[(size_t)(3 * jat + c) * nat + cnat[ia]] += frc * dr2[c][1] * cnd[ia];
                        }
            }
        }
    }
    free(id);
}

/*
 * atm.f90:38-429 get_atm_dispersion_hessian: every triple term as a function
 * K(u1, u2, u3) of the squared distances u1 = |R_j - R_i|^2, u2 = |R_k - R_i|^2,
 * u3 = |R_k - R_j|^2 (switching function, angular factor, zero damping), its
 * gradient and Hessian in u, contracted with du/dR; plus the C6 derivatives.
 */
static void atm_hessian(int nat, const int *num, const double *xyz, int ntr, const double *trans, double cutoff,
                        double width, double s9, double rs9, double alp, const double *c6, const double *dc6,
                        const double *d2c6, const double *d2c6ij, double *hess, double *dEdcn, double *dEdcndr,
                        double *dEdcndcn)
{
    if (fabs(s9) < EPS) return;
    static const int sigp[3][3] = {{-1, -1, 0}, {1, 0, -1}, {0, 1, 1}};  /* sigp(atom, pair) */
    static const double cl[3][3] = {{1, -1, 1}, {1, 1, -1}, {-1, 1, 1}};  /* cl(l, pair) */
    int ndim = 3 * nat;
    double cutoff2 = cutoff * cutoff, alp3 = alp / 3.0, aexp = 0.5 * alp3;
    for (int iat = 0; iat < nat; ++iat) {
        int zi = num[iat];
        for (int jat = 0; jat <= iat; ++jat) {
            if (!owns_pair(iat, jat)) continue; /* atm.f90:136 */
            int zj = num[jat];
            for (int jtr = 0; jtr < ntr; ++jtr) {
                double vec[3][3], u[3];  /* vec[pair][xyz] */
                for (int c = 0; c < 3; ++c) vec[0][c] = xyz[3 * jat + c] + trans[3 * jtr + c] - xyz[3 * iat + c];
                u[0] = vec[0][0] * vec[0][0] + vec[0][1] * vec[0][1] + vec[0][2] * vec[0][2];
                if (u[0] > cutoff2 || u[0] < EPS) continue;
                for (int kat = 0; kat <= jat; ++kat) {
                    int zk = num[kat];
                    double triple = triple_scale(iat, jat, kat);
                    double r0 = rs9 * d3o_rvdw(zj, zi) * rs9 * d3o_rvdw(zk, zi) * rs9 * d3o_rvdw(zk, zj);
                    for (int ktr = 0; ktr < ntr; ++ktr) {
                        for (int c = 0; c < 3; ++c)
                            vec[1][c] = xyz[3 * kat + c] + trans[3 * ktr + c] - xyz[3 * iat + c];
                        u[1] = vec[1][0] * vec[1][0] + vec[1][1] * vec[1][1] + vec[1][2] * vec[1][2];
                        if (u[1] > cutoff2 || u[1] < EPS) continue;
                        for (int c = 0; c < 3; ++c) vec[2][c] = vec[1][c] - vec[0][c];
                        u[2] = vec[2][0] * vec[2][0] + vec[2][1] * vec[2][1] + vec[2][2] * vec[2][2];
                        if (u[2] > cutoff2 || u[2] < EPS) continue;
                        double c6t[3] = {c6[FIDX(jat, iat)], c6[FIDX(kat, iat)], c6[FIDX(kat, jat)]};
                        if (fabs(c6t[0]) < EPS || fabs(c6t[1]) < EPS || fabs(c6t[2]) < EPS) continue;
                        double swp[3], dswp[3], d2swp[3];
                        for (int ip = 0; ip < 3; ++ip) smooth_cutoff_r2(u[ip], cutoff, width, &swp[ip], &dswp[ip], &d2swp[ip]);
                        double sval = swp[0] * swp[1] * swp[2], sp[3], spq[3][3];
                        sp[0] = dswp[0] * swp[1] * swp[2];
                        sp[1] = swp[0] * dswp[1] * swp[2];
                        sp[2] = swp[0] * swp[1] * dswp[2];
                        spq[0][0] = d2swp[0] * swp[1] * swp[2];
                        spq[1][1] = swp[0] * d2swp[1] * swp[2];
                        spq[2][2] = swp[0] * swp[1] * d2swp[2];
                        spq[0][1] = spq[1][0] = dswp[0] * dswp[1] * swp[2];
                        spq[0][2] = spq[2][0] = dswp[0] * swp[1] * dswp[2];
                        spq[1][2] = spq[2][1] = swp[0] * dswp[1] * dswp[2];
                        double ww = u[0] * u[1] * u[2];
                        double wp_[3] = {u[1] * u[2], u[0] * u[2], u[0] * u[1]};
                        double wpq[3][3] = {{0, u[2], u[1]}, {u[2], 0, u[0]}, {u[1], u[0], 0}};
                        double lval[3] = {u[0] + u[2] - u[1], u[0] - u[2] + u[1], -u[0] + u[2] + u[1]};
                        double nval = lval[0] * lval[1] * lval[2], np_[3], npq[3][3];
                        for (int ip = 0; ip < 3; ++ip)
                            np_[ip] = cl[0][ip] * lval[1] * lval[2] + cl[1][ip] * lval[0] * lval[2]
                                    + cl[2][ip] * lval[0] * lval[1];
                        for (int ip = 0; ip < 3; ++ip)
                            for (int iq = 0; iq < 3; ++iq) {
                                double tmp = 0.0;
                                for (int il = 0; il < 3; ++il)
                                    for (int im = 0; im < 3; ++im) {
                                        if (il == im) continue;
                                        tmp += cl[il][ip] * cl[im][iq] * lval[3 - il - im];
                                    }
                                npq[ip][iq] = tmp;
                            }
                        double gv = pow(ww, -2.5), gd = -2.5 * pow(ww, -3.5), gdd = 8.75 * pow(ww, -4.5);
                        double hv = pow(ww, -1.5), hd = -1.5 * pow(ww, -2.5), hdd = 3.75 * pow(ww, -3.5);
                        double ang = 0.375 * nval * gv + hv, angp[3], angpq[3][3];
                        for (int ip = 0; ip < 3; ++ip)
                            angp[ip] = 0.375 * (np_[ip] * gv + nval * gd * wp_[ip]) + hd * wp_[ip];
                        for (int ip = 0; ip < 3; ++ip)
                            for (int iq = 0; iq < 3; ++iq)
                                angpq[ip][iq] = 0.375 * (npq[ip][iq] * gv + np_[ip] * gd * wp_[iq] + np_[iq] * gd * wp_[ip]
                                                         + nval * (gdd * wp_[ip] * wp_[iq] + gd * wpq[ip][iq]))
                                              + hdd * wp_[ip] * wp_[iq] + hd * wpq[ip][iq];
                        double cval = 6.0 * pow(r0, alp3), tval = cval * pow(ww, -aexp), tp[3], tpq[3][3];
                        for (int ip = 0; ip < 3; ++ip) tp[ip] = -aexp * tval * wp_[ip] / ww;
                        for (int ip = 0; ip < 3; ++ip)
                            for (int iq = 0; iq < 3; ++iq)
                                tpq[ip][iq] = aexp * (aexp + 1.0) * tval * wp_[ip] * wp_[iq] / (ww * ww)
                                            - aexp * tval * wpq[ip][iq] / ww;
                        double fd = 1.0 / (1.0 + tval), fdp[3], fdpq[3][3];
                        for (int ip = 0; ip < 3; ++ip) fdp[ip] = -tp[ip] * fd * fd;
                        for (int ip = 0; ip < 3; ++ip)
                            for (int iq = 0; iq < 3; ++iq)
                                fdpq[ip][iq] = -tpq[ip][iq] * fd * fd + 2.0 * tp[ip] * tp[iq] * fd * fd * fd;
                        double av = ang * fd, ap[3], apq[3][3];
                        for (int ip = 0; ip < 3; ++ip) ap[ip] = angp[ip] * fd + ang * fdp[ip];
                        for (int ip = 0; ip < 3; ++ip)
                            for (int iq = 0; iq < 3; ++iq)
                                apq[ip][iq] = angpq[ip][iq] * fd + angp[ip] * fdp[iq] + angp[iq] * fdp[ip]
                                            + ang * fdpq[ip][iq];
                        double kv = sval * av, kp[3], kpq[3][3];
                        for (int ip = 0; ip < 3; ++ip) kp[ip] = sp[ip] * av + sval * ap[ip];
                        for (int ip = 0; ip < 3; ++ip)
                            for (int iq = 0; iq < 3; ++iq)
                                kpq[ip][iq] = spq[ip][iq] * av + sp[ip] * ap[iq] + sp[iq] * ap[ip] + sval * apq[ip][iq];
                        double qv = sqrt(fabs(c6t[0] * c6t[1] * c6t[2])), qm[3], qmn[3][3];
                        for (int ip = 0; ip < 3; ++ip) qm[ip] = 0.5 * qv / c6t[ip];
                        for (int ip = 0; ip < 3; ++ip)
                            for (int iq = 0; iq < 3; ++iq)
                                qmn[ip][iq] = (ip == iq) ? -0.25 * qv / (c6t[ip] * c6t[ip])
                                                         : 0.25 * qv / (c6t[ip] * c6t[iq]);
                        double pref = s9 * triple;
                        int at[3] = {iat, jat, kat};
                        double grad[3][3][3];  /* grad[atom][pair][xyz] = du_pair / dR_atom */
                        for (int ia = 0; ia < 3; ++ia)
                            for (int ip = 0; ip < 3; ++ip)
                                for (int c = 0; c < 3; ++c) grad[ia][ip][c] = 2.0 * sigp[ia][ip] * vec[ip][c];
                        double pq = pref * qv, gk[3][3][3], dg[3][3];
                        for (int ia = 0; ia < 3; ++ia)
                            for (int c = 0; c < 3; ++c)
                                for (int iq = 0; iq < 3; ++iq) {
                                    double tmp = 0.0;
                                    for (int ip = 0; ip < 3; ++ip) tmp += kpq[ip][iq] * grad[ia][ip][c];
                                    gk[ia][iq][c] = pq * tmp;
                                }
                        for (int ia = 0; ia < 3; ++ia)
                            for (int ib = 0; ib < 3; ++ib) {
                                double tmp = 0.0;
                                for (int ip = 0; ip < 3; ++ip) tmp += kp[ip] * sigp[ia][ip] * sigp[ib][ip];
                                dg[ia][ib] = 2.0 * pq * tmp;
                            }
                        for (int ia = 0; ia < 3; ++ia)
                            for (int ib = 0; ib < 3; ++ib)
                                for (int ic = 0; ic < 3; ++ic)
                                    for (int jc = 0; jc < 3; ++jc) {
                                        double tmp = 0.0;
                                        for (int iq = 0; iq < 3; ++iq) tmp += gk[ia][iq][ic] * grad[ib][iq][jc];
                                        if (ic == jc) tmp += dg[ia][ib];
                                        hess[HIDX(3 * at[ia] + ic, 3 * at[ib] + jc)] += tmp;
                                    }
                        double ec[3], ecc[3][3], euc[3][3], egr[3][3][3];
                        for (int ip = 0; ip < 3; ++ip) {
                            ec[ip] = pref * qm[ip] * kv;
                            for (int iq = 0; iq < 3; ++iq) {
                                ecc[ip][iq] = pref * qmn[ip][iq] * kv;
                                euc[iq][ip] = pref * qm[ip] * kp[iq];
                            }
                        }
                        for (int ia = 0; ia < 3; ++ia)
                            for (int c = 0; c < 3; ++c)
                                for (int im = 0; im < 3; ++im) {
                                    double tmp = 0.0;
                                    for (int ip = 0; ip < 3; ++ip) tmp += euc[ip][im] * grad[ia][ip][c];
                                    egr[ia][im][c] = tmp;
                                }
                        int pat[3][2] = {{iat, jat}, {iat, kat}, {jat, kat}};
                        int nent = 0, ent_atom[6], ent_pair[6];
                        double ent_coef[6];
                        for (int ip = 0; ip < 3; ++ip) {
                            int a0 = pat[ip][0], a1 = pat[ip][1];
                            if (a0 != a1) {
                                ent_atom[nent] = a0; ent_pair[nent] = ip; ent_coef[nent++] = dc6[FIDX(a0, a1)];
                                ent_atom[nent] = a1; ent_pair[nent] = ip; ent_coef[nent++] = dc6[FIDX(a1, a0)];
                            } else {
                                ent_atom[nent] = a0; ent_pair[nent] = ip; ent_coef[nent++] = 2.0 * dc6[FIDX(a0, a0)];
                            }
                        }
                        for (int ie = 0; ie < nent; ++ie) {
                            dEdcn[ent_atom[ie]] += ec[ent_pair[ie]] * ent_coef[ie];
                            for (int jf = 0; jf < nent; ++jf)
                                dEdcndcn[(size_t)ent_atom[ie] * nat + ent_atom[jf]] +=
                                    ecc[ent_pair[ie]][ent_pair[jf]] * ent_coef[ie] * ent_coef[jf];
                            for (int ia = 0; ia < 3; ++ia)
                                for (int c = 0; c < 3; ++c)
                                    dEdcndr[(size_t)(3 * at[ia] + c) * nat + ent_atom[ie]] +=
                                        egr[ia][ent_pair[ie]][c] * ent_coef[ie];
                        }
                        for (int ip = 0; ip < 3; ++ip) {
                            int a0 = pat[ip][0], a1 = pat[ip][1];
                            if (a0 != a1) {
                                dEdcndcn[(size_t)a0 * nat + a0] += ec[ip] * d2c6[FIDX(a0, a1)];
                                dEdcndcn[(size_t)a1 * nat + a1] += ec[ip] * d2c6[FIDX(a1, a0)];
                                dEdcndcn[(size_t)a0 * nat + a1] += ec[ip] * d2c6ij[FIDX(a0, a1)];
                                dEdcndcn[(size_t)a1 * nat + a0] += ec[ip] * d2c6ij[FIDX(a0, a1)];
                            } else {
                                dEdcndcn[(size_t)a0 * nat + a0] +=
                                    ec[ip] * (2.0 * d2c6[FIDX(a0, a0)] + 2.0 * d2c6ij[FIDX(a0, a0)]);
                            }
                        }
                    }
                }
            }
        }
    }
}

/* ncoord.f90:236-306 add_coordination_number_hessian: dE/dCN times the
 * second derivatives of the coordination numbers. */
static void cn_hessian(int nat, const int *num, const double *xyz, int ntr, const double *trans, double cutoff,
                       const double *dEdcn, double *hess)
{
    int ndim = 3 * nat;
    double cutoff2 = cutoff * cutoff;
    for (int iat = 0; iat < nat; ++iat)
        for (int jat = 0; jat < iat; ++jat) {
            double rc = d3o_rcov(num[iat]) + d3o_rcov(num[jat]);
            double dEdcnij = dEdcn[iat] + dEdcn[jat];
            for (int itr = 0; itr < ntr; ++itr) {
                double v[3];
                for (int c = 0; c < 3; ++c) v[c] = xyz[3 * iat + c] - (xyz[3 * jat + c] + trans[3 * itr + c]);
                double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                if (r2 > cutoff2 || r2 < 1.0e-12) continue;
                double r1 = sqrt(r2);
                double expterm = exp(-KCN * (rc / r1 - 1.0));
                double cf = 1.0 / (1.0 + expterm), tmp = cf * (1.0 - cf);
                double dcf = -tmp * KCN * rc / r2;
                double d2cf = tmp * (1.0 - 2.0 * cf) * (KCN * rc) * (KCN * rc) / (r2 * r2)
                            + 2.0 * tmp * KCN * rc / (r2 * r1);
                double block[3][3];
                for (int a = 0; a < 3; ++a) {
                    for (int b = 0; b < 3; ++b)
                        block[a][b] = dEdcnij * (d2cf * v[a] * v[b] / r2 - dcf * v[a] * v[b] / (r2 * r1));
                    block[a][a] += dEdcnij * dcf / r1;
                }
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        hess[HIDX(3 * iat + a, 3 * iat + b)] += block[a][b];
                        hess[HIDX(3 * jat + a, 3 * jat + b)] += block[a][b];
                        hess[HIDX(3 * iat + a, 3 * jat + b)] -= block[a][b];
                        hess[HIDX(3 * jat + a, 3 * iat + b)] -= block[a][b];
                    }
            }
        }
}

/*
 * disp.f90:236-308 get_dispersion_hessian (the energy as in
 * d3o_get_dispersion): hessian is 3 nat x 3 nat in Fortran order, the array
 * the C API of the reference returns.  Same arguments as d3o_get_dispersion.
 */
int d3o_get_dispersion_hessian(int nat, const int *num, const double *xyz, const double *lattice,
                               const int *periodic, int kind, const double *par, const double *cut,
                               const int *ghost, double *energy, double *hessian)
{
    d3o_param p = {kind, par[0], par[1], par[2], par[3], par[4], par[5], par[6], par[7], par[8]};
    int ndim = 3 * nat, mref = 0;
    for (int i = 0; i < nat; ++i)
        if (g_nref[num[i] - 1] > mref) mref = g_nref[num[i] - 1];
    double zero_lat[9] = {0};
    int zero_pbc[3] = {0, 0, 0};
    if (!lattice) lattice = zero_lat;
    if (!periodic) periodic = zero_pbc;
    if (energy)
        d3o_get_dispersion(nat, num, xyz, lattice, periodic, kind, par, cut, ghost, energy, NULL, NULL, NULL,
                           NULL, NULL);
    size_t nn = (size_t)nat * nat;
    double *cn = (double *)calloc(nat, sizeof(double));
    double *dr = (double *)calloc(3 * nn, sizeof(double));
    double *trans;
    int ntr = lattice_alloc(periodic, lattice, cut[0], &trans);
    d3o_cn(nat, num, xyz, ntr, trans, cut[0], cn);
    cn_dcndr(nat, num, xyz, ntr, trans, cut[0], dr);
    free(trans);
    double *gw = (double *)calloc((size_t)nat * mref, sizeof(double));
    double *gwd = (double *)calloc((size_t)nat * mref, sizeof(double));
    double *gwd2 = (double *)calloc((size_t)nat * mref, sizeof(double));
    weights2(nat, num, cn, mref, gw, gwd, gwd2);
    double *c6 = (double *)calloc(nn, sizeof(double)), *dc6 = (double *)calloc(nn, sizeof(double));
    double *d2c6 = (double *)calloc(nn, sizeof(double)), *d2c6ij = (double *)calloc(nn, sizeof(double));
    c6_second(nat, num, mref, gw, gwd, gwd2, ghost, c6, dc6, d2c6, d2c6ij);
    memset(hessian, 0, sizeof(double) * (size_t)ndim * ndim);
    double *dEdcn = (double *)calloc(nat, sizeof(double));
    double *dEdcndr = (double *)calloc(3 * nn, sizeof(double));
    double *dEdcndcn = (double *)calloc(nn, sizeof(double));
    if (two_body_active(&p)) {
        ntr = lattice_alloc(periodic, lattice, cut[1], &trans);
        disp2_hessian(&p, nat, num, xyz, ntr, trans, cut[1], cut[3], c6, dc6, d2c6, d2c6ij, hessian, dEdcn,
                      dEdcndr, dEdcndcn);
        free(trans);
    }
    ntr = lattice_alloc(periodic, lattice, cut[2], &trans);
    atm_hessian(nat, num, xyz, ntr, trans, cut[2], cut[4], p.s9, 4.0 / 3.0, p.alp + 2.0, c6, dc6, d2c6, d2c6ij,
                hessian, dEdcn, dEdcndr, dEdcndcn);
    free(trans);
    ntr = lattice_alloc(periodic, lattice, cut[0], &trans);
    cn_hessian(nat, num, xyz, ntr, trans, cut[0], dEdcn, hessian);
    free(trans);
    /* hessian += dEdcndr dr^T + dr dEdcndr^T + dr dEdcndcn dr^T (disp.f90:303-306) */
    double *tmp = (double *)calloc(3 * nn, sizeof(double));  /* dEdcndcn dr^T as [nat][3 nat] */
    for (int k = 0; k < nat; ++k)
        for (int y = 0; y < ndim; ++y) {
            double s = 0.0;
            for (int l = 0; l < nat; ++l) s += dEdcndcn[(size_t)k * nat + l] * dr[(size_t)y * nat + l];
            tmp[(size_t)k * ndim + y] = s;
        }
    for (int y = 0; y < ndim; ++y)
        for (int x = 0; x < ndim; ++x) {
            double s = 0.0;
            for (int k = 0; k < nat; ++k)
                s += dEdcndr[(size_t)x * nat + k] * dr[(size_t)y * nat + k]
                   + dr[(size_t)x * nat + k] * dEdcndr[(size_t)y * nat + k]
                   + dr[(size_t)x * nat + k] * tmp[(size_t)k * ndim + y];
            hessian[HIDX(x, y)] += s;
        }
    free(tmp);
    free(cn); free(dr); free(gw); free(gwd); free(gwd2); free(c6); free(dc6); free(d2c6); free(d2c6ij);
    free(dEdcn); free(dEdcndr); free(dEdcndcn);
    return 0;
}

/*
 * The three-body part of get_dispersion_hessian before the CN chain rule
 * (disp.f90:288-298): the C6 matrices of c6_second and the four accumulators
 * of atm_hessian (atm.f90:38-429), all column-major as FIDX / HIDX, for
 * checks of per-triple Hessian code.  cut as d3o_get_dispersion_hessian.
 */
int d3o_atm_hessian_parts(int nat, const int *num, const double *xyz, const double *lattice,
                          const int *periodic, const double *cut, double s9, double alp, double *c6, double *dc6,
                          double *d2c6, double *d2c6ij, double *hessian, double *dEdcn, double *dEdcndr,
                          double *dEdcndcn)
{
    int ndim = 3 * nat, mref = 0;
    for (int i = 0; i < nat; ++i)
        if (g_nref[num[i] - 1] > mref) mref = g_nref[num[i] - 1];
    double zero_lat[9] = {0};
    int zero_pbc[3] = {0, 0, 0};
    if (!lattice) lattice = zero_lat;
    if (!periodic) periodic = zero_pbc;
    double *cn = (double *)calloc(nat, sizeof(double));
    double *trans;
    int ntr = lattice_alloc(periodic, lattice, cut[0], &trans);
    d3o_cn(nat, num, xyz, ntr, trans, cut[0], cn);
    free(trans);
    double *gw = (double *)calloc((size_t)nat * mref, sizeof(double));
    double *gwd = (double *)calloc((size_t)nat * mref, sizeof(double));
    double *gwd2 = (double *)calloc((size_t)nat * mref, sizeof(double));
    weights2(nat, num, cn, mref, gw, gwd, gwd2);
    c6_second(nat, num, mref, gw, gwd, gwd2, NULL, c6, dc6, d2c6, d2c6ij);
    memset(hessian, 0, sizeof(double) * (size_t)ndim * ndim);
    memset(dEdcn, 0, sizeof(double) * nat);
    memset(dEdcndr, 0, sizeof(double) * 3 * (size_t)nat * nat);
    memset(dEdcndcn, 0, sizeof(double) * (size_t)nat * nat);
    ntr = lattice_alloc(periodic, lattice, cut[2], &trans);
    atm_hessian(nat, num, xyz, ntr, trans, cut[2], cut[4], s9, 4.0 / 3.0, alp + 2.0, c6, dc6, d2c6, d2c6ij,
                hessian, dEdcn, dEdcndr, dEdcndcn);
    free(trans);
    free(cn); free(gw); free(gwd); free(gwd2);
    return 0;
}

/* ------------------------------------------------------ reciprocal space */
/* fourier/jacobi.f90:40-122 symmetric_eigendecomposition: cyclic Jacobi
 * rotations (at most 100 sweeps), eigenvalues sorted by decreasing magnitude.
 * amat, evec: ndim x ndim column-major. */
static void jacobi_eig(int ndim, const double *amat, double *eval, double *evec)
{
#define A_(r, c) work[(size_t)(r) + (size_t)(c) * ndim]
#define V_(r, c) evec[(size_t)(r) + (size_t)(c) * ndim]
    memset(evec, 0, sizeof(double) * ndim * ndim);
    for (int p = 0; p < ndim; ++p) V_(p, p) = 1.0;
    memset(eval, 0, sizeof(double) * ndim);
    double anorm = 0.0;
    for (int k = 0; k < ndim * ndim; ++k) anorm += amat[k] * amat[k];
    anorm = sqrt(anorm);
    if (anorm <= 0.0) return;
    double skip = 0.01 * EPS * anorm;
    double *work = (double *)malloc(sizeof(double) * ndim * ndim);
    memcpy(work, amat, sizeof(double) * ndim * ndim);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int p = 0; p < ndim - 1; ++p)
            for (int q = p + 1; q < ndim; ++q) off += A_(p, q) * A_(p, q);
        if (sqrt(2.0 * off) <= EPS * anorm) break;
        for (int p = 0; p < ndim - 1; ++p)
            for (int q = p + 1; q < ndim; ++q) {
                if (fabs(A_(p, q)) <= skip) continue;
                double theta = 0.5 * (A_(q, q) - A_(p, p)) / A_(p, q);
                double t = copysign(1.0, theta) / (fabs(theta) + sqrt(1.0 + theta * theta));
                double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
                for (int j = 0; j < ndim; ++j) {
                    double ajp = A_(j, p), ajq = A_(j, q);
                    A_(j, p) = c * ajp - sn * ajq;
                    A_(j, q) = sn * ajp + c * ajq;
                }
                for (int j = 0; j < ndim; ++j) {
                    double apj = A_(p, j), aqj = A_(q, j);
                    A_(p, j) = c * apj - sn * aqj;
                    A_(q, j) = sn * apj + c * aqj;
                }
                A_(p, q) = 0.0;
                A_(q, p) = 0.0;
                for (int j = 0; j < ndim; ++j) {
                    double vjp = V_(j, p), vjq = V_(j, q);
                    V_(j, p) = c * vjp - sn * vjq;
                    V_(j, q) = sn * vjp + c * vjq;
                }
            }
    }
    for (int p = 0; p < ndim; ++p) eval[p] = A_(p, p);
    for (int p = 0; p < ndim - 1; ++p) {
        int imax = p;
        for (int k = p + 1; k < ndim; ++k)
            if (fabs(eval[k]) > fabs(eval[imax])) imax = k;
        if (imax == p) continue;
        double t = eval[p]; eval[p] = eval[imax]; eval[imax] = t;
        for (int j = 0; j < ndim; ++j) { t = V_(j, p); V_(j, p) = V_(j, imax); V_(j, imax) = t; }
    }
    free(work);
#undef A_
#undef V_
}

/* fourier/decomposition.f90:104-175 new_lowrank_c6 over the species of the
 * structure (first appearance order): the reference C6 of all (species,
 * reference) pairs as a symmetric matrix, its eigenpairs by decreasing
 * magnitude, kept until the largest relative error (floor 1e-6 of the largest
 * entry) is at most the tolerance, or the requested rank.  vec is
 * [nid][MAX_REF][rank]. */
typedef struct { int rank, nid, ndim; double error, *lambda, *vec; } d3o_lowrank;

static void lowrank_new(int nat, const int *num, int want_rank, double tol, d3o_lowrank *lr, int *id)
{
    species_ids(nat, num, id);
    int nid = 0;
    for (int i = 0; i < nat; ++i) if (id[i] > nid) nid = id[i];
    int zsp[MAX_ELEM] = {0};
    for (int i = 0; i < nat; ++i) zsp[id[i] - 1] = num[i];
    int ndim = 0, spmap[MAX_ELEM * MAX_REF], refmap[MAX_ELEM * MAX_REF];
    for (int sp = 0; sp < nid; ++sp)
        for (int r = 0; r < g_nref[zsp[sp] - 1]; ++r) { spmap[ndim] = sp; refmap[ndim] = r; ++ndim; }
    double *amat = (double *)calloc((size_t)ndim * ndim, sizeof(double));
    for (int ii = 0; ii < ndim; ++ii)
        for (int jj = 0; jj < ndim; ++jj)
            amat[jj + ii * ndim] = d3o_get_c6(refmap[jj] + 1, refmap[ii] + 1, zsp[spmap[jj]], zsp[spmap[ii]]);
    double *eval = (double *)malloc(sizeof(double) * ndim), *evec = (double *)malloc(sizeof(double) * ndim * ndim);
    jacobi_eig(ndim, amat, eval, evec);
    double amax = 0.0;
    for (int k = 0; k < ndim * ndim; ++k) if (fabs(amat[k]) > amax) amax = fabs(amat[k]);
    double floor_ = 1.0e-6 * amax;
    double *res = (double *)malloc(sizeof(double) * ndim * ndim);
    memcpy(res, amat, sizeof(double) * ndim * ndim);
    int nrank = ndim;
    double err = 0.0;
    for (int il = 0; il < ndim; ++il) {
        for (int ii = 0; ii < ndim; ++ii)
            for (int jj = 0; jj < ndim; ++jj)
                res[jj + ii * ndim] -= eval[il] * evec[jj + il * ndim] * evec[ii + il * ndim];
        err = 0.0;
        for (int k = 0; k < ndim * ndim; ++k) {
            double e = fabs(res[k]) / fmax(fabs(amat[k]), floor_);
            if (e > err) err = e;
        }
        if (want_rank > 0) {
            if (il + 1 >= (want_rank < ndim ? want_rank : ndim)) { nrank = il + 1; break; }
        } else if (err <= tol) {
            nrank = il + 1;
            break;
        }
    }
    lr->rank = nrank;
    lr->nid = nid;
    lr->ndim = ndim;
    lr->error = err;
    lr->lambda = (double *)malloc(sizeof(double) * nrank);
    lr->vec = (double *)calloc((size_t)nid * MAX_REF * nrank, sizeof(double));
    for (int il = 0; il < nrank; ++il) {
        lr->lambda[il] = eval[il];
        for (int ii = 0; ii < ndim; ++ii)
            lr->vec[((size_t)spmap[ii] * MAX_REF + refmap[ii]) * nrank + il] = evec[ii + il * ndim];
    }
    free(amat); free(eval); free(evec); free(res);
}

/* decomposition.f90:202-242 get_weights: c6l[i][l] = vec(:, id_i, l) . w_i (and
 * dc6l with the weight derivatives); ghost atoms have zero rows. */
static void lowrank_weights(const d3o_lowrank *lr, int nat, const int *num, const int *id, const int *ghost,
                            int mref, const double *gw, const double *gwd, double *c6l, double *dc6l)
{
    int R = lr->rank;
    for (int i = 0; i < nat; ++i)
        for (int l = 0; l < R; ++l) {
            double a = 0.0, d = 0.0;
            if (!(ghost && ghost[i]))
                for (int r = 0; r < g_nref[num[i] - 1]; ++r) {
                    double v = lr->vec[((size_t)(id[i] - 1) * MAX_REF + r) * R + l];
                    a += v * gw[i * mref + r];
                    if (gwd) d += v * gwd[i * mref + r];
                }
            c6l[(size_t)i * R + l] = a;
            if (dc6l) dc6l[(size_t)i * R + l] = d;
        }
}

/* decomposition.f90:247-296 get_atomic_c6 of the low-rank model (Fortran order). */
static void lowrank_c6(const d3o_lowrank *lr, int nat, const double *c6l, const double *dc6l, double *c6,
                       double *dc6)
{
    int R = lr->rank;
    for (int i = 0; i < nat; ++i)
        for (int j = 0; j <= i; ++j) {
            double a = 0.0, di = 0.0, dj = 0.0;
            for (int l = 0; l < R; ++l) {
                a += lr->lambda[l] * c6l[(size_t)i * R + l] * c6l[(size_t)j * R + l];
                if (dc6l) {
                    di += lr->lambda[l] * dc6l[(size_t)i * R + l] * c6l[(size_t)j * R + l];
                    dj += lr->lambda[l] * c6l[(size_t)i * R + l] * dc6l[(size_t)j * R + l];
                }
            }
            c6[FIDX(i, j)] = a; c6[FIDX(j, i)] = a;
            if (dc6) { dc6[FIDX(i, j)] = di; dc6[FIDX(j, i)] = dj; }
        }
}

/* fourier/kernel.f90: a term pref * r^m / (r^alpha + radius^alpha). */
typedef struct { double pref; int m, alpha; double radius; } d3o_term;

/* kernel.f90:83-129 get_fourier_transform */
static void fourier_transform(const d3o_term *t, double k, double *phi, double *dphi)
{
    *phi = 0.0; *dphi = 0.0;
    if (fabs(t->pref) <= 0.0) return;
    int nexp = t->m - t->alpha + 2;
    double pre = 4.0 * M_PI * M_PI / t->alpha * t->pref * pow(t->radius, nexp);
    if (k <= 0.0) {
        *phi = pre * t->radius / sin(M_PI * (t->m + 3) / (double)t->alpha);
        return;
    }
    int npole = t->alpha / 2;
    double krad = k * t->radius, acc = 0.0, dacc = 0.0;
    for (int ip = 0; ip < npole; ++ip) {
        double theta = M_PI * (2 * ip + 1) / (double)t->alpha;
        double ct = cos(theta), st = sin(theta);
        double arg = 0.5 * M_PI + nexp * theta + krad * ct;
        double damp = exp(-krad * st);
        acc += sin(arg) * damp;
        dacc += cos(arg + theta) * damp;
    }
    *phi = pre * acc / k;
    *dphi = (pre * t->radius * dacc - *phi) / k;
}

/* kernel.f90:133-147 get_potential_zero, :151-164 get_reciprocal_cutoff */
static double potential_zero(const d3o_term *t) { return t->m > 0 ? 0.0 : t->pref / pow(t->radius, t->alpha); }
static double reciprocal_cutoff(const d3o_term *t, double tol)
{
    return -log(tol) / (t->radius * sin(M_PI / (double)t->alpha));
}

/* get_fourier_terms of the rational (rational.f90) and zero (zero.f90:127-163)
 * damping for the species of atoms zi, zj; returns the number of terms. */
static int fourier_terms(const d3o_param *p, int zi, int zj, d3o_term *t)
{
    if (fabs(p->s6) < EPS && fabs(p->s8) < EPS) return 0;
    double rrij = 3 * d3o_r4r2(zi) * d3o_r4r2(zj);
    if (p->kind == K_RATIONAL) {
        double r0 = p->a1 * sqrt(rrij) + p->a2;
        t[0] = (d3o_term){p->s6, 0, 6, r0};
        t[1] = (d3o_term){p->s8 * rrij, 0, 8, r0};
    } else {
        int alp6 = (int)lround(p->alp), alp8 = alp6 + 2;
        double r0 = d3o_rvdw(zj, zi);
        t[0] = (d3o_term){p->s6, alp6 - 6, alp6, pow(6.0, 1.0 / alp6) * p->a1 * r0};
        t[1] = (d3o_term){p->s8 * rrij, alp8 - 8, alp8, pow(6.0, 1.0 / alp8) * p->a2 * r0};
    }
    return 2;
}

/*
 * fourier/ewald.f90:47-252 get_dispersion_ewald with damping.f90:306-385 (the
 * default reciprocal cutoff from the damping radii, kcut_tolerance 1e-8):
 * the two-body energy of the 3D-periodic system as a sum over reciprocal
 * lattice vectors of the structure factors of the low-rank C6 weights.
 * Returns nonzero when the damping function has no Fourier representation.
 */
static int ewald2(const d3o_param *p, int nat, const int *num, const int *id, const double *xyz,
                  const double *latflat, const d3o_lowrank *lr, const double *c6l, const double *dc6l, double kcut_req,
                  double *energy, double *dEdcn, double *gradient, double *sigma, int grad)
{
    int nid = lr->nid, R = lr->rank;
    int zsp[MAX_ELEM] = {0};
    for (int i = 0; i < nat; ++i) zsp[id[i] - 1] = num[i];
    d3o_term *terms = (d3o_term *)calloc((size_t)nid * nid * 2, sizeof(d3o_term));
    int *nterm = (int *)calloc((size_t)nid * nid, sizeof(int));
    double kcut = 0.0;
    for (int a = 0; a < nid; ++a)
        for (int b = 0; b < nid; ++b) {
            nterm[a * nid + b] = fourier_terms(p, zsp[a], zsp[b], terms + 2 * (a * nid + b));
            for (int t = 0; t < nterm[a * nid + b]; ++t) {
                const d3o_term *tt = terms + 2 * (a * nid + b) + t;
                /* kernel.f90:63-78 is_supported_term */
                if (!(tt->alpha > 0 && tt->m >= 0 && tt->alpha % 2 == 0 && tt->m % 2 == 0 && tt->m + 3 < tt->alpha
                      && tt->radius > 0.0)) {
                    free(terms); free(nterm);
                    return 2;
                }
                double kc = reciprocal_cutoff(tt, 1.0e-8);
                if (kc > kcut) kcut = kc;
            }
        }
    if (kcut <= 0.0) { free(terms); free(nterm); return 0; }
    if (kcut_req > 0.0) kcut = kcut_req;
    /* reciprocal vectors b_j (a_i . b_j = 2 pi delta_ij): row j of 2 pi M^-1,
     * M the matrix whose columns are the lattice vectors (ewald.f90:118) */
    double L[3][3], inv[3][3], rec[9];
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) L[r][c] = latflat[3 * r + c];
    matinv3(L, inv);
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) rec[3 * r + c] = 2.0 * M_PI * inv[r][c];
    double vol = fabs(L[0][0] * (L[1][1] * L[2][2] - L[1][2] * L[2][1]) - L[0][1] * (L[1][0] * L[2][2] - L[1][2] * L[2][0])
                      + L[0][2] * (L[1][0] * L[2][1] - L[1][1] * L[2][0]));
    int pbc3[3] = {1, 1, 1};
    int nk = d3o_lattice_points(pbc3, rec, kcut, NULL, 0);
    double *kp = (double *)malloc(sizeof(double) * 3 * nk);
    d3o_lattice_points(pbc3, rec, kcut, kp, nk);
    double *esum = (double *)calloc(nat, sizeof(double)), *dsum = (double *)calloc(nat, sizeof(double));
    double *gsum = (double *)calloc(3 * (size_t)nat, sizeof(double)), ssum[9] = {0};
    for (int i = 0; i < nat && owns_index(0); ++i) { /* ewald.f90:127: the self term on the first part */
        double pval = 0.0;
        int a = id[i] - 1;
        for (int t = 0; t < nterm[a * nid + a]; ++t) pval += potential_zero(terms + 2 * (a * nid + a) + t);
        for (int l = 0; l < R; ++l) {
            esum[i] += 0.5 * lr->lambda[l] * c6l[(size_t)i * R + l] * c6l[(size_t)i * R + l] * pval;
            if (grad) dsum[i] += lr->lambda[l] * c6l[(size_t)i * R + l] * dc6l[(size_t)i * R + l] * pval;
        }
    }
    double erecip = 0.0;
    for (int i = 0; i < nat; ++i) erecip -= esum[i];
    /* ewald.f90:150-240: vectors split over threads, per-thread sums */
#pragma omp parallel
    {
    double *phihat = (double *)malloc(sizeof(double) * nid * nid), *dphihat = (double *)malloc(sizeof(double) * nid * nid);
    double *sfr = (double *)malloc(sizeof(double) * R * nid), *sfi = (double *)malloc(sizeof(double) * R * nid);
    double *fr = (double *)malloc(sizeof(double) * R * nid), *fi = (double *)malloc(sizeof(double) * R * nid);
    double *cr = (double *)malloc(sizeof(double) * nat), *ci = (double *)malloc(sizeof(double) * nat);
    double *el = (double *)calloc(nat, sizeof(double)), *dl = (double *)calloc(nat, sizeof(double));
    double *gl = (double *)calloc(3 * (size_t)nat, sizeof(double)), sl[9] = {0};
#pragma omp for schedule(dynamic, 16)
    for (int ik = 0; ik < nk; ++ik) {
        if (ik % g_sample_stride) continue; /* bounded timing sample */
        if (!owns_index(ik)) continue;      /* ewald.f90:161 */
        const double *kv = kp + 3 * ik;
        double knorm = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
        if (knorm > kcut) continue;
        for (int a = 0; a < nid; ++a)
            for (int b = 0; b <= a; ++b) {
                double ph = 0.0, dph = 0.0;
                for (int t = 0; t < nterm[a * nid + b]; ++t) {
                    double pv, dpv;
                    fourier_transform(terms + 2 * (a * nid + b) + t, knorm, &pv, &dpv);
                    ph += pv; dph += dpv;
                }
                phihat[b * nid + a] = phihat[a * nid + b] = ph;
                dphihat[b * nid + a] = dphihat[a * nid + b] = dph;
            }
        memset(sfr, 0, sizeof(double) * R * nid);
        memset(sfi, 0, sizeof(double) * R * nid);
        for (int i = 0; i < nat; ++i) {
            double kr = kv[0] * xyz[3 * i] + kv[1] * xyz[3 * i + 1] + kv[2] * xyz[3 * i + 2];
            cr[i] = cos(kr); ci[i] = sin(kr);
            int a = id[i] - 1;
            for (int l = 0; l < R; ++l) {
                sfr[l * nid + a] += c6l[(size_t)i * R + l] * cr[i];
                sfi[l * nid + a] += c6l[(size_t)i * R + l] * ci[i];
            }
        }
        memset(fr, 0, sizeof(double) * R * nid);
        memset(fi, 0, sizeof(double) * R * nid);
        for (int a = 0; a < nid; ++a)
            for (int b = 0; b < nid; ++b)
                for (int l = 0; l < R; ++l) {
                    fr[l * nid + a] += phihat[b * nid + a] * sfr[l * nid + b];
                    fi[l * nid + a] += phihat[b * nid + a] * sfi[l * nid + b];
                }
        for (int i = 0; i < nat; ++i) {
            int a = id[i] - 1;
            for (int l = 0; l < R; ++l) {
                /* zval = phase * conj(fvec) */
                double zre = cr[i] * fr[l * nid + a] + ci[i] * fi[l * nid + a];
                double zim = ci[i] * fr[l * nid + a] - cr[i] * fi[l * nid + a];
                el[i] -= 0.5 * lr->lambda[l] * c6l[(size_t)i * R + l] * zre / vol;
                if (grad) {
                    dl[i] -= lr->lambda[l] * dc6l[(size_t)i * R + l] * zre / vol;
                    for (int c = 0; c < 3; ++c)
                        gl[3 * i + c] += lr->lambda[l] * c6l[(size_t)i * R + l] * zim * kv[c] / vol;
                }
            }
        }
        if (grad && knorm > 0.0) {
            double qq = 0.0;
            for (int a = 0; a < nid; ++a)
                for (int b = 0; b < nid; ++b)
                    for (int l = 0; l < R; ++l)
                        qq += lr->lambda[l] * dphihat[b * nid + a]
                            * (sfr[l * nid + a] * sfr[l * nid + b] + sfi[l * nid + a] * sfi[l * nid + b]);
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) sl[c * 3 + r] += 0.5 * qq * kv[r] * kv[c] / (knorm * vol);
        }
    }
#pragma omp critical(d3o_ewald_)
    {
    for (int i = 0; i < nat; ++i) {
        esum[i] += el[i];
        dsum[i] += dl[i];
        for (int c = 0; c < 3; ++c) gsum[3 * i + c] += gl[3 * i + c];
    }
    for (int k = 0; k < 9; ++k) ssum[k] += sl[k];
    }
    free(phihat); free(dphihat); free(sfr); free(sfi); free(fr); free(fi); free(cr); free(ci); free(el); free(dl);
    free(gl);
    }
    for (int i = 0; i < nat; ++i) energy[i] += esum[i];
    if (grad) {
        for (int i = 0; i < nat; ++i) erecip += esum[i];
        for (int c = 0; c < 3; ++c) ssum[c * 3 + c] -= erecip;
        for (int i = 0; i < nat; ++i) dEdcn[i] += dsum[i];
        for (int k = 0; k < 3 * nat; ++k) gradient[k] += gsum[k];
        for (int k = 0; k < 9; ++k) sigma[k] += ssum[k];
    }
    free(terms); free(nterm); free(kp); free(esum); free(dsum); free(gsum);
    return 0;
}

/* ------------------------------------------------------ particle mesh */
#define SPME_ORDER 6 /* spme.f90:47 spline_order */

/* fft.f90:48-60 valid_mesh_size: the next power of two */
static int valid_mesh(int npoint)
{
    int n = 1;
    while (n < npoint) n *= 2;
    return n;
}

/* spme.f90:457-499 fill_bspline: cardinal B-spline weights of order 6 at the
 * offset w in [0, 1) and their derivatives (from the order-5 weights). */
static void fill_bspline(double w, double *theta, double *dtheta)
{
    for (int k = 0; k < SPME_ORDER; ++k) theta[k] = 0.0;
    theta[0] = 1.0 - w;
    theta[1] = w;
    for (int korder = 3; korder <= SPME_ORDER - 1; ++korder) {
        double div = 1.0 / (korder - 1);
        theta[korder - 1] = div * w * theta[korder - 2];
        for (int j = 1; j <= korder - 2; ++j)
            theta[korder - j - 1] = div * ((w + j) * theta[korder - j - 2] + (korder - j - w) * theta[korder - j - 1]);
        theta[0] = div * (1.0 - w) * theta[0];
    }
    dtheta[0] = -theta[0];
    for (int j = 2; j <= SPME_ORDER; ++j) dtheta[j - 1] = theta[j - 2] - theta[j - 1];
    double div = 1.0 / (SPME_ORDER - 1);
    theta[SPME_ORDER - 1] = div * w * theta[SPME_ORDER - 2];
    for (int j = 1; j <= SPME_ORDER - 2; ++j)
        theta[SPME_ORDER - j - 1] =
            div * ((w + j) * theta[SPME_ORDER - j - 2] + (SPME_ORDER - j - w) * theta[SPME_ORDER - j - 1]);
    theta[0] = div * (1.0 - w) * theta[0];
}

/* Radix-2 transform of a strided complex sequence in place (re/im arrays),
 * exp(sign 2 pi i j k / n), unnormalised; fft.f90:136-188 (isign = sign). */
static void fft_1d(double *re, double *im, int n, size_t stride, int sign)
{
    if (n <= 1) return;
    for (int i = 1, j = 0; i < n; ++i) { /* bit reversal */
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i * stride]; re[i * stride] = re[j * stride]; re[j * stride] = t;
            t = im[i * stride]; im[i * stride] = im[j * stride]; im[j * stride] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len / 2;
        for (int k = 0; k < half; ++k) {
            double ang = sign * 2.0 * M_PI * k / len;
            double wr = cos(ang), wi = sin(ang);
            for (int i = k; i < n; i += len) {
                size_t a = (size_t)i * stride, b = (size_t)(i + half) * stride;
                double tr = wr * re[b] - wi * im[b], ti = wr * im[b] + wi * re[b];
                re[b] = re[a] - tr; im[b] = im[a] - ti;
                re[a] += tr; im[a] += ti;
            }
        }
    }
}

/* fft.f90:86-131 fft_3d on a grid [m3][m2][m1] (index g1 fastest). */
static void fft_3d(double *re, double *im, const int *m, int sign)
{
    for (int i3 = 0; i3 < m[2]; ++i3)
        for (int i2 = 0; i2 < m[1]; ++i2) {
            size_t o = ((size_t)i3 * m[1] + i2) * m[0];
            fft_1d(re + o, im + o, m[0], 1, sign);
        }
    for (int i3 = 0; i3 < m[2]; ++i3)
        for (int i1 = 0; i1 < m[0]; ++i1) {
            size_t o = (size_t)i3 * m[1] * m[0] + i1;
            fft_1d(re + o, im + o, m[1], m[0], sign);
        }
    for (int i2 = 0; i2 < m[1]; ++i2)
        for (int i1 = 0; i1 < m[0]; ++i1) {
            size_t o = (size_t)i2 * m[0] + i1;
            fft_1d(re + o, im + o, m[2], (size_t)m[0] * m[1], sign);
        }
}

/* fourier/spme.f90:78-328 get_dispersion_spme with spme.f90:55-74 (mesh from
 * the cutoff) and :346-529 (mesh kernel, spline weights, Euler factors): the
 * low-rank factors spread onto one mesh per species, transformed, contracted
 * with the kernel, transformed back and gathered onto the atoms, term by
 * term of the expansion.  mesh_req > 0 fixes the mesh points per direction. */
static int spme2(const d3o_param *p, int nat, const int *num, const int *id, const double *xyz,
                 const double *latflat, const d3o_lowrank *lr, const double *c6l, const double *dc6l,
                 double kcut_req, int mesh_req, const int *ghost, double *energy, double *dEdcn, double *gradient,
                 double *sigma, int grad)
{
    int nid = lr->nid, R = lr->rank;
    int zsp[MAX_ELEM] = {0};
    for (int i = 0; i < nat; ++i) zsp[id[i] - 1] = num[i];
    d3o_term *terms = (d3o_term *)calloc((size_t)nid * nid * 2, sizeof(d3o_term));
    int *nterm = (int *)calloc((size_t)nid * nid, sizeof(int));
    double kcut = 0.0;
    for (int a = 0; a < nid; ++a)
        for (int b = 0; b < nid; ++b) {
            nterm[a * nid + b] = fourier_terms(p, zsp[a], zsp[b], terms + 2 * (a * nid + b));
            for (int t = 0; t < nterm[a * nid + b]; ++t) {
                const d3o_term *tt = terms + 2 * (a * nid + b) + t;
                if (!(tt->alpha > 0 && tt->m >= 0 && tt->alpha % 2 == 0 && tt->m % 2 == 0 && tt->m + 3 < tt->alpha
                      && tt->radius > 0.0)) {
                    free(terms); free(nterm);
                    return 2;
                }
                double kc = reciprocal_cutoff(tt, 1.0e-8);
                if (kc > kcut) kcut = kc;
            }
        }
    if (kcut <= 0.0) { free(terms); free(nterm); return 0; }
    if (kcut_req > 0.0) kcut = kcut_req;
    double L[3][3], inv[3][3];
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) L[r][c] = latflat[3 * r + c];
    matinv3(L, inv); /* inv = M^-1, M columns = lattice vectors */
    double rec[9]; /* rows: b_j = 2 pi row j of M^-1 */
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) rec[3 * r + c] = 2.0 * M_PI * inv[r][c];
    double vol = fabs(L[0][0] * (L[1][1] * L[2][2] - L[1][2] * L[2][1]) - L[0][1] * (L[1][0] * L[2][2] - L[1][2] * L[2][0])
                      + L[0][2] * (L[1][0] * L[2][1] - L[1][1] * L[2][0]));
    int m[3];
    for (int d = 0; d < 3; ++d) {
        double len = sqrt(L[d][0] * L[d][0] + L[d][1] * L[d][1] + L[d][2] * L[d][2]);
        m[d] = mesh_req > 0 ? valid_mesh(mesh_req) : valid_mesh((int)ceil(kcut * len / M_PI));
    }
    size_t ng = (size_t)m[0] * m[1] * m[2];
    /* spline weights per atom and direction */
    int *base = (int *)malloc(sizeof(int) * 3 * nat);
    double *th = (double *)malloc(sizeof(double) * 3 * SPME_ORDER * nat);
    double *dth = (double *)malloc(sizeof(double) * 3 * SPME_ORDER * nat);
    for (int i = 0; i < nat; ++i)
        for (int d = 0; d < 3; ++d) {
            double f = inv[d][0] * xyz[3 * i] + inv[d][1] * xyz[3 * i + 1] + inv[d][2] * xyz[3 * i + 2];
            double u = m[d] * (f - floor(f));
            base[3 * i + d] = (int)floor(u) - SPME_ORDER;
            fill_bspline(u - floor(u), th + (3 * i + d) * SPME_ORDER, dth + (3 * i + d) * SPME_ORDER);
        }
    /* Euler factors */
    double *bfac = (double *)malloc(sizeof(double) * 3 * (m[0] > m[1] ? (m[0] > m[2] ? m[0] : m[2]) : (m[1] > m[2] ? m[1] : m[2])));
    int mmax = m[0] > m[1] ? (m[0] > m[2] ? m[0] : m[2]) : (m[1] > m[2] ? m[1] : m[2]);
    {
        double spl[SPME_ORDER], dspl[SPME_ORDER];
        fill_bspline(0.0, spl, dspl);
        for (int d = 0; d < 3; ++d)
            for (int im = 0; im < m[d]; ++im) {
                double dr = 0.0, di = 0.0;
                for (int jm = 0; jm < SPME_ORDER - 1; ++jm) {
                    double arg = 2.0 * M_PI * im * jm / m[d];
                    dr += spl[jm] * cos(arg);
                    di += spl[jm] * sin(arg);
                }
                bfac[d * mmax + im] = 1.0 / (dr * dr + di * di);
            }
    }
    /* mesh kernel */
    int npair = nid * (nid + 1) / 2;
    double *phik = (double *)malloc(sizeof(double) * npair * ng);
    double *dphik = (double *)malloc(sizeof(double) * npair * ng);
    double *kvecs = (double *)malloc(sizeof(double) * 3 * ng);
    for (int i3 = 0; i3 < m[2]; ++i3)
        for (int i2 = 0; i2 < m[1]; ++i2)
            for (int i1 = 0; i1 < m[0]; ++i1) {
                size_t g = ((size_t)i3 * m[1] + i2) * m[0] + i1;
                int j1 = i1 - m[0] * ((2 * i1) / m[0]), j2 = i2 - m[1] * ((2 * i2) / m[1]);
                int j3 = i3 - m[2] * ((2 * i3) / m[2]);
                double kv[3];
                for (int c = 0; c < 3; ++c) kv[c] = rec[c] * j1 + rec[3 + c] * j2 + rec[6 + c] * j3;
                memcpy(kvecs + 3 * g, kv, sizeof kv);
                double kn = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
                for (int a = 0; a < nid; ++a)
                    for (int b = 0; b <= a; ++b) {
                        double ph = 0.0, dph = 0.0;
                        for (int t = 0; t < nterm[a * nid + b]; ++t) {
                            double pv, dpv;
                            fourier_transform(terms + 2 * (a * nid + b) + t, kn, &pv, &dpv);
                            ph += pv; dph += dpv;
                        }
                        int pr = a * (a + 1) / 2 + b;
                        phik[(size_t)pr * ng + g] = ph;
                        dphik[(size_t)pr * ng + g] = dph;
                    }
            }
    double dudr[3][3]; /* du_d / dx_c = m_d rec_d[c] / 2 pi */
    for (int d = 0; d < 3; ++d) for (int c = 0; c < 3; ++c) dudr[d][c] = m[d] * rec[3 * d + c] / (2.0 * M_PI);
    double *esum = (double *)calloc(nat, sizeof(double)), *dsum = (double *)calloc(nat, sizeof(double));
    double *gsum = (double *)calloc(3 * (size_t)nat, sizeof(double)), ssum[9] = {0};
    for (int i = 0; i < nat; ++i) { /* spme.f90:157-171 self-interaction */
        int a = id[i] - 1;
        double pval = 0.0;
        for (int t = 0; t < nterm[a * nid + a]; ++t) pval += potential_zero(terms + 2 * (a * nid + a) + t);
        for (int l = 0; l < R; ++l) {
            if (!owns_index(l)) continue; /* spme.f90:166 */
            esum[i] += 0.5 * lr->lambda[l] * c6l[(size_t)i * R + l] * c6l[(size_t)i * R + l] * pval;
            if (grad) dsum[i] += lr->lambda[l] * c6l[(size_t)i * R + l] * dc6l[(size_t)i * R + l] * pval;
        }
    }
    double *qr = (double *)malloc(sizeof(double) * nid * ng), *qi = (double *)malloc(sizeof(double) * nid * ng);
    double *tr = (double *)malloc(sizeof(double) * ng), *ti = (double *)malloc(sizeof(double) * ng);
    double erecip = 0.0, escale = -0.5 / vol;
    for (int l = 0; l < R; ++l) {
        if (!owns_index(l)) continue; /* spme.f90:198 */
        memset(qr, 0, sizeof(double) * nid * ng);
        memset(qi, 0, sizeof(double) * nid * ng);
        for (int i = 0; i < nat; ++i) {
            if (ghost && ghost[i]) continue;
            int a = id[i] - 1;
            for (int j3 = 1; j3 <= SPME_ORDER; ++j3) {
                int g3 = ((base[3 * i + 2] + j3) % m[2] + m[2]) % m[2];
                for (int j2 = 1; j2 <= SPME_ORDER; ++j2) {
                    int g2 = ((base[3 * i + 1] + j2) % m[1] + m[1]) % m[1];
                    for (int j1 = 1; j1 <= SPME_ORDER; ++j1) {
                        int g1 = ((base[3 * i] + j1) % m[0] + m[0]) % m[0];
                        double w = th[(3 * i) * SPME_ORDER + j1 - 1] * th[(3 * i + 1) * SPME_ORDER + j2 - 1]
                                 * th[(3 * i + 2) * SPME_ORDER + j3 - 1];
                        qr[(size_t)a * ng + ((size_t)g3 * m[1] + g2) * m[0] + g1] += c6l[(size_t)i * R + l] * w;
                    }
                }
            }
        }
        for (int a = 0; a < nid; ++a) fft_3d(qr + (size_t)a * ng, qi + (size_t)a * ng, m, +1);
        for (int a = 0; a < nid; ++a) {
            for (size_t g = 0; g < ng; ++g) {
                int i1 = (int)(g % m[0]), i2 = (int)((g / m[0]) % m[1]), i3 = (int)(g / ((size_t)m[0] * m[1]));
                double sfac = bfac[i1] * bfac[mmax + i2] * bfac[2 * mmax + i3];
                double xr = 0.0, xi = 0.0;
                for (int b = 0; b < nid; ++b) {
                    int pr = a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a;
                    double kk = sfac * phik[(size_t)pr * ng + g];
                    xr += kk * qr[(size_t)b * ng + g];
                    xi += kk * qi[(size_t)b * ng + g];
                }
                tr[g] = xr; ti[g] = xi;
            }
            fft_3d(tr, ti, m, -1);
            for (int i = 0; i < nat; ++i) {
                if ((ghost && ghost[i]) || id[i] - 1 != a) continue;
                double w = 0.0, dw[3] = {0, 0, 0};
                for (int j3 = 1; j3 <= SPME_ORDER; ++j3) {
                    int g3 = ((base[3 * i + 2] + j3) % m[2] + m[2]) % m[2];
                    for (int j2 = 1; j2 <= SPME_ORDER; ++j2) {
                        int g2 = ((base[3 * i + 1] + j2) % m[1] + m[1]) % m[1];
                        for (int j1 = 1; j1 <= SPME_ORDER; ++j1) {
                            int g1 = ((base[3 * i] + j1) % m[0] + m[0]) % m[0];
                            double sf = tr[((size_t)g3 * m[1] + g2) * m[0] + g1];
                            double t1 = th[(3 * i) * SPME_ORDER + j1 - 1], t2 = th[(3 * i + 1) * SPME_ORDER + j2 - 1];
                            double t3 = th[(3 * i + 2) * SPME_ORDER + j3 - 1];
                            w += t1 * t2 * t3 * sf;
                            if (grad) {
                                double d1 = dth[(3 * i) * SPME_ORDER + j1 - 1] * t2 * t3;
                                double d2 = t1 * dth[(3 * i + 1) * SPME_ORDER + j2 - 1] * t3;
                                double d3 = t1 * t2 * dth[(3 * i + 2) * SPME_ORDER + j3 - 1];
                                for (int c = 0; c < 3; ++c)
                                    dw[c] += (dudr[0][c] * d1 + dudr[1][c] * d2 + dudr[2][c] * d3) * sf;
                            }
                        }
                    }
                }
                esum[i] += escale * lr->lambda[l] * c6l[(size_t)i * R + l] * w;
                erecip += escale * lr->lambda[l] * c6l[(size_t)i * R + l] * w;
                if (grad) {
                    dsum[i] += 2.0 * escale * lr->lambda[l] * dc6l[(size_t)i * R + l] * w;
                    for (int c = 0; c < 3; ++c)
                        gsum[3 * i + c] += 2.0 * escale * lr->lambda[l] * c6l[(size_t)i * R + l] * dw[c];
                }
            }
        }
        if (grad) { /* spme.f90:284-314 */
            for (size_t g = 0; g < ng; ++g) {
                const double *kv = kvecs + 3 * g;
                double kn = sqrt(kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2]);
                if (kn <= 0.0) continue;
                int i1 = (int)(g % m[0]), i2 = (int)((g / m[0]) % m[1]), i3 = (int)(g / ((size_t)m[0] * m[1]));
                double sfac = bfac[i1] * bfac[mmax + i2] * bfac[2 * mmax + i3];
                double qq = 0.0;
                for (int a = 0; a < nid; ++a)
                    for (int b = 0; b < nid; ++b) {
                        int pr = a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a;
                        qq += lr->lambda[l] * dphik[(size_t)pr * ng + g] * sfac
                            * (qr[(size_t)a * ng + g] * qr[(size_t)b * ng + g] + qi[(size_t)a * ng + g] * qi[(size_t)b * ng + g]);
                    }
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) ssum[c * 3 + r] += 0.5 * qq * kv[r] * kv[c] / (kn * vol);
            }
        }
    }
    for (int i = 0; i < nat; ++i) energy[i] += esum[i];
    if (grad) {
        for (int c = 0; c < 3; ++c) ssum[c * 3 + c] -= erecip;
        for (int i = 0; i < nat; ++i) dEdcn[i] += dsum[i];
        for (int k = 0; k < 3 * nat; ++k) gradient[k] += gsum[k];
        for (int k = 0; k < 9; ++k) sigma[k] += ssum[k];
    }
    free(terms); free(nterm); free(base); free(th); free(dth); free(bfac); free(phik); free(dphik); free(kvecs);
    free(esum); free(dsum); free(gsum); free(qr); free(qi); free(tr); free(ti);
    return 0;
}

/* The oracle's transform (for its check against an independent FFT): grid
 * [m3][m2][m1] complex, sign of the exponent. */
void d3o_fft3d(double *re, double *im, const int *m, int sign) { fft_3d(re, im, m, sign); }
void d3o_bspline(double w, double *theta, double *dtheta) { fill_bspline(w, theta, dtheta); }

/*
 * disp.f90:56-185 with a low-rank model (model.f90:185-196, 379-382): all C6
 * from the low-rank expansion; under three-dimensional periodicity the
 * two-body term is the reciprocal lattice sum (disp.f90:104-111, 142-145),
 * otherwise the real-space sum.  ewald = {rank, tolerance, kcut, mesh} with
 * the defaults of api.f90:447-451 for non-positive values; mesh < 0 sums the
 * reciprocal vectors directly, mesh >= 0 on the particle mesh (0: mesh from
 * the cutoff, damping.f90:365-380).  Returns 1 when the
 * damping function does not support the Ewald summation, 2 when a pair
 * potential has no closed-form transform (damping.f90:345-352).
 */
int d3o_get_dispersion_ewald(int nat, const int *num, const double *xyz, const double *lattice,
                             const int *periodic, int kind, const double *par, const double *cut,
                             const int *ghost, const double *ewald, double *energy, double *gradient,
                             double *sigma, double *energies_out)
{
    d3o_param p = {kind, par[0], par[1], par[2], par[3], par[4], par[5], par[6], par[7], par[8]};
    double zero_lat[9] = {0};
    int zero_pbc[3] = {0, 0, 0};
    if (!lattice) lattice = zero_lat;
    if (!periodic) periodic = zero_pbc;
    int recip = periodic[0] && periodic[1] && periodic[2];
    if (recip && !(kind == K_RATIONAL || (kind == K_ZERO && fabs(p.alp - lround(p.alp)) <= sqrt(EPS))))
        return 1;
    int grad = (gradient != NULL) || (sigma != NULL);
    int mref = 0;
    for (int i = 0; i < nat; ++i)
        if (g_nref[num[i] - 1] > mref) mref = g_nref[num[i] - 1];
    int *id = (int *)malloc(sizeof(int) * nat);
    d3o_lowrank lr;
    lowrank_new(nat, num, ewald[0] > 0 ? (int)ewald[0] : 0, ewald[1] > 0.0 ? ewald[1] : 1.0e-4, &lr, id);
    double *cn = (double *)calloc(nat, sizeof(double)), *trans;
    int ntr = lattice_alloc(periodic, lattice, cut[0], &trans);
    d3o_cn(nat, num, xyz, ntr, trans, cut[0], cn);
    free(trans);
    double *gw = (double *)calloc((size_t)nat * mref, sizeof(double));
    double *gwd = grad ? (double *)calloc((size_t)nat * mref, sizeof(double)) : NULL;
    d3o_weights(nat, num, cn, mref, gw, gwd);
    double *c6l = (double *)calloc((size_t)nat * lr.rank, sizeof(double));
    double *dc6l = (double *)calloc((size_t)nat * lr.rank, sizeof(double));
    lowrank_weights(&lr, nat, num, id, ghost, mref, gw, gwd, c6l, grad ? dc6l : NULL);
    double *c6 = (double *)calloc((size_t)nat * nat, sizeof(double));
    double *dc6 = grad ? (double *)calloc((size_t)nat * nat, sizeof(double)) : NULL;
    lowrank_c6(&lr, nat, c6l, grad ? dc6l : NULL, c6, dc6);
    double *en = (double *)calloc(nat, sizeof(double));
    double *dEdcn = grad ? (double *)calloc(nat, sizeof(double)) : NULL;
    double *gl = grad ? (double *)calloc(3 * (size_t)nat, sizeof(double)) : NULL;
    double sl[9] = {0};
    int rc = 0;
    if (recip && ewald[3] >= 0.0) {
        rc = spme2(&p, nat, num, id, xyz, lattice, &lr, c6l, dc6l, ewald[2], (int)ewald[3], ghost, en, dEdcn, gl, sl,
                   grad);
    } else if (recip) {
        rc = ewald2(&p, nat, num, id, xyz, lattice, &lr, c6l, dc6l, ewald[2], en, dEdcn, gl, sl, grad);
    } else {
        ntr = lattice_alloc(periodic, lattice, cut[1], &trans);
        disp2(&p, nat, num, xyz, ntr, trans, cut[1], cut[3], c6, dc6, en, dEdcn, gl, sl, grad);
        free(trans);
    }
    if (rc == 0) {
        ntr = lattice_alloc(periodic, lattice, cut[2], &trans);
        atm(nat, num, xyz, ntr, trans, cut[2], cut[4], p.s9, 4.0 / 3.0, p.alp + 2.0, c6, dc6, en, dEdcn, gl, sl,
            grad);
        free(trans);
        if (grad) {
            ntr = lattice_alloc(periodic, lattice, cut[0], &trans);
            d3o_cn_derivs(nat, num, xyz, ntr, trans, cut[0], dEdcn, gl, sl);
            free(trans);
            if (gradient) memcpy(gradient, gl, sizeof(double) * 3 * nat);
            if (sigma) memcpy(sigma, sl, sizeof sl);
        }
        double e = 0.0;
        for (int i = 0; i < nat; ++i) e += en[i];
        *energy = e;
        if (energies_out) memcpy(energies_out, en, sizeof(double) * nat);
    }
    free(id); free(lr.lambda); free(lr.vec); free(cn); free(gw); free(gwd); free(c6l); free(dc6l); free(c6);
    free(dc6); free(en); free(dEdcn); free(gl);
    return rc;
}

/* The low-rank expansion of a structure's species (decomposition.f90:104-160):
 * returns the rank; error receives the largest relative error of the kept
 * expansion, lambda (if not null, >= ndim entries) the kept eigenvalues and
 * ndim_out the size of the reference C6 matrix. */
int d3o_lowrank_info(int nat, const int *num, int rank, double tolerance, double *error, double *lambda,
                     int *ndim_out)
{
    int *id = (int *)malloc(sizeof(int) * nat);
    d3o_lowrank lr;
    lowrank_new(nat, num, rank, tolerance > 0.0 ? tolerance : 1.0e-4, &lr, id);
    if (error) *error = lr.error;
    if (ndim_out) *ndim_out = lr.ndim;
    if (lambda) memcpy(lambda, lr.lambda, sizeof(double) * lr.rank);
    int r = lr.rank;
    free(id); free(lr.lambda); free(lr.vec);
    return r;
}

/* Pair C6 of the low-rank expansion from given reference weights gw
 * [nat][mref] (decomposition.f90:247-296); c6 nat x nat. */
void d3o_lowrank_atomic_c6(int nat, const int *num, int rank, double tolerance, int mref, const double *gw,
                           const int *ghost, double *c6)
{
    int *id = (int *)malloc(sizeof(int) * nat);
    d3o_lowrank lr;
    lowrank_new(nat, num, rank, tolerance > 0.0 ? tolerance : 1.0e-4, &lr, id);
    double *c6l = (double *)calloc((size_t)nat * lr.rank, sizeof(double));
    lowrank_weights(&lr, nat, num, id, ghost, mref, gw, NULL, c6l, NULL);
    lowrank_c6(&lr, nat, c6l, NULL, c6, NULL);
    free(id); free(lr.lambda); free(lr.vec); free(c6l);
}

/* Eigenpairs of a symmetric matrix (column-major) by the cyclic Jacobi
 * method of the reference, sorted by decreasing magnitude. */
void d3o_jacobi(int ndim, const double *amat, double *eval, double *evec) { jacobi_eig(ndim, amat, eval, evec); }

/* Fourier transform of pref r^m / (r^alpha + radius^alpha) at wave number k
 * (kernel.f90:83-129), its value at the origin and reciprocal cutoff. */
void d3o_fourier_term(double pref, int m, int alpha, double radius, double k, double *phi, double *dphi,
                      double *pot0, double *kcut)
{
    d3o_term t = {pref, m, alpha, radius};
    fourier_transform(&t, k, phi, dphi);
    if (pot0) *pot0 = potential_zero(&t);
    if (kcut) *kcut = reciprocal_cutoff(&t, 1.0e-8);
}

/* Count of the terms the reference loops visit (the work units). */
void d3o_count_terms(int nat, const double *xyz, const double *lattice, const int *periodic,
                     const double *cut, long long *n_cn, long long *n_pair, long long *n_triple)
{
    double *trans;
    double c2;
    int ntr;
    long long cnt = 0;
    ntr = lattice_alloc(periodic, lattice, cut[0], &trans);
    c2 = cut[0] * cut[0];
#pragma omp parallel for reduction(+ : cnt) schedule(static)
    for (int i = 0; i < nat; ++i)
        for (int j = 0; j <= i; ++j)
            for (int t = 0; t < ntr; ++t) {
                double v[3];
                for (int c = 0; c < 3; ++c) v[c] = xyz[3 * i + c] - (xyz[3 * j + c] + trans[3 * t + c]);
                double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                if (!(r2 > c2 || r2 < 1.0e-12)) ++cnt;
            }
    free(trans);
    *n_cn = cnt;
    cnt = 0;
    ntr = lattice_alloc(periodic, lattice, cut[1], &trans);
    c2 = cut[1] * cut[1];
#pragma omp parallel for reduction(+ : cnt) schedule(static)
    for (int i = 0; i < nat; ++i)
        for (int j = 0; j <= i; ++j)
            for (int t = 0; t < ntr; ++t) {
                double v[3];
                for (int c = 0; c < 3; ++c) v[c] = xyz[3 * i + c] - (xyz[3 * j + c] + trans[3 * t + c]);
                double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                if (!(r2 > c2 || r2 < EPS)) ++cnt;
            }
    free(trans);
    *n_pair = cnt;
    if (!n_triple) return;
    cnt = 0;
    ntr = lattice_alloc(periodic, lattice, cut[2], &trans);
    c2 = cut[2] * cut[2];
#pragma omp parallel for reduction(+ : cnt) schedule(dynamic)
    for (int i = 0; i < nat; ++i)
        for (int j = 0; j <= i; ++j)
            for (int tj = 0; tj < ntr; ++tj) {
                double vij[3];
                for (int c = 0; c < 3; ++c) vij[c] = xyz[3 * j + c] + trans[3 * tj + c] - xyz[3 * i + c];
                double r2ij = vij[0] * vij[0] + vij[1] * vij[1] + vij[2] * vij[2];
                if (r2ij > c2 || r2ij < EPS) continue;
                for (int k = 0; k <= j; ++k)
                    for (int tk = 0; tk < ntr; ++tk) {
                        double vik[3], vjk[3];
                        for (int c = 0; c < 3; ++c) vik[c] = xyz[3 * k + c] + trans[3 * tk + c] - xyz[3 * i + c];
                        double r2ik = vik[0] * vik[0] + vik[1] * vik[1] + vik[2] * vik[2];
                        if (r2ik > c2 || r2ik < EPS) continue;
                        for (int c = 0; c < 3; ++c) vjk[c] = vik[c] - vij[c];
                        double r2jk = vjk[0] * vjk[0] + vjk[1] * vjk[1] + vjk[2] * vjk[2];
                        if (r2jk > c2 || r2jk < EPS) continue;
                        ++cnt;
                    }
            }
    free(trans);
    *n_triple = cnt;
}

int d3o_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Threads of the following parallel loops (timing of the baseline). */
void d3o_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
