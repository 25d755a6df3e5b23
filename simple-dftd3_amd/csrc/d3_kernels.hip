// HIP kernels of the D3 energy / gradient / virial path (gfx950, wave64).
//
// Reference: simple-dftd3 v1.6.0, src/dftd3/disp.f90:56-185 and the loops it
// calls.  Every kernel visits the reference's pair / triple terms: the same
// lattice images, the same pair-vector arithmetic for the cutoff tests and
// the same triangular triple index set; only the order of summation differs.
// Term arithmetic runs in the template type R (double: plain path, float:
// mixed path); energies, gradients, virials and dE/dCN are accumulated in
// double in both paths, and every cutoff test uses the double-precision
// squared distance, so both paths visit exactly the reference's terms.
#include <hip/hip_runtime.h>

#include <cmath>

#include "d3_kernels.hpp"

namespace d3 {
namespace {

constexpr double kEps = 2.220446049250313e-16;  // epsilon(1.0_wp)
constexpr double kKcn = 16.0;                   // ncoord.f90:31
constexpr double kCnFloor = 1.0e-12;            // ncoord.f90:127
constexpr double kRs9 = 4.0 / 3.0;              // rational.f90:73, zero.f90:73

// ------------------------------------------------------------ helpers
template <typename R> struct Vec4T;
template <> struct Vec4T<double> { using type = double4; };
template <> struct Vec4T<float> { using type = float4; };
template <typename R> using R4 = typename Vec4T<R>::type;

template <typename R> __device__ __forceinline__ R4<R> make_r4(R x, R y, R z, R w);
template <> __device__ __forceinline__ double4 make_r4<double>(double x, double y, double z, double w) {
  return make_double4(x, y, z, w);
}
template <> __device__ __forceinline__ float4 make_r4<float>(float x, float y, float z, float w) {
  return make_float4(x, y, z, w);
}

// One slot (j, T) of the three-body list L(i) of atom i with the quantities
// of the pair (i, j) that every triple of the slot uses.
template <typename R>
struct Slot {
  double vx, vy, vz;  // x_j + T - x_i (atm.f90:711), FP64 for the exact cutoff tests
  R r2;               // |v|^2
  R rinv;             // 1 / |v|
  R pdmp;             // (r0_ij / r_ij)^(alp/3), r0_ij = 4/3 rvdw_ij, alp = alp6 + 2
  R sc6;              // sqrt(C6_ij)
  R qi;               // (dC6_ij / dCN_i) / C6_ij
  R qx;               // (dC6_ij / dCN_j) / C6_ij
  int atom;           // j (global index)
  int pad;
};

__device__ __forceinline__ int wave_lane() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Pair vector of the reference tuple for the ordered pair (i, j) and image T
// of the calling thread.  For j <= i the reference visits (iat=i, jat=j, T)
// with vec = x_i - (x_j + T) (rational.f90:462, ncoord.f90:124); for j > i it
// visits (iat=j, jat=i, -T) with vec = x_j - (x_i - T).  Returns r2 exactly
// as the reference evaluates it (no contraction) and u, the vector from the
// image of j to i.
__device__ __forceinline__ double pair_vec(const double4& xi, const double4& xj, const double4& t, bool swap,
                                           double& u0, double& u1, double& u2) {
#pragma clang fp contract(off)
  double v0, v1, v2;
  if (!swap) {
    v0 = xi.x - (xj.x + t.x);
    v1 = xi.y - (xj.y + t.y);
    v2 = xi.z - (xj.z + t.z);
    u0 = v0; u1 = v1; u2 = v2;
  } else {
    v0 = xj.x - (xi.x - t.x);
    v1 = xj.y - (xi.y - t.y);
    v2 = xj.z - (xi.z - t.z);
    u0 = -v0; u1 = -v1; u2 = -v2;
  }
  return v0 * v0 + v1 * v1 + v2 * v2;
}

// Three-body vector x_j + T - x_i (atm.f90:711,721) and its squared length,
// evaluated as the reference does.
__device__ __forceinline__ double atm_vec(const double4& xj, const double4& t, const double4& xi, double& v0,
                                          double& v1, double& v2) {
#pragma clang fp contract(off)
  v0 = xj.x + t.x - xi.x;
  v1 = xj.y + t.y - xi.y;
  v2 = xj.z + t.z - xi.z;
  return v0 * v0 + v1 * v1 + v2 * v2;
}

// vjk = vik - vij (atm.f90:723) and its squared length, as the reference.
__device__ __forceinline__ double diff_norm2(double bx, double by, double bz, double ax, double ay, double az,
                                             double& v0, double& v1, double& v2) {
#pragma clang fp contract(off)
  v0 = bx - ax;
  v1 = by - ay;
  v2 = bz - az;
  return v0 * v0 + v1 * v1 + v2 * v2;
}

// cutoff.f90:73-99 smooth_cutoff
template <typename R>
__device__ __forceinline__ void smooth_cutoff(R r, R cutoff, R width, R& sw, R& dswdr) {
  if (width <= R(0) || width >= cutoff || r <= cutoff - width) {
    sw = R(1); dswdr = R(0);
  } else if (r >= cutoff) {
    sw = R(0); dswdr = R(0);
  } else {
    R x = (cutoff - r) / width;
    sw = x * x * x * (R(10) + x * (R(-15) + R(6) * x));
    dswdr = R(-30) * x * x * (R(1) - x) * (R(1) - x) / width;
  }
}

template <typename R> __device__ __forceinline__ R d_exp(R x);
template <> __device__ __forceinline__ double d_exp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float d_exp<float>(float x) { return __expf(x); }
template <typename R> __device__ __forceinline__ R d_sqrt(R x);
template <> __device__ __forceinline__ double d_sqrt<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float d_sqrt<float>(float x) { return __fsqrt_rn(x); }
template <typename R> __device__ __forceinline__ R d_rsqrt(R x);
template <> __device__ __forceinline__ double d_rsqrt<double>(double x) { return rsqrt(x); }
template <> __device__ __forceinline__ float d_rsqrt<float>(float x) { return __frsqrt_rn(x); }
template <typename R> __device__ __forceinline__ R d_rcp(R x);
template <> __device__ __forceinline__ double d_rcp<double>(double x) { return 1.0 / x; }
template <> __device__ __forceinline__ float d_rcp<float>(float x) { return __frcp_rn(x); }
// x**y for x > 0
template <typename R> __device__ __forceinline__ R d_pow(R x, R y);
template <> __device__ __forceinline__ double d_pow<double>(double x, double y) { return pow(x, y); }
template <> __device__ __forceinline__ float d_pow<float>(float x, float y) { return exp2f(y * __log2f(x)); }

// atm.f90:954-980
__device__ __forceinline__ double triple_scale(int ii, int jj, int kk) {
  if (ii == jj) return ii == kk ? 1.0 / 6.0 : 0.5;
  if (ii != kk && jj != kk) return 1.0;
  return 0.5;
}

// Sum v over the block and add the total to *dst (thread 0).  Every thread
// of the block must call it.
__device__ __forceinline__ void block_add(double* __restrict__ dst, double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if (wave_lane() == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
    if (t != 0.0) atomicAdd(dst, t);
  }
  __syncthreads();
}

int blocks_for(long long n, int per_block) { return static_cast<int>((n + per_block - 1) / per_block); }

// --------------------------------------------------------------- packing
__global__ void pack_positions(const double* __restrict__ xyz, int n, double4* __restrict__ pos) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pos[i] = make_double4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.0);
}

// ------------------------------------------------------- coordination number
// ncoord.f90:86-144, one wavefront per atom i over all partners j and images
// (each ordered pair once; the reference's triangle counts it for both atoms).
template <typename R>
__global__ __launch_bounds__(kBlock) void cn_kernel(DevBatch b, double* __restrict__ cn) {
  const int lane = wave_lane();
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= b.natoms) return;
  const int s = b.sys[i];
  const SysGeom g = b.geom[s];
  const double cut = b.par[s].cut_cn;
  const double cut2 = cut * cut;
  const double4 xi = b.pos[i];
  const R rci = R(b.rcov[b.sp[i]]);
  double acc = 0.0;
  for (int j = g.atom_begin + lane; j < g.atom_begin + g.natoms; j += kWave) {
    const double4 xj = b.pos[j];
    const R rc = rci + R(b.rcov[b.sp[j]]);
    const bool swap = j > i;
    for (int t = 0; t < g.img_cn_count; ++t) {
      double u0, u1, u2;
      const double r2 = pair_vec(xi, xj, b.img[g.img_cn_begin + t], swap, u0, u1, u2);
      if (r2 > cut2 || r2 < kCnFloor) continue;
      const R r1 = d_sqrt(R(r2));
      acc += double(d_rcp(R(1) + d_exp(-R(kKcn) * (rc / r1 - R(1)))));
    }
  }
  acc = wave_sum(acc);
  if (lane == 0) cn[i] = acc;
}

// ------------------------------------------------------------ weights
// model.f90:228-287 weight_references with wf = 4 (model.f90:80), FP64.
__global__ void weights_kernel(DevBatch b, const double* __restrict__ cn, double* __restrict__ gw,
                               double* __restrict__ dgw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.natoms) return;
  const double wf = 4.0, wf2 = 2.0 * wf;
  const int s = b.sp[i];
  const int nr = b.nref[s];
  const double* rcn = b.refcn + s * kMaxRef;
  const double c = cn[i];
  double gwt[kMaxRef];
  double norm = 0.0, dnorm = 0.0, maxcn = -1.0e300;
  for (int r = 0; r < nr; ++r) {
    const double d = c - rcn[r];
    const double w = exp(-wf * (d * d));
    gwt[r] = w;
    norm += w;
    dnorm += wf2 * (rcn[r] - c) * w;
    maxcn = fmax(maxcn, rcn[r]);
  }
  norm = 1.0 / norm;
  for (int r = 0; r < b.mref; ++r) {
    double gwk = 0.0, dgwk = 0.0;
    if (r < nr) {
      const double expw = gwt[r];
      const double expd = wf2 * (rcn[r] - c) * expw;
      gwk = expw * norm;
      if (isnan(gwk) || isinf(gwk)) gwk = (maxcn == rcn[r]) ? 1.0 : 0.0;  // model.f90:260-266
      dgwk = expd * norm - expw * dnorm * (norm * norm);
      if (isnan(dgwk) || isinf(dgwk)) dgwk = 0.0;  // model.f90:269-273
    }
    gw[i * b.mref + r] = gwk;
    if (dgw) dgw[i * b.mref + r] = dgwk;
  }
}

// ---------------------------------------------------------- pair records
// model.f90:432-462: C6_ij = sum_kl w_ik w_jl C6ref_kl and its derivatives,
// stored per ordered pair as {sqrt(C6_ij), (dC6_ij/dCN_i)/C6_ij,
// (dC6_ij/dCN_j)/C6_ij, (4/3 rvdw_ij)^(alp/3)} (the last two feed the ATM
// term, atm.f90:667-668,811-816).  One block per row i of the system block.
template <typename R>
__global__ __launch_bounds__(kBlock) void pair_kernel(DevBatch b, const double* __restrict__ gw,
                                                      const double* __restrict__ dgw, R4<R>* __restrict__ pair) {
  const int i = blockIdx.x;
  const int s = b.sys[i];
  const SysGeom g = b.geom[s];
  const double alp3 = (b.par[s].alp + 2.0) / 3.0;
  const int n = g.natoms, il = i - g.atom_begin;
  const int spi = b.sp[i];
  const int mref = b.mref;
  const int nri = b.nref[spi];
  double gi[kMaxRef], dgi[kMaxRef];
  for (int r = 0; r < kMaxRef; ++r) {
    gi[r] = r < nri ? gw[i * mref + r] : 0.0;
    dgi[r] = (r < nri && dgw) ? dgw[i * mref + r] : 0.0;
  }
  const bool gi_ghost = b.ghost && b.ghost[i];
  R4<R>* row = pair + g.c6_begin + static_cast<long long>(il) * n;
  for (int jl = threadIdx.x; jl < n; jl += blockDim.x) {
    const int j = g.atom_begin + jl;
    const int spj = b.sp[j];
    const int nrj = b.nref[spj];
    const double* ref = b.c6ref + (spi * b.nsp + spj) * (kMaxRef * kMaxRef);
    double cc = 0.0, di = 0.0, dj = 0.0;
    if (!(gi_ghost || (b.ghost && b.ghost[j]))) {  // model.f90:436-440
      for (int ri = 0; ri < nri; ++ri) {
        for (int rj = 0; rj < nrj; ++rj) {
          const double w = gw[j * mref + rj];
          const double rc = ref[ri * kMaxRef + rj];
          cc += gi[ri] * w * rc;
          if (dgw) {
            di += dgi[ri] * w * rc;
            dj += gi[ri] * dgw[j * mref + rj] * rc;
          }
        }
      }
    }
    const double ic = cc > 0.0 ? 1.0 / cc : 0.0;
    const double r0p = pow(kRs9 * b.rvdw[spj * b.nsp + spi], alp3);
    row[jl] = make_r4<R>(R(sqrt(cc)), R(di * ic), R(dj * ic), R(r0p));
  }
}

// --------------------------------------------------------------- two-body
// rational.f90:378-514 (BJ) and zero.f90:409-553 (zero damping), E+grad+sigma;
// rational.f90:282-374 / zero.f90:306-405 when kGrad is false.  One wavefront
// per atom i; atom i collects the ordered pairs (i, j, T) it takes part in.
template <typename R, bool kGrad>
__global__ __launch_bounds__(kBlock) void disp2_kernel(DevBatch b, const R4<R>* __restrict__ pair,
                                                       double* __restrict__ e_atom, double* __restrict__ dedcn,
                                                       double* __restrict__ grad, double* __restrict__ sig_atom) {
  const int lane = wave_lane();
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= b.natoms) return;
  const int s = b.sys[i];
  const SysGeom g = b.geom[s];
  const SysParam p = b.par[s];
  const int n = g.natoms, il = i - g.atom_begin;
  const double cut2 = p.cut2 * p.cut2;
  const double4 xi = b.pos[i];
  const int spi = b.sp[i];
  const R r4r2i = R(b.r4r2[spi]);
  double acc_e = 0.0, acc_d = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
  double sxx = 0.0, syy = 0.0, szz = 0.0, sxy = 0.0, sxz = 0.0, syz = 0.0;
  const bool active = !(fabs(p.s6) < kEps && fabs(p.s8) < kEps);  // rational.f90:267
  const R s6 = R(p.s6), s8 = R(p.s8), cutr = R(p.cut2), width = R(p.width2);
  const R4<R>* row = pair + g.c6_begin + static_cast<long long>(il) * n;
  if (active) {
    for (int jl = lane; jl < n; jl += kWave) {
      const int j = g.atom_begin + jl;
      const double4 xj = b.pos[j];
      const int spj = b.sp[j];
      const R rrij = R(3) * r4r2i * R(b.r4r2[spj]);
      const R4<R> q = row[jl];
      const R c6ij = q.x * q.x;
      const R dc6i = q.y * c6ij;  // dC6_ij/dCN_i
      R r0a = R(0), r0b = R(0);   // BJ: r0^6, r0^8; zero: rs6*R0, rs8*R0
      if (p.kind == kRational) {
        const R r0 = R(p.a1) * d_sqrt(rrij) + R(p.a2);
        const R r02 = r0 * r0;
        r0a = r02 * r02 * r02;
        r0b = r0a * r02;
      } else {
        const R r0 = R(b.rvdw[spj * b.nsp + spi]);
        r0a = R(p.a1) * r0;
        r0b = R(p.a2) * r0;
      }
      const bool swap = j > i;
      for (int t = 0; t < g.img2_count; ++t) {
        double u0, u1, u2;
        const double r2d = pair_vec(xi, xj, b.img[g.img2_begin + t], swap, u0, u1, u2);
        if (r2d > cut2 || r2d < kEps) continue;
        const R r2 = R(r2d);
        const R r1 = d_sqrt(r2);
        R sw, dswdr;
        smooth_cutoff(r1, cutr, width, sw, dswdr);
        R edisp0, gdisp0 = R(0);
        if (p.kind == kRational) {
          const R r4 = r2 * r2;
          const R t6 = d_rcp(r4 * r2 + r0a);
          const R t8 = d_rcp(r4 * r4 + r0b);
          edisp0 = s6 * t6 + s8 * rrij * t8;
          if (kGrad) {
            const R d6 = R(-6) * r4 * t6 * t6;
            const R d8 = R(-8) * r4 * r2 * t8 * t8;
            gdisp0 = s6 * d6 + s8 * rrij * d8;
          }
        } else {
          const R alp6 = R(p.alp), alp8 = R(p.alp) + R(2);
          const R r6 = r2 * r2 * r2, r8 = r6 * r2;
          const R ir1 = d_rcp(r1);
          const R t6 = d_pow(r0a * ir1, alp6);
          const R t8 = d_pow(r0b * ir1, alp8);
          const R f6 = d_rcp(R(1) + R(6) * t6);
          const R f8 = d_rcp(R(1) + R(6) * t8);
          const R ir6 = d_rcp(r6), ir8 = d_rcp(r8);
          edisp0 = s6 * f6 * ir6 + s8 * rrij * f8 * ir8;
          if (kGrad) {
            const R ir2 = d_rcp(r2);
            const R d6 = R(-6) * f6 * ir2 + R(6) * alp6 * t6 * f6 * f6 * ir2;
            const R d8 = R(-8) * f8 * ir2 + R(6) * alp8 * t8 * f8 * f8 * ir2;
            gdisp0 = s6 * d6 * ir6 + s8 * rrij * d8 * ir8;
          }
        }
        const R edisp = sw * edisp0;
        acc_e += double(-c6ij * edisp * R(0.5));
        if (kGrad) {
          const R gdisp = sw * gdisp0 + dswdr * edisp0 / r1;
          const R f = -c6ij * gdisp;  // dG = f * vec
          acc_d -= double(dc6i * edisp);
          const R v0 = R(u0), v1 = R(u1), v2 = R(u2);
          if (j != i) {
            g0 += double(f * v0); g1 += double(f * v1); g2 += double(f * v2);
          }
          const R h = R(0.5) * f;  // dS = 0.5 dG (x) vec
          sxx += double(h * v0 * v0); syy += double(h * v1 * v1); szz += double(h * v2 * v2);
          sxy += double(h * v0 * v1); sxz += double(h * v0 * v2); syz += double(h * v1 * v2);
        }
      }
    }
  }
  acc_e = wave_sum(acc_e);
  if (kGrad) {
    acc_d = wave_sum(acc_d);
    g0 = wave_sum(g0); g1 = wave_sum(g1); g2 = wave_sum(g2);
    sxx = wave_sum(sxx); syy = wave_sum(syy); szz = wave_sum(szz);
    sxy = wave_sum(sxy); sxz = wave_sum(sxz); syz = wave_sum(syz);
  }
  if (lane == 0) {
    e_atom[i] = acc_e;
    if (kGrad) {
      dedcn[i] = acc_d;
      grad[3 * i] = g0; grad[3 * i + 1] = g1; grad[3 * i + 2] = g2;
      double* sg = sig_atom + 6 * i;
      sg[0] = sxx; sg[1] = syy; sg[2] = szz; sg[3] = sxy; sg[4] = sxz; sg[5] = syz;
    }
  }
}

// ------------------------------------------------------ three-body lists
// L(i) = {(j, T): j <= i, eps <= |x_j + T - x_i|^2 <= cut3^2, C6_ij > 0}
// (atm.f90:706-716; pairs with a vanishing C6 -- ghost atoms -- contribute
// nothing to either branch of the reference, atm.f90:725-726), in (j, T)
// order.  The triples of atom i are the unordered pairs {a, b} of distinct
// slots of L(i) with eps <= |v_b - v_a|^2 <= cut3^2 (atm.f90:717-742): a pair
// of different atoms is the reference's triple with the larger index as j, a
// pair of two images of the same atom stands for the two triples (j, Tj, j,
// Tk) and (j, Tk, j, Tj) of the reference, which have equal contributions.
template <bool kFill, typename R>
__global__ __launch_bounds__(kBlock) void atm_list_kernel(DevBatch b, const R4<R>* __restrict__ pair,
                                                          long long* __restrict__ count,
                                                          const long long* __restrict__ offset,
                                                          Slot<R>* __restrict__ slots) {
  const int lane = wave_lane();
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= b.natoms) return;
  const int s = b.sys[i];
  const SysGeom g = b.geom[s];
  const SysParam p = b.par[s];
  if (fabs(p.s9) < kEps) {
    if (!kFill && lane == 0) count[i] = 0;
    return;
  }
  const double cut2 = p.cut3 * p.cut3;
  const R malp6 = -(R(p.alp) + R(2)) / R(6);  // exponent of r2 in (r0/r)^(alp/3)
  const double4 xi = b.pos[i];
  const int n = g.natoms, il = i - g.atom_begin;
  const R4<R>* row = pair ? pair + g.c6_begin + static_cast<long long>(il) * n : nullptr;
  Slot<R>* out = kFill ? slots + offset[i] : nullptr;
  long long total = 0;
  for (int j0 = g.atom_begin; j0 <= i; j0 += kWave) {
    const int j = j0 + lane;
    int cnt = 0;
    bool valid = j <= i;
    R4<R> q{};
    if (valid && row) {
      q = row[j - g.atom_begin];
      valid = !(double(q.x) * double(q.x) < kEps);
    }
    const double4 xj = valid ? b.pos[j] : xi;
    if (valid) {
      for (int t = 0; t < g.img3_count; ++t) {
        double v0, v1, v2;
        const double r2 = atm_vec(xj, b.img[g.img3_begin + t], xi, v0, v1, v2);
        if (r2 > cut2 || r2 < kEps) continue;
        ++cnt;
      }
    }
    int incl = cnt;  // inclusive scan of the lane counts
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      int y = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += y;
    }
    const int wtot = __shfl(incl, kWave - 1, kWave);
    if (kFill && valid && cnt > 0) {
      long long w = total + incl - cnt;
      for (int t = 0; t < g.img3_count; ++t) {
        double v0, v1, v2;
        const double r2 = atm_vec(xj, b.img[g.img3_begin + t], xi, v0, v1, v2);
        if (r2 > cut2 || r2 < kEps) continue;
        Slot<R> e;
        e.vx = v0; e.vy = v1; e.vz = v2;
        e.r2 = R(r2);
        e.rinv = R(1.0 / sqrt(r2));
        e.pdmp = q.w * d_pow(R(r2), malp6);
        e.sc6 = q.x;
        e.qi = q.y;
        e.qx = q.z;
        e.atom = j;
        e.pad = 0;
        out[w++] = e;
      }
    }
    total += wtot;
  }
  if (!kFill && lane == 0) count[i] = total;
}

// Exclusive scan of n counts into offset[0..n] (single block).
__global__ void scan_kernel(const long long* __restrict__ count, int n, long long* __restrict__ offset) {
  __shared__ long long part[1024];
  const int tid = threadIdx.x, nt = blockDim.x;
  const int chunk = (n + nt - 1) / nt;
  const int lo = min(n, tid * chunk), hi = min(n, lo + chunk);
  long long sum = 0;
  for (int k = lo; k < hi; ++k) sum += count[k];
  part[tid] = sum;
  __syncthreads();
  if (tid == 0) {
    long long run = 0;
    for (int k = 0; k < nt; ++k) {
      long long v = part[k];
      part[k] = run;
      run += v;
    }
    offset[n] = run;
  }
  __syncthreads();
  long long run = part[tid];
  for (int k = lo; k < hi; ++k) {
    offset[k] = run;
    run += count[k];
  }
}

// Number of slot blocks (work items) per atom.
__global__ void atm_items_kernel(const long long* __restrict__ offset, int n, long long* __restrict__ items) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) items[i] = (offset[i + 1] - offset[i] + kTile - 1) / kTile;
}

// --------------------------------------------------------- three-body term
template <typename R>
struct TileLds {  // b-slots of the current tile of one wavefront
  double vx[kTile], vy[kTile], vz[kTile];
  R r2[kTile], rinv[kTile], pdmp[kTile], sc6[kTile];
  int atom[kTile];
  double acc[6][kTile];  // F, E, Gx, Gy, Gz, Dq of the b-slots
};

// atm.f90:625-825 get_atm_dispersion_derivs (kGrad) and atm.f90:500-621,
// with rs9 = 4/3 and alp + 2 (rational.f90:563, zero.f90:602).
//
// Work item = (atom i, slot block P of L(i)); the four wavefronts of the
// block take the slot blocks Q >= P in turn.  In a 64 x 64 tile, lane l owns
// slot a = P*64 + l and at step s pairs it with b = Q*64 + (l + s) mod 64, so
// the b-slots of a step are distinct: their contributions are added into LDS
// without conflicts, the a-slot contributions stay in registers.  A diagonal
// tile (Q == P) runs the steps 1..32 (32 only for l < 32) so that every
// unordered pair is visited once.
//
// Per triple the contributions are condensed into per-slot scalars: with
// f_xy the radial factor of pair xy (dG_xy = f_xy v_xy, atm.f90:754-790),
//   F_s   = sum f_is   over the triples of slot s   -> grad_x += F_s v_s,
//                                                      grad_i -= F_s v_s,
//                                                      sigma  += F_s v_s v_s
//   E_s   = sum dE                                  -> dEdcn_x -= E_s qx_s/2,
//                                                      dEdcn_i -= E_s qi_s/2
//   G_s   = sum -/+ f_jk v_jk, Dq_s = sum dE q_jk  -> grad_x += G_s,
//                                                      dEdcn_x -= Dq_s/2
// so a triple costs one gather of the (j, k) pair record, one rsqrt, one pow
// and one reciprocal, and the global atomics happen once per slot and tile.
// kSmooth adds the switching function of a nonzero width3 (cutoff.f90:73-99,
// atm.f90:720-790).
template <typename R, bool kGrad, bool kSmooth>
__global__ __launch_bounds__(kBlock) void atm_kernel(DevBatch b, const long long* __restrict__ item_off,
                                                     const long long* __restrict__ offset,
                                                     const Slot<R>* __restrict__ slots,
                                                     const R4<R>* __restrict__ pair, double* __restrict__ e_atom,
                                                     double* __restrict__ dedcn, double* __restrict__ grad,
                                                     double* __restrict__ sig_atom) {
  __shared__ TileLds<R> lds[kWavesPerBlock];
  __shared__ double comb[kGrad ? kWavesPerBlock : 1][6][kTile];
  __shared__ double red[kWavesPerBlock];
  const long long item = blockIdx.x;
  if (item >= item_off[b.natoms]) return;  // whole block
  int lo = 0, hi = b.natoms;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (item_off[mid] <= item) lo = mid;
    else hi = mid;
  }
  const int i = lo;
  const int P = static_cast<int>(item - item_off[i]);
  const int s = b.sys[i];
  const SysGeom g = b.geom[s];
  const SysParam p = b.par[s];
  const int n = g.natoms;
  const long long L0 = offset[i];
  const int m = static_cast<int>(offset[i + 1] - L0);
  const int nb = (m + kTile - 1) / kTile;
  const double cut2 = p.cut3 * p.cut3;
  const R alp = R(p.alp) + R(2);
  const R malp6 = -alp / R(6);
  const R s9 = R(p.s9);
  const int w = threadIdx.x >> 6, l = wave_lane();
  TileLds<R>& T = lds[w];

  const int a = P * kTile + l;
  const bool aval = a < m;
  Slot<R> sa{};
  if (aval) sa = slots[L0 + a];
  const int jat = sa.atom;
  const R4<R>* prow = aval ? pair + g.c6_begin + static_cast<long long>(jat - g.atom_begin) * n - g.atom_begin
                           : pair;
  const R ar2 = sa.r2, arinv = sa.rinv, apdmp = sa.pdmp, asc6 = sa.sc6;
  const R air2 = arinv * arinv;
  const R cutr = R(p.cut3), width = R(p.width3);
  R swa = R(1), dswa = R(0);  // switching of r_ij and d sw / dr / r
  if (kSmooth && aval) {
    smooth_cutoff(R(1) / arinv, cutr, width, swa, dswa);
    dswa *= arinv;
  }
  double ei = 0.0;
  double Fa = 0.0, Ea = 0.0, Gax = 0.0, Gay = 0.0, Gaz = 0.0, Dqa = 0.0;
  double gix = 0.0, giy = 0.0, giz = 0.0, di = 0.0;
  double sxx = 0.0, syy = 0.0, szz = 0.0, sxy = 0.0, sxz = 0.0, syz = 0.0;

  for (int Q = P + w; Q < nb; Q += kWavesPerBlock) {
    const int bq = Q * kTile + l;
    Slot<R> sq{};
    sq.atom = -1;
    if (bq < m) sq = slots[L0 + bq];
    T.vx[l] = sq.vx; T.vy[l] = sq.vy; T.vz[l] = sq.vz;
    T.r2[l] = sq.r2; T.rinv[l] = sq.rinv; T.pdmp[l] = sq.pdmp; T.sc6[l] = sq.sc6;
    T.atom[l] = sq.atom;
    if (kGrad)
      for (int c = 0; c < 6; ++c) T.acc[c][l] = 0.0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const bool diag = Q == P;
    const int s0 = diag ? 1 : 0, s1 = diag ? kTile / 2 : kTile - 1;
    if (aval) {
      for (int st = s0; st <= s1; ++st) {
        const int bl = (l + st) & (kTile - 1);
        const int kat = T.atom[bl];
        if (kat < 0 || (diag && st == kTile / 2 && l >= kTile / 2)) continue;
        double vjxd, vjyd, vjzd;
        const double Cd = diff_norm2(T.vx[bl], T.vy[bl], T.vz[bl], sa.vx, sa.vy, sa.vz, vjxd, vjyd, vjzd);
        if (Cd > cut2 || Cd < kEps) continue;
        const R4<R> q = prow[kat];
        const R C = R(Cd);
        const R A = ar2, B = T.r2[bl];
        const R sc = d_rsqrt(C);
        const R pc = q.w * d_pow(C, malp6);
        const R ir1 = arinv * T.rinv[bl] * sc;
        const R ir2 = ir1 * ir1, ir3 = ir1 * ir2, ir5 = ir3 * ir2;
        const R t1 = A + C - B, t2 = A - C + B, t3 = -A + C + B;
        const R ang = R(0.375) * t1 * t2 * t3 * ir5 + ir3;
        const R pp = apdmp * T.pdmp[bl] * pc;
        const R fd = d_rcp(R(1) + R(6) * pp);
        const R c9 = -s9 * asc6 * T.sc6[bl] * q.x;
        const R wgt = R(triple_scale(i, jat, kat) * (jat == kat ? 2.0 : 1.0));
        const R cw = c9 * wgt;
        R sw = R(1), swb = R(1), dswb = R(0), swc = R(1), dswc = R(0);
        if (kSmooth) {
          smooth_cutoff(R(1) / T.rinv[bl], cutr, width, swb, dswb);
          dswb *= T.rinv[bl];
          smooth_cutoff(C * sc, cutr, width, swc, dswc);
          dswc *= sc;
          sw = swa * swb * swc;
        }
        const R dE = ang * fd * cw * sw;
        ei -= double(dE);
        if (kGrad) {
          const R dfd = R(-2) * alp * pp * fd * fd;
          const R A2 = A * A, B2 = B * B, C2 = C * C;
          const R dang_ij = R(-0.375) * (A2 * A + A2 * (C + B) + A * (R(3) * C2 + R(2) * C * B + R(3) * B2)
                                         - R(5) * (C - B) * (C - B) * (C + B)) * ir5;
          const R dang_ik = R(-0.375) * (B2 * B + B2 * (C + A) + B * (R(3) * C2 + R(2) * C * A + R(3) * A2)
                                         - R(5) * (C - A) * (C - A) * (C + A)) * ir5;
          const R dang_jk = R(-0.375) * (C2 * C + C2 * (B + A) + C * (R(3) * B2 + R(2) * B * A + R(3) * A2)
                                         - R(5) * (B - A) * (B - A) * (B + A)) * ir5;
          const R angd = ang * dfd;
          const R brinv = T.rinv[bl];
          R fij = cw * (angd - dang_ij * fd) * air2;
          R fik = cw * (angd - dang_ik * fd) * (brinv * brinv);
          R fjk = cw * (angd - dang_jk * fd) * (sc * sc);
          if (kSmooth) {  // atm.f90:754-790, switching-function terms
            const R e0 = ang * fd * cw;
            fij = sw * fij - e0 * dswa * swb * swc;
            fik = sw * fik - e0 * dswb * swa * swc;
            fjk = sw * fjk - e0 * dswc * swa * swb;
          }
          const R vjx = R(vjxd), vjy = R(vjyd), vjz = R(vjzd);
          const R gx = fjk * vjx, gy = fjk * vjy, gz = fjk * vjz;
          Fa += double(fij);
          Ea += double(dE);
          Gax -= double(gx); Gay -= double(gy); Gaz -= double(gz);
          Dqa += double(dE * q.y);
          atomicAdd(&T.acc[0][bl], double(fik));
          atomicAdd(&T.acc[1][bl], double(dE));
          atomicAdd(&T.acc[2][bl], double(gx));
          atomicAdd(&T.acc[3][bl], double(gy));
          atomicAdd(&T.acc[4][bl], double(gz));
          atomicAdd(&T.acc[5][bl], double(dE * q.z));
          sxx += double(gx * vjx); syy += double(gy * vjy); szz += double(gz * vjz);
          sxy += double(gx * vjy); sxz += double(gx * vjz); syz += double(gy * vjz);
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    if (kGrad && bq < m) {
      const double F = T.acc[0][l], E = T.acc[1][l];
      if (F != 0.0 || E != 0.0) {
        const int k = sq.atom;
        const double vx = sq.vx, vy = sq.vy, vz = sq.vz;
        atomicAdd(grad + 3 * k, F * vx + T.acc[2][l]);
        atomicAdd(grad + 3 * k + 1, F * vy + T.acc[3][l]);
        atomicAdd(grad + 3 * k + 2, F * vz + T.acc[4][l]);
        atomicAdd(dedcn + k, -0.5 * (E * double(sq.qx) + T.acc[5][l]));
        gix -= F * vx; giy -= F * vy; giz -= F * vz;
        di -= 0.5 * E * double(sq.qi);
        sxx += F * vx * vx; syy += F * vy * vy; szz += F * vz * vz;
        sxy += F * vx * vy; sxz += F * vx * vz; syz += F * vy * vz;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
  }
  if (kGrad) {
    comb[w][0][l] = Fa; comb[w][1][l] = Ea; comb[w][2][l] = Gax;
    comb[w][3][l] = Gay; comb[w][4][l] = Gaz; comb[w][5][l] = Dqa;
  }
  __syncthreads();
  if (kGrad && w == 0 && aval) {
    double F = 0, E = 0, Gx = 0, Gy = 0, Gz = 0, Dq = 0;
    for (int v = 0; v < kWavesPerBlock; ++v) {
      F += comb[v][0][l]; E += comb[v][1][l]; Gx += comb[v][2][l];
      Gy += comb[v][3][l]; Gz += comb[v][4][l]; Dq += comb[v][5][l];
    }
    if (F != 0.0 || E != 0.0) {
      const double vx = sa.vx, vy = sa.vy, vz = sa.vz;
      atomicAdd(grad + 3 * jat, F * vx + Gx);
      atomicAdd(grad + 3 * jat + 1, F * vy + Gy);
      atomicAdd(grad + 3 * jat + 2, F * vz + Gz);
      atomicAdd(dedcn + jat, -0.5 * (E * double(sa.qx) + Dq));
      gix -= F * vx; giy -= F * vy; giz -= F * vz;
      di -= 0.5 * E * double(sa.qi);
      sxx += F * vx * vx; syy += F * vy * vy; szz += F * vz * vz;
      sxy += F * vx * vy; sxz += F * vx * vz; syz += F * vy * vz;
    }
  }
  block_add(e_atom + i, ei, red);
  if (kGrad) {
    block_add(grad + 3 * i, gix, red);
    block_add(grad + 3 * i + 1, giy, red);
    block_add(grad + 3 * i + 2, giz, red);
    block_add(dedcn + i, di, red);
    block_add(sig_atom + 6 * i + 0, sxx, red);
    block_add(sig_atom + 6 * i + 1, syy, red);
    block_add(sig_atom + 6 * i + 2, szz, red);
    block_add(sig_atom + 6 * i + 3, sxy, red);
    block_add(sig_atom + 6 * i + 4, sxz, red);
    block_add(sig_atom + 6 * i + 5, syz, red);
  }
}

// ------------------------------------------------------ dCN/dR chain rule
// ncoord.f90:152-228 add_coordination_number_derivs, one wavefront per atom
// (recomputes the CN derivative, no 3 x N x N dCN/dR array).
template <typename R>
__global__ __launch_bounds__(kBlock) void cn_chain_kernel(DevBatch b, const double* __restrict__ dedcn,
                                                          double* __restrict__ grad, double* __restrict__ sig_atom) {
  const int lane = wave_lane();
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= b.natoms) return;
  const int s = b.sys[i];
  const SysGeom g = b.geom[s];
  const double cut = b.par[s].cut_cn;
  const double cut2 = cut * cut;
  const double4 xi = b.pos[i];
  const R rci = R(b.rcov[b.sp[i]]);
  const double dei = dedcn[i];
  double g0 = 0.0, g1 = 0.0, g2 = 0.0;
  double sxx = 0.0, syy = 0.0, szz = 0.0, sxy = 0.0, sxz = 0.0, syz = 0.0;
  for (int j = g.atom_begin + lane; j < g.atom_begin + g.natoms; j += kWave) {
    const double4 xj = b.pos[j];
    const R rc = rci + R(b.rcov[b.sp[j]]);
    const bool swap = j > i;
    // the reference adds ds*(dEdcn_i + dEdcn_j) once per unordered pair and
    // ds*dEdcn_i once per self-image (ncoord.f90:214-216)
    const R de = R(dei + dedcn[j]);
    const R wsig = (j == i) ? R(dei) : R(0.5) * de;
    for (int t = 0; t < g.img_cn_count; ++t) {
      double u0, u1, u2;
      const double r2d = pair_vec(xi, xj, b.img[g.img_cn_begin + t], swap, u0, u1, u2);
      if (r2d > cut2 || r2d < kCnFloor) continue;
      const R r2 = R(r2d);
      const R r1 = d_sqrt(r2);
      const R expterm = d_exp(-R(kKcn) * (rc / r1 - R(1)));
      const R cf = d_rcp(R(1) + expterm);
      const R dcf = -cf * (R(1) - cf) * R(kKcn) * rc / r2;
      const R f = dcf / r1;  // countd = f * vec
      const R v0 = R(u0), v1 = R(u1), v2 = R(u2);
      if (j != i) {
        g0 += double(f * v0 * de); g1 += double(f * v1 * de); g2 += double(f * v2 * de);
      }
      const R h = f * wsig;
      sxx += double(h * v0 * v0); syy += double(h * v1 * v1); szz += double(h * v2 * v2);
      sxy += double(h * v0 * v1); sxz += double(h * v0 * v2); syz += double(h * v1 * v2);
    }
  }
  g0 = wave_sum(g0); g1 = wave_sum(g1); g2 = wave_sum(g2);
  sxx = wave_sum(sxx); syy = wave_sum(syy); szz = wave_sum(szz);
  sxy = wave_sum(sxy); sxz = wave_sum(sxz); syz = wave_sum(syz);
  if (lane == 0) {
    grad[3 * i] += g0; grad[3 * i + 1] += g1; grad[3 * i + 2] += g2;
    double* sg = sig_atom + 6 * i;
    sg[0] += sxx; sg[1] += syy; sg[2] += szz; sg[3] += sxy; sg[4] += sxz; sg[5] += syz;
  }
}

// --------------------------------------------------------- final reduction
// disp.f90:397 energy = sum(energies) and the virial sum, per system.
__global__ __launch_bounds__(kBlock) void reduce_kernel(DevBatch b, const double* __restrict__ e_atom,
                                                        const double* __restrict__ sig_atom,
                                                        double* __restrict__ energy, double* __restrict__ sigma) {
  __shared__ double red[7][kWavesPerBlock];
  const int s = blockIdx.x;
  const SysGeom g = b.geom[s];
  double v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = g.atom_begin + threadIdx.x; i < g.atom_begin + g.natoms; i += blockDim.x) {
    v[0] += e_atom[i];
    if (sigma)
      for (int c = 0; c < 6; ++c) v[1 + c] += sig_atom[6 * i + c];
  }
  for (int c = 0; c < 7; ++c) {
    double x = wave_sum(v[c]);
    if (wave_lane() == 0) red[c][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[7];
    for (int c = 0; c < 7; ++c) {
      t[c] = 0.0;
      for (int w = 0; w < kWavesPerBlock; ++w) t[c] += red[c][w];
    }
    energy[s] = t[0];
    if (sigma) {
      double* sg = sigma + 9 * s;  // sigma(a,b) at [b*3+a], symmetric
      sg[0] = t[1]; sg[4] = t[2]; sg[8] = t[3];
      sg[1] = sg[3] = t[4];
      sg[2] = sg[6] = t[5];
      sg[5] = sg[7] = t[6];
    }
  }
}

// ------------------------------------------------------------ term counts
// Work units as the reference's loops visit them: CN terms (iat, jat <= iat,
// T) with 1e-12 <= r2 <= cn^2 (ncoord.f90:118-127) and two-body terms with
// eps <= r2 <= disp2^2 (rational.f90:450-464), counted per system.
__global__ __launch_bounds__(kBlock) void count_pairs_kernel(DevBatch b, unsigned long long* __restrict__ counts) {
  const int lane = wave_lane();
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= b.natoms) return;
  const int s = b.sys[i];
  const SysGeom g = b.geom[s];
  const SysParam p = b.par[s];
  const double cn2 = p.cut_cn * p.cut_cn, c22 = p.cut2 * p.cut2;
  const double4 xi = b.pos[i];
  unsigned long long ncn = 0, n2 = 0;
  for (int j = g.atom_begin + lane; j <= i; j += kWave) {
    const double4 xj = b.pos[j];
    double u0, u1, u2;
    for (int t = 0; t < g.img_cn_count; ++t) {
      const double r2 = pair_vec(xi, xj, b.img[g.img_cn_begin + t], false, u0, u1, u2);
      ncn += (r2 > cn2 || r2 < kCnFloor) ? 0 : 1;
    }
    for (int t = 0; t < g.img2_count; ++t) {
      const double r2 = pair_vec(xi, xj, b.img[g.img2_begin + t], false, u0, u1, u2);
      n2 += (r2 > c22 || r2 < kEps) ? 0 : 1;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    ncn += __shfl_xor(ncn, off, kWave);
    n2 += __shfl_xor(n2, off, kWave);
  }
  if (lane == 0) {
    atomicAdd(counts + 3 * s, ncn);
    atomicAdd(counts + 3 * s + 1, n2);
  }
}

// ATM triple terms of the reference (atm.f90:706-742): unordered slot pairs
// of L(i) within the cutoff, two for a pair of images of one atom.
template <typename R>
__global__ __launch_bounds__(kBlock) void count_triples_kernel(DevBatch b, const long long* __restrict__ offset,
                                                               const Slot<R>* __restrict__ slots,
                                                               unsigned long long* __restrict__ counts) {
  __shared__ unsigned long long red[kWavesPerBlock];
  const int i = blockIdx.x;
  const int s = b.sys[i];
  const double cut2 = b.par[s].cut3 * b.par[s].cut3;
  const long long L0 = offset[i], L1 = offset[i + 1];
  unsigned long long n = 0;
  for (long long a = L0 + threadIdx.x; a < L1; a += blockDim.x) {
    const Slot<R> sa = slots[a];
    for (long long bb = a + 1; bb < L1; ++bb) {
      double v0, v1, v2;
      const Slot<R>& sb = slots[bb];
      const double r2 = diff_norm2(sb.vx, sb.vy, sb.vz, sa.vx, sa.vy, sa.vz, v0, v1, v2);
      if (r2 > cut2 || r2 < kEps) continue;
      n += sb.atom == sa.atom ? 2 : 1;
    }
  }
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, kWave);
  if (wave_lane() == 0) red[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kWavesPerBlock; ++w) t += red[w];
    if (t) atomicAdd(counts + 3 * s + 2, t);
  }
}

}  // namespace

// ---------------------------------------------------------------- launchers
size_t pair_record_bytes(bool mixed) { return mixed ? sizeof(float4) : sizeof(double4); }
size_t slot_record_bytes(bool mixed) { return mixed ? sizeof(Slot<float>) : sizeof(Slot<double>); }

void launch_pack(const double* xyz, int n, double4* pos, hipStream_t st) {
  hipLaunchKernelGGL(pack_positions, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, xyz, n, pos);
}

void launch_cn(const DevBatch& b, bool mixed, double* cn, hipStream_t st) {
  const dim3 grid(blocks_for(b.natoms, kWavesPerBlock)), blk(kBlock);
  if (mixed) hipLaunchKernelGGL(cn_kernel<float>, grid, blk, 0, st, b, cn);
  else hipLaunchKernelGGL(cn_kernel<double>, grid, blk, 0, st, b, cn);
}

void launch_weights(const DevBatch& b, const double* cn, double* gw, double* dgw, hipStream_t st) {
  hipLaunchKernelGGL(weights_kernel, dim3(blocks_for(b.natoms, kBlock)), dim3(kBlock), 0, st, b, cn, gw, dgw);
}

void launch_pair(const DevBatch& b, bool mixed, const double* gw, const double* dgw, void* pair, hipStream_t st) {
  if (mixed)
    hipLaunchKernelGGL(pair_kernel<float>, dim3(b.natoms), dim3(kBlock), 0, st, b, gw, dgw,
                       static_cast<float4*>(pair));
  else
    hipLaunchKernelGGL(pair_kernel<double>, dim3(b.natoms), dim3(kBlock), 0, st, b, gw, dgw,
                       static_cast<double4*>(pair));
}

void launch_disp2(const DevBatch& b, bool mixed, bool grad, const DevWork& w, hipStream_t st) {
  const dim3 grid(blocks_for(b.natoms, kWavesPerBlock)), blk(kBlock);
  const float4* pf = static_cast<const float4*>(w.pair);
  const double4* pd = static_cast<const double4*>(w.pair);
  if (mixed && grad)
    hipLaunchKernelGGL((disp2_kernel<float, true>), grid, blk, 0, st, b, pf, w.e_atom, w.dedcn, w.grad, w.sig_atom);
  else if (mixed)
    hipLaunchKernelGGL((disp2_kernel<float, false>), grid, blk, 0, st, b, pf, w.e_atom, w.dedcn, w.grad, w.sig_atom);
  else if (grad)
    hipLaunchKernelGGL((disp2_kernel<double, true>), grid, blk, 0, st, b, pd, w.e_atom, w.dedcn, w.grad, w.sig_atom);
  else
    hipLaunchKernelGGL((disp2_kernel<double, false>), grid, blk, 0, st, b, pd, w.e_atom, w.dedcn, w.grad,
                       w.sig_atom);
}

void launch_atm_count(const DevBatch& b, const void* pair, bool mixed, long long* count, hipStream_t st) {
  const dim3 grid(blocks_for(b.natoms, kWavesPerBlock)), blk(kBlock);
  if (mixed)
    hipLaunchKernelGGL((atm_list_kernel<false, float>), grid, blk, 0, st, b, static_cast<const float4*>(pair), count,
                       (const long long*)nullptr, (Slot<float>*)nullptr);
  else
    hipLaunchKernelGGL((atm_list_kernel<false, double>), grid, blk, 0, st, b, static_cast<const double4*>(pair),
                       count, (const long long*)nullptr, (Slot<double>*)nullptr);
}

void launch_scan(const long long* count, int n, long long* offset, hipStream_t st) {
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, count, n, offset);
}

void launch_atm_fill(const DevBatch& b, const void* pair, bool mixed, const long long* offset, void* slots,
                     hipStream_t st) {
  const dim3 grid(blocks_for(b.natoms, kWavesPerBlock)), blk(kBlock);
  if (mixed)
    hipLaunchKernelGGL((atm_list_kernel<true, float>), grid, blk, 0, st, b, static_cast<const float4*>(pair),
                       (long long*)nullptr, offset, static_cast<Slot<float>*>(slots));
  else
    hipLaunchKernelGGL((atm_list_kernel<true, double>), grid, blk, 0, st, b, static_cast<const double4*>(pair),
                       (long long*)nullptr, offset, static_cast<Slot<double>*>(slots));
}

void launch_atm_items(const long long* offset, int n, long long* item_count, hipStream_t st) {
  hipLaunchKernelGGL(atm_items_kernel, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, offset, n, item_count);
}

namespace {
template <typename R, bool kGrad, bool kSmooth>
void launch_atm_t(const DevBatch& b, const long long* item_off, long long max_items, const long long* offset,
                  const void* slots, const DevWork& w, hipStream_t st) {
  hipLaunchKernelGGL((atm_kernel<R, kGrad, kSmooth>), dim3(static_cast<unsigned>(max_items)), dim3(kBlock), 0, st,
                     b, item_off, offset, static_cast<const Slot<R>*>(slots), static_cast<const R4<R>*>(w.pair),
                     w.e_atom, w.dedcn, w.grad, w.sig_atom);
}
}  // namespace

void launch_atm(const DevBatch& b, bool mixed, bool grad, bool smooth, const long long* item_off,
                long long max_items, const long long* offset, const void* slots, const DevWork& w, hipStream_t st) {
  if (max_items <= 0) return;
  using L = void (*)(const DevBatch&, const long long*, long long, const long long*, const void*, const DevWork&,
                     hipStream_t);
  static const L table[8] = {
      launch_atm_t<double, false, false>, launch_atm_t<double, false, true>, launch_atm_t<double, true, false>,
      launch_atm_t<double, true, true>,   launch_atm_t<float, false, false>, launch_atm_t<float, false, true>,
      launch_atm_t<float, true, false>,   launch_atm_t<float, true, true>};
  table[(mixed ? 4 : 0) + (grad ? 2 : 0) + (smooth ? 1 : 0)](b, item_off, max_items, offset, slots, w, st);
}

void launch_cn_chain(const DevBatch& b, bool mixed, const DevWork& w, hipStream_t st) {
  const dim3 grid(blocks_for(b.natoms, kWavesPerBlock)), blk(kBlock);
  if (mixed) hipLaunchKernelGGL(cn_chain_kernel<float>, grid, blk, 0, st, b, w.dedcn, w.grad, w.sig_atom);
  else hipLaunchKernelGGL(cn_chain_kernel<double>, grid, blk, 0, st, b, w.dedcn, w.grad, w.sig_atom);
}

void launch_reduce(const DevBatch& b, const DevWork& w, double* energy, double* sigma, hipStream_t st) {
  hipLaunchKernelGGL(reduce_kernel, dim3(b.nsys), dim3(kBlock), 0, st, b, w.e_atom, w.sig_atom, energy, sigma);
}

void launch_count_pairs(const DevBatch& b, unsigned long long* counts, hipStream_t st) {
  hipLaunchKernelGGL(count_pairs_kernel, dim3(blocks_for(b.natoms, kWavesPerBlock)), dim3(kBlock), 0, st, b, counts);
}

void launch_count_triples(const DevBatch& b, bool mixed, const long long* offset, const void* slots,
                          unsigned long long* counts, hipStream_t st) {
  if (mixed)
    hipLaunchKernelGGL(count_triples_kernel<float>, dim3(b.natoms), dim3(kBlock), 0, st, b, offset,
                       static_cast<const Slot<float>*>(slots), counts);
  else
    hipLaunchKernelGGL(count_triples_kernel<double>, dim3(b.natoms), dim3(kBlock), 0, st, b, offset,
                       static_cast<const Slot<double>*>(slots), counts);
}

}  // namespace d3
