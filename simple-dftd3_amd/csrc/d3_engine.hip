// HIP kernels and batch engine of the D3 energy / gradient / virial path.
//
// Reference: simple-dftd3 v1.6.0, src/dftd3/disp.f90:56-185 and the loops it
// calls.  Every kernel evaluates the same set of pair / triple terms as the
// reference (same lattice images, same pair-vector arithmetic for the cutoff
// tests, same triangular triple index set); only the order of summation
// differs.  Arithmetic of a term runs in the template type R (double for the
// plain path, float for the mixed path); energies, gradients, virials and
// dE/dCN are accumulated in double in both paths, and the cutoff tests always
// use the double-precision squared distance so that both paths visit exactly
// the reference's terms.
//
// Data layout (device memory of one batch; systems are consecutive):
//   pos4[N]      double4 (x, y, z, 0) bohr
//   sp[N], sys[N] species index and system index of every atom
//   img[]        double4 lattice translations, per system and cutoff
//   c6[blk + il*n + jl]   C6_ij         (symmetric n x n block per system)
//   dc6[...]              dC6_ij/dCN_i  (model.f90:432-462 dc6dcn(i,j))
//   dc6t[...]             dC6_ij/dCN_j  (= dc6dcn(j,i), transposed copy)
//   gw/dgw[N][mref]       Gaussian weights and dweights/dCN
//   e_atom[N], dedcn[N], sig_atom[N][6] per-atom accumulators
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "d3_engine.hpp"

namespace d3 {

void hip_check(hipError_t err, const char* what) {
  if (err != hipSuccess) throw Error(std::string("HIP error in ") + what + ": " + hipGetErrorString(err));
}

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr double kEps = 2.220446049250313e-16;  // epsilon(1.0_wp)
constexpr double kKcn = 16.0;                   // ncoord.f90:31
constexpr double kCnFloor = 1.0e-12;            // ncoord.f90:127

struct DevBatch {
  int natoms, nsys, nsp, mref;
  const double4* pos;
  const int* sp;
  const int* sys;
  const SysGeom* geom;
  const SysParam* par;
  const double4* img;
  const double* rcov;   // [nsp]
  const double* r4r2;   // [nsp]
  const int* nref;      // [nsp]
  const double* refcn;  // [nsp][kMaxRef]
  const double* c6ref;  // [nsp*nsp][kMaxRef*kMaxRef], [(si*nsp+sj)][ri*7+rj]
  const double* rvdw;   // [nsp*nsp]
};

struct AtmEntry {  // one (j, image) of the three-body neighbour list of atom i
  double vx, vy, vz;  // x_j + T - x_i (atm.f90:711)
  int j;              // partner atom (global index), j <= i
  int hi;             // end (local slot) of the group of atom j in the list
};
static_assert(sizeof(AtmEntry) == 32, "AtmEntry layout");

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ int wave_lane() { return threadIdx.x & (kWave - 1); }

// Pair vector of the reference tuple for the ordered pair (i, j) and image T
// of the calling thread.  For j <= i the reference visits (iat=i, jat=j, T)
// with vec = x_i - (x_j + T) (rational.f90:462, ncoord.f90:124); for j > i it
// visits (iat=j, jat=i, -T) with vec = x_j - (x_i - T).  Returns r2 exactly
// as the reference evaluates it (no contraction) and u, the vector from the
// image of j to i.
__device__ __forceinline__ double pair_vec(const double4& xi, const double4& xj, const double4& t,
                                           bool swap, double& u0, double& u1, double& u2) {
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

__device__ __forceinline__ double norm2_exact(double a, double b, double c) {
#pragma clang fp contract(off)
  return a * a + b * b + c * c;
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
template <typename R> __device__ __forceinline__ R d_rcp(R x);
template <> __device__ __forceinline__ double d_rcp<double>(double x) { return 1.0 / x; }
template <> __device__ __forceinline__ float d_rcp<float>(float x) { return __frcp_rn(x); }
// x**y for x > 0
template <typename R> __device__ __forceinline__ R d_pow(R x, R y);
template <> __device__ __forceinline__ double d_pow<double>(double x, double y) { return pow(x, y); }
template <> __device__ __forceinline__ float d_pow<float>(float x, float y) { return exp2f(y * __log2f(x)); }

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
      const double4 tv = b.img[g.img_cn_begin + t];
      double u0, u1, u2;
      const double r2 = pair_vec(xi, xj, tv, swap, u0, u1, u2);
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

// -------------------------------------------------------------------- C6
// model.f90:432-462: C6_ij = sum_kl w_ik w_jl C6ref_kl and its derivatives,
// one block per row i of the system's n x n block.
__global__ __launch_bounds__(kBlock) void c6_kernel(DevBatch b, const double* __restrict__ gw,
                                                    const double* __restrict__ dgw,
                                                    const int* __restrict__ ghost,
                                                    double* __restrict__ c6, double* __restrict__ dc6,
                                                    double* __restrict__ dc6t) {
  const int i = blockIdx.x;
  const int s = b.sys[i];
  const SysGeom g = b.geom[s];
  const int n = g.natoms, il = i - g.atom_begin;
  const int spi = b.sp[i];
  const int mref = b.mref;
  const int nri = b.nref[spi];
  double gi[kMaxRef], dgi[kMaxRef];
  for (int r = 0; r < kMaxRef; ++r) {
    gi[r] = r < nri ? gw[i * mref + r] : 0.0;
    dgi[r] = (r < nri && dgw) ? dgw[i * mref + r] : 0.0;
  }
  const bool gi_ghost = ghost && ghost[i];
  double* crow = c6 + g.c6_begin + static_cast<long long>(il) * n;
  for (int jl = threadIdx.x; jl < n; jl += blockDim.x) {
    const int j = g.atom_begin + jl;
    const int spj = b.sp[j];
    const int nrj = b.nref[spj];
    const double* ref = b.c6ref + (spi * b.nsp + spj) * (kMaxRef * kMaxRef);
    double cc = 0.0, di = 0.0, dj = 0.0;
    if (!(gi_ghost || (ghost && ghost[j]))) {
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
    crow[jl] = cc;
    if (dgw) {
      dc6[g.c6_begin + static_cast<long long>(il) * n + jl] = di;
      dc6t[g.c6_begin + static_cast<long long>(il) * n + jl] = dj;
    }
  }
}

// --------------------------------------------------------------- two-body
// rational.f90:378-514 (BJ) and zero.f90:409-553 (zero damping), E+grad+sigma;
// rational.f90:282-374 / zero.f90:306-405 when kGrad is false.  One wavefront
// per atom i; atom i collects the ordered pairs (i, j, T) it takes part in.
template <typename R, bool kGrad>
__global__ __launch_bounds__(kBlock) void disp2_kernel(DevBatch b, const double* __restrict__ c6,
                                                       const double* __restrict__ dc6,
                                                       double* __restrict__ e_atom,
                                                       double* __restrict__ dedcn,
                                                       double* __restrict__ grad,
                                                       double* __restrict__ sig_atom) {
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
  if (active) {
    for (int jl = lane; jl < n; jl += kWave) {
      const int j = g.atom_begin + jl;
      const double4 xj = b.pos[j];
      const int spj = b.sp[j];
      const R rrij = R(3) * r4r2i * R(b.r4r2[spj]);
      const R c6ij = R(c6[g.c6_begin + static_cast<long long>(il) * n + jl]);
      const R dc6i = kGrad ? R(dc6[g.c6_begin + static_cast<long long>(il) * n + jl]) : R(0);
      R r0a = R(0), r0b = R(0);  // BJ: r0^6, r0^8; zero: rs6*R0, rs8*R0
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
        const double4 tv = b.img[g.img2_begin + t];
        double u0, u1, u2;
        const double r2d = pair_vec(xi, xj, tv, swap, u0, u1, u2);
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
// Neighbour list L(i) = {(j, T): j <= i, eps <= |x_j + T - x_i|^2 <= cut3^2}
// (atm.f90:706-716), sorted by j then image.  The triples of atom i are the
// pairs (a, b) of L(i) with atom(b) <= atom(a) and eps <= |v_b - v_a|^2 <=
// cut3^2 (atm.f90:717-742), i.e. exactly the reference's index set.
template <bool kFill>
__global__ __launch_bounds__(kBlock) void atm_list_kernel(DevBatch b, long long* __restrict__ count,
                                                          const long long* __restrict__ offset,
                                                          AtmEntry* __restrict__ list) {
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
  const double4 xi = b.pos[i];
  AtmEntry* out = kFill ? list + offset[i] : nullptr;
  long long total = 0;
  for (int j0 = g.atom_begin; j0 <= i; j0 += kWave) {
    const int j = j0 + lane;
    int cnt = 0;
    const bool valid = j <= i;
    double4 xj = valid ? b.pos[j] : xi;
    if (valid) {
      for (int t = 0; t < g.img3_count; ++t) {
        double v0, v1, v2;
        const double r2 = atm_vec(xj, b.img[g.img3_begin + t], xi, v0, v1, v2);
        if (r2 > cut2 || r2 < kEps) continue;
        ++cnt;
      }
    }
    // exclusive scan of the lane counts
    int incl = cnt;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      int y = __shfl_up(incl, off, kWave);
      if (lane >= off) incl += y;
    }
    const int excl = incl - cnt;
    const int wtot = __shfl(incl, kWave - 1, kWave);
    if (kFill && valid && cnt > 0) {
      long long w = total + excl;
      const int hi = static_cast<int>(total + incl);
      for (int t = 0; t < g.img3_count; ++t) {
        double v0, v1, v2;
        const double r2 = atm_vec(xj, b.img[g.img3_begin + t], xi, v0, v1, v2);
        if (r2 > cut2 || r2 < kEps) continue;
        AtmEntry e;
        e.vx = v0; e.vy = v1; e.vz = v2; e.j = j; e.hi = hi;
        out[w++] = e;
      }
    }
    total += wtot;
  }
  if (!kFill && lane == 0) count[i] = total;
}

// Exclusive scan of n counts into offset[0..n] (single block, n moderate).
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
  }
  __syncthreads();
  long long run = part[tid];
  for (int k = lo; k < hi; ++k) {
    offset[k] = run;
    run += count[k];
  }
  if (hi == n && lo < hi) offset[n] = run;
  if (n == 0 && tid == 0) offset[0] = 0;
}

// vjk = vik - vij (atm.f90:723) and its squared length, as the reference.
__device__ __forceinline__ double diff_norm2(const AtmEntry& b, const AtmEntry& a, double& v0, double& v1,
                                             double& v2) {
#pragma clang fp contract(off)
  v0 = b.vx - a.vx;
  v1 = b.vy - a.vy;
  v2 = b.vz - a.vz;
  return v0 * v0 + v1 * v1 + v2 * v2;
}

// atm.f90:954-980
__device__ __forceinline__ double triple_scale(int ii, int jj, int kk) {
  if (ii == jj) return ii == kk ? 1.0 / 6.0 : 0.5;
  if (ii != kk && jj != kk) return 1.0;
  return 0.5;
}

__device__ __forceinline__ void block_add(double* __restrict__ dst, double v, double* red) {
  // sum v over the block and atomically add the total to *dst (thread 0)
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if (wave_lane() == 0) red[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += red[k];
    atomicAdd(dst, t);
  }
  __syncthreads();
}

// atm.f90:625-825 get_atm_dispersion_derivs (kGrad) and atm.f90:500-621,
// with rs9 = 4/3 and alp + 2 (rational.f90:563, zero.f90:602).  One block per
// atom i; each thread owns slots a = (j, Tj) of L(i) and loops over the
// slots b = (k, Tk) with k <= j.  Contributions to i and j stay in
// registers, contributions to k are added atomically.  The triple energy is
// booked on atom i (the atomic split of atm.f90:794-796 does not change the
// total).
template <typename R, bool kGrad>
__global__ __launch_bounds__(kBlock) void atm_kernel(DevBatch b, const long long* __restrict__ offset,
                                                     const AtmEntry* __restrict__ list,
                                                     const double* __restrict__ c6,
                                                     const double* __restrict__ dc6,
                                                     const double* __restrict__ dc6t,
                                                     double* __restrict__ e_atom,
                                                     double* __restrict__ dedcn,
                                                     double* __restrict__ grad,
                                                     double* __restrict__ sig_atom) {
  __shared__ double red[kWavesPerBlock];
  const int i = blockIdx.x;
  const int s = b.sys[i];
  const SysParam p = b.par[s];
  if (fabs(p.s9) < kEps) return;
  const SysGeom g = b.geom[s];
  const int n = g.natoms, il = i - g.atom_begin;
  const long long blk = g.c6_begin;
  const long long L0 = offset[i], L1 = offset[i + 1];
  const double cut2 = p.cut3 * p.cut3;
  const R cutr = R(p.cut3), width = R(p.width3);
  const R alp = R(p.alp) + R(2), alp3 = alp / R(3);
  const R rs9 = R(4) / R(3);
  const R s9 = R(p.s9);
  const int spi = b.sp[i];
  const double* c6row_i = c6 + blk + static_cast<long long>(il) * n;
  const double* drow_i = dc6 + blk + static_cast<long long>(il) * n;    // dC6_ik/dCN_i
  const double* dtrow_i = dc6t + blk + static_cast<long long>(il) * n;  // dC6_ik/dCN_k
  double ei = 0.0, di = 0.0, gi0 = 0.0, gi1 = 0.0, gi2 = 0.0;
  double sxx = 0.0, syy = 0.0, szz = 0.0, sxy = 0.0, sxz = 0.0, syz = 0.0;
  for (long long a = L0 + threadIdx.x; a < L1; a += blockDim.x) {
    const AtmEntry ea = list[a];
    const int j = ea.j, jl = j - g.atom_begin;
    const int spj = b.sp[j];
    const R vij0 = R(ea.vx), vij1 = R(ea.vy), vij2 = R(ea.vz);
    const double r2ijd = norm2_exact(ea.vx, ea.vy, ea.vz);
    const R r2ij = R(r2ijd);
    const R rij = d_sqrt(r2ij);
    R swij, dswijdr;
    smooth_cutoff(rij, cutr, width, swij, dswijdr);
    const double c6ijd = c6row_i[jl];
    const R c6ij = R(c6ijd);
    const R r0ij = rs9 * R(b.rvdw[spj * b.nsp + spi]);
    R qij_i = R(0), qij_j = R(0);
    if (kGrad) {
      qij_i = R(drow_i[jl]) / c6ij;   // dc6dcn(iat,jat)/c6ij
      qij_j = R(dtrow_i[jl]) / c6ij;  // dc6dcn(jat,iat)/c6ij
    }
    double dj = 0.0, gj0 = 0.0, gj1 = 0.0, gj2 = 0.0;
    const long long bend = L0 + ea.hi;
    for (long long bb = L0; bb < bend; ++bb) {
      const AtmEntry eb = list[bb];
      double vjk0d, vjk1d, vjk2d;
      const double r2jkd = diff_norm2(eb, ea, vjk0d, vjk1d, vjk2d);
      if (r2jkd > cut2 || r2jkd < kEps) continue;
      const int k = eb.j, kl = k - g.atom_begin;
      const long long kj = blk + static_cast<long long>(kl) * n + jl;  // row k, column j
      const double c6ikd = c6row_i[kl];
      const double c6jkd = c6[kj];
      if (kGrad && (fabs(c6ijd) < kEps || fabs(c6ikd) < kEps || fabs(c6jkd) < kEps)) continue;
      const R c6ik = R(c6ikd), c6jk = R(c6jkd);
      const int spk = b.sp[k];
      const R c9 = -s9 * d_sqrt(fabs(c6ij * c6ik * c6jk));
      const R r0ik = rs9 * R(b.rvdw[spk * b.nsp + spi]);
      const R r0jk = rs9 * R(b.rvdw[spk * b.nsp + spj]);
      const R r0 = r0ij * r0ik * r0jk;
      const R triple = R(triple_scale(i, j, k));
      const R vik0 = R(eb.vx), vik1 = R(eb.vy), vik2 = R(eb.vz);
      const R vjk0 = R(vjk0d), vjk1 = R(vjk1d), vjk2 = R(vjk2d);
      const R r2ik = R(norm2_exact(eb.vx, eb.vy, eb.vz));
      const R r2jk = R(r2jkd);
      const R rik = d_sqrt(r2ik), rjk = d_sqrt(r2jk);
      R swik, dswikdr, swjk, dswjkdr;
      smooth_cutoff(rik, cutr, width, swik, dswikdr);
      smooth_cutoff(rjk, cutr, width, swjk, dswjkdr);
      const R sw = swij * swik * swjk;
      const R r2 = r2ij * r2ik * r2jk;
      const R r1 = d_sqrt(r2);
      const R r3 = r2 * r1;
      const R r5 = r3 * r2;
      const R ir5 = d_rcp(r5);
      const R r0r1alp3 = d_pow(r0 / r1, alp3);
      const R fdmp = d_rcp(R(1) + R(6) * r0r1alp3);
      const R ang = R(0.375) * (r2ij + r2jk - r2ik) * (r2ij - r2jk + r2ik) * (-r2ij + r2jk + r2ik) * ir5
                  + d_rcp(r3);
      const R rr = ang * fdmp;
      if (!kGrad) {
        ei -= double(rr * c9 * triple * sw);
        continue;
      }
      const R dfdmp = R(-2) * alp * r0r1alp3 * fdmp * fdmp;
      const R dE0 = rr * c9;
      R dang;
      dang = R(-0.375) * (r2ij * r2ij * r2ij + r2ij * r2ij * (r2jk + r2ik)
                          + r2ij * (R(3) * r2jk * r2jk + R(2) * r2jk * r2ik + R(3) * r2ik * r2ik)
                          - R(5) * (r2jk - r2ik) * (r2jk - r2ik) * (r2jk + r2ik)) * ir5;
      const R fij = sw * c9 * (-dang * fdmp + ang * dfdmp) / r2ij - dE0 * dswijdr / rij * swik * swjk;
      dang = R(-0.375) * (r2ik * r2ik * r2ik + r2ik * r2ik * (r2jk + r2ij)
                          + r2ik * (R(3) * r2jk * r2jk + R(2) * r2jk * r2ij + R(3) * r2ij * r2ij)
                          - R(5) * (r2jk - r2ij) * (r2jk - r2ij) * (r2jk + r2ij)) * ir5;
      const R fik = sw * c9 * (-dang * fdmp + ang * dfdmp) / r2ik - dE0 * dswikdr / rik * swij * swjk;
      dang = R(-0.375) * (r2jk * r2jk * r2jk + r2jk * r2jk * (r2ik + r2ij)
                          + r2jk * (R(3) * r2ik * r2ik + R(2) * r2ik * r2ij + R(3) * r2ij * r2ij)
                          - R(5) * (r2ik - r2ij) * (r2ik - r2ij) * (r2ik + r2ij)) * ir5;
      const R fjk = sw * c9 * (-dang * fdmp + ang * dfdmp) / r2jk - dE0 * dswjkdr / rjk * swij * swik;
      const R dE = dE0 * triple * sw;
      ei -= double(dE);
      // gradients (atm.f90:798-800), dG = f * v
      const R gij0 = fij * vij0, gij1 = fij * vij1, gij2 = fij * vij2;
      const R gik0 = fik * vik0, gik1 = fik * vik1, gik2 = fik * vik2;
      const R gjk0 = fjk * vjk0, gjk1 = fjk * vjk1, gjk2 = fjk * vjk2;
      gi0 -= double((gij0 + gik0) * triple);
      gi1 -= double((gij1 + gik1) * triple);
      gi2 -= double((gij2 + gik2) * triple);
      gj0 += double((gij0 - gjk0) * triple);
      gj1 += double((gij1 - gjk1) * triple);
      gj2 += double((gij2 - gjk2) * triple);
      atomicAdd(grad + 3 * k, double((gik0 + gjk0) * triple));
      atomicAdd(grad + 3 * k + 1, double((gik1 + gjk1) * triple));
      atomicAdd(grad + 3 * k + 2, double((gik2 + gjk2) * triple));
      // virial (atm.f90:802-809), symmetric part
      const R fij_t = fij * triple, fik_t = fik * triple, fjk_t = fjk * triple;
      sxx += double(fij_t * vij0 * vij0 + fik_t * vik0 * vik0 + fjk_t * vjk0 * vjk0);
      syy += double(fij_t * vij1 * vij1 + fik_t * vik1 * vik1 + fjk_t * vjk1 * vjk1);
      szz += double(fij_t * vij2 * vij2 + fik_t * vik2 * vik2 + fjk_t * vjk2 * vjk2);
      sxy += double(fij_t * vij0 * vij1 + fik_t * vik0 * vik1 + fjk_t * vjk0 * vjk1);
      sxz += double(fij_t * vij0 * vij2 + fik_t * vik0 * vik2 + fjk_t * vjk0 * vjk2);
      syz += double(fij_t * vij1 * vij2 + fik_t * vik1 * vik2 + fjk_t * vjk1 * vjk2);
      // dE/dCN (atm.f90:811-816)
      const R ic6ik = d_rcp(c6ik), ic6jk = d_rcp(c6jk);
      const R hde = dE * R(0.5);
      di -= double(hde * (qij_i + R(drow_i[kl]) * ic6ik));
      dj -= double(hde * (qij_j + R(dc6t[kj]) * ic6jk));   // dc6dcn(jat,kat)/c6jk
      atomicAdd(dedcn + k, -double(hde * (R(dtrow_i[kl]) * ic6ik + R(dc6[kj]) * ic6jk)));
    }
    if (kGrad) {
      atomicAdd(dedcn + j, dj);
      atomicAdd(grad + 3 * j, gj0);
      atomicAdd(grad + 3 * j + 1, gj1);
      atomicAdd(grad + 3 * j + 2, gj2);
    }
  }
  block_add(e_atom + i, ei, red);
  if (kGrad) {
    block_add(dedcn + i, di, red);
    block_add(grad + 3 * i, gi0, red);
    block_add(grad + 3 * i + 1, gi1, red);
    block_add(grad + 3 * i + 2, gi2, red);
    block_add(sig_atom + 6 * i + 0, sxx, red);
    block_add(sig_atom + 6 * i + 1, syy, red);
    block_add(sig_atom + 6 * i + 2, szz, red);
    block_add(sig_atom + 6 * i + 3, sxy, red);
    block_add(sig_atom + 6 * i + 4, sxz, red);
    block_add(sig_atom + 6 * i + 5, syz, red);
  }
}

// ------------------------------------------------------ dCN/dR chain rule
// ncoord.f90:152-228 add_coordination_number_derivs, one wavefront per atom.
template <typename R>
__global__ __launch_bounds__(kBlock) void cn_chain_kernel(DevBatch b, const double* __restrict__ dedcn,
                                                          double* __restrict__ grad,
                                                          double* __restrict__ sig_atom) {
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
    // weight of the virial term: the reference adds ds*(dEdcn_i + dEdcn_j)
    // once per unordered pair and ds*dEdcn_i per self-image
    const R de = R(dei + dedcn[j]);
    const R wsig = (j == i) ? R(dei) : R(0.5) * de;
    for (int t = 0; t < g.img_cn_count; ++t) {
      const double4 tv = b.img[g.img_cn_begin + t];
      double u0, u1, u2;
      const double r2d = pair_vec(xi, xj, tv, swap, u0, u1, u2);
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
                                                        double* __restrict__ energy,
                                                        double* __restrict__ sigma) {
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
      double* sg = sigma + 9 * s;  // sigma(a,b) at [b*3+a]
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

// ATM triple terms (atm.f90:706-742): ordered (a, b) of L(i) with
// atom(b) <= atom(a) and eps <= r2_ab <= cut3^2.
__global__ __launch_bounds__(kBlock) void count_triples_kernel(DevBatch b, const long long* __restrict__ offset,
                                                               const AtmEntry* __restrict__ list,
                                                               unsigned long long* __restrict__ counts) {
  __shared__ unsigned long long red[kWavesPerBlock];
  const int i = blockIdx.x;
  const int s = b.sys[i];
  const SysParam p = b.par[s];
  const double cut2 = p.cut3 * p.cut3;
  const long long L0 = offset[i], L1 = offset[i + 1];
  unsigned long long n = 0;
  for (long long a = L0 + threadIdx.x; a < L1; a += blockDim.x) {
    const AtmEntry ea = list[a];
    for (long long bb = L0; bb < L0 + ea.hi; ++bb) {
      double v0, v1, v2;
      const double r2 = diff_norm2(list[bb], ea, v0, v1, v2);
      n += (r2 > cut2 || r2 < kEps) ? 0 : 1;
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

// ------------------------------------------------------------ device memory
class DeviceBuffer {
 public:
  ~DeviceBuffer() { if (ptr_) (void)hipFree(ptr_); }
  void* get(size_t bytes) {
    if (bytes > cap_) {
      if (ptr_) hip_check(hipFree(ptr_), "hipFree");
      ptr_ = nullptr;
      cap_ = 0;
      size_t want = std::max<size_t>(bytes + bytes / 4, 256);
      hip_check(hipMalloc(&ptr_, want), "hipMalloc");
      cap_ = want;
    }
    return ptr_;
  }
  template <typename T> T* as(size_t count) { return static_cast<T*>(get(std::max<size_t>(count, 1) * sizeof(T))); }

 private:
  void* ptr_ = nullptr;
  size_t cap_ = 0;
};

int blocks_for(long long n, int per_block) { return static_cast<int>((n + per_block - 1) / per_block); }

}  // namespace

// cutoff.f90:140-227 get_lattice_points (host, FP64, same operation order).
std::vector<double> lattice_points(const int periodic[3], const double lattice[9], double rthr) {
  const int np = periodic[0] + periodic[1] + periodic[2];
  if (np == 0) return {0.0, 0.0, 0.0};
  double lat[3][3];
  std::memcpy(lat, lattice, sizeof lat);
  auto cross = [](const double* a, const double* c, double* out) {
    out[0] = a[1] * c[2] - c[1] * a[2];
    out[1] = a[2] * c[0] - c[2] * a[0];
    out[2] = a[0] * c[1] - c[0] * a[1];
  };
  auto translations = [&](double l[3][3], int rep[3]) {
    double nrm[3][3];
    cross(l[1], l[2], nrm[0]);
    cross(l[2], l[0], nrm[1]);
    cross(l[0], l[1], nrm[2]);
    for (int k = 0; k < 3; ++k) {
      double len = std::sqrt(nrm[k][0] * nrm[k][0] + nrm[k][1] * nrm[k][1] + nrm[k][2] * nrm[k][2]);
      for (int c = 0; c < 3; ++c) nrm[k][c] = nrm[k][c] / len;
    }
    for (int k = 0; k < 3; ++k) {
      double cs = nrm[k][0] * l[k][0] + nrm[k][1] * l[k][1] + nrm[k][2] * l[k][2];
      rep[k] = static_cast<int>(std::ceil(std::fabs(rthr / cs)));
    }
  };
  int rep[3] = {0, 0, 0};
  if (np == 3) {
    translations(lat, rep);
  } else if (np == 1) {
    for (int k = 0; k < 3; ++k) {
      double len = std::sqrt(lat[k][0] * lat[k][0] + lat[k][1] * lat[k][1] + lat[k][2] * lat[k][2]);
      rep[k] = periodic[k] ? static_cast<int>(std::ceil(rthr / len)) : 0;
    }
  } else {
    double nrm[3][3];
    for (int k = 0; k < 3; ++k) cross(lat[(k + 1) % 3], lat[(k + 2) % 3], nrm[k]);
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c)
        if (periodic[c]) nrm[k][c] = lat[k][c];  // where(spread(periodic, 2, 3))
    translations(nrm, rep);
    for (int k = 0; k < 3; ++k)
      if (!periodic[k]) rep[k] = 0;
  }
  std::vector<double> out;
  out.reserve(3 * (2 * rep[0] + 1) * (2 * rep[1] + 1) * (2 * rep[2] + 1));
  for (int ix = -rep[0]; ix <= rep[0]; ++ix)
    for (int iy = -rep[1]; iy <= rep[1]; ++iy)
      for (int iz = -rep[2]; iz <= rep[2]; ++iz)
        for (int c = 0; c < 3; ++c) out.push_back(lat[0][c] * ix + lat[1][c] * iy + lat[2][c] * iz);
  return out;
}

// Pinned host buffer for the metadata upload; an event recorded after the copy
// guards its reuse by the next call.
class PinnedStaging {
 public:
  ~PinnedStaging() {
    if (event_) {
      (void)hipEventSynchronize(event_);
      (void)hipEventDestroy(event_);
    }
    if (ptr_) (void)hipHostFree(ptr_);
  }
  char* get(size_t bytes) {
    if (event_) hip_check(hipEventSynchronize(event_), "staging reuse");
    if (bytes > cap_) {
      if (ptr_) hip_check(hipHostFree(ptr_), "hipHostFree");
      ptr_ = nullptr;
      cap_ = 0;
      size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
      hip_check(hipHostMalloc(reinterpret_cast<void**>(&ptr_), want, hipHostMallocDefault), "hipHostMalloc");
      cap_ = want;
    }
    return ptr_;
  }
  void record(hipStream_t stream) {
    if (!event_) hip_check(hipEventCreateWithFlags(&event_, hipEventDisableTiming), "hipEventCreate");
    hip_check(hipEventRecord(event_, stream), "hipEventRecord");
  }

 private:
  char* ptr_ = nullptr;
  size_t cap_ = 0;
  hipEvent_t event_ = nullptr;
};


// Optional per-kernel timing with HIP events recorded on the launch stream.
class KernelTimer {
 public:
  ~KernelTimer() { clear(); }
  bool enabled = false;
  void begin(hipStream_t st) {
    if (!enabled) return;
    hipEvent_t a;
    hip_check(hipEventCreate(&a), "hipEventCreate");
    hip_check(hipEventRecord(a, st), "hipEventRecord");
    open_ = a;
  }
  void end(const char* name, hipStream_t st) {
    if (!enabled) return;
    hipEvent_t b;
    hip_check(hipEventCreate(&b), "hipEventCreate");
    hip_check(hipEventRecord(b, st), "hipEventRecord");
    pending_.push_back({name, open_, b});
  }
  void collect() {
    for (auto& p : pending_) {
      hip_check(hipEventSynchronize(p.b), "hipEventSynchronize");
      float ms = 0.0f;
      hip_check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
      auto& t = totals_[p.name];
      t.first += ms;
      t.second += 1;
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    pending_.clear();
  }
  void clear() {
    for (auto& p : pending_) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    pending_.clear();
    totals_.clear();
  }
  std::pair<double, long long> get(const std::string& name) const {
    auto it = totals_.find(name);
    return it == totals_.end() ? std::make_pair(0.0, 0LL) : it->second;
  }

 private:
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  hipEvent_t open_ = nullptr;
  std::vector<Rec> pending_;
  std::map<std::string, std::pair<double, long long>> totals_;
};

// Device description of a prepared batch.
struct Prepared {
  DevBatch db;
  int natoms = 0, nsys = 0, nsp = 0, mref = 0;
  long long c6_total = 0;
  bool any_atm = false;
  const int* ghost = nullptr;
};

struct Engine::Impl {
  DeviceBuffer meta, pos, cn, gw, dgw, c6, dc6, dc6t, e_atom, dedcn, grad, sig, cnt, off, list, counts;
  std::vector<char> staging;
  PinnedStaging pinned;
  KernelTimer timer;

  Prepared prepare(const std::vector<SystemDesc>& systems, const int* numbers, const double* positions_dev,
                   hipStream_t stream, const int* ghost_host);
  long long build_atm_lists(const Prepared& pb, hipStream_t stream, long long** off, AtmEntry** list);

  template <typename F>
  void timed(const char* name, hipStream_t st, F&& launch) {
    timer.begin(st);
    launch();
    timer.end(name, st);
  }
};

Engine::Engine() : impl_(new Impl) {}
Engine::~Engine() { delete impl_; }

void Engine::set_kernel_timing(bool enable) {
  impl_->timer.clear();
  impl_->timer.enabled = enable;
}

std::pair<double, long long> Engine::kernel_time(const std::string& name) const { return impl_->timer.get(name); }

namespace {
template <typename T>
size_t stage(std::vector<char>& buf, const std::vector<T>& v) {
  size_t off = (buf.size() + 255) & ~size_t(255);
  buf.resize(off + v.size() * sizeof(T));
  if (!v.empty()) std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
  return off;
}
}  // namespace

// Host planning of a batch: species of the batch, per-species tables, lattice
// images per system and cutoff, C6 block offsets; one metadata upload and the
// packing of the positions.
Prepared Engine::Impl::prepare(const std::vector<SystemDesc>& systems, const int* numbers,
                               const double* positions_dev, hipStream_t stream, const int* ghost_host) {
  Prepared pb;
  const int nsys = static_cast<int>(systems.size());
  int natoms = 0;
  for (const auto& s : systems) {
    if (s.natoms <= 0) throw Error("Invalid molecular structure data provided");
    natoms += s.natoms;
  }
  const auto& tab = element_tables();
  std::map<int, int> zmap;
  std::vector<int> zlist;
  std::vector<int> sp(natoms), sysidx(natoms);
  for (int k = 0, a = 0; k < nsys; ++k)
    for (int q = 0; q < systems[k].natoms; ++q, ++a) {
      const int z = numbers[a];
      if (z < 1 || z > kMaxElem) throw Error("Unsupported element with atomic number " + std::to_string(z));
      auto it = zmap.find(z);
      if (it == zmap.end()) {
        it = zmap.emplace(z, static_cast<int>(zlist.size())).first;
        zlist.push_back(z);
      }
      sp[a] = it->second;
      sysidx[a] = k;
    }
  const int nsp = static_cast<int>(zlist.size());
  if (nsp > kMaxSpecies) throw Error("Too many distinct elements in one batch");
  int mref = 0;
  std::vector<double> rcov(nsp), r4r2(nsp), refcn(nsp * kMaxRef, -1.0),
      c6ref(size_t(nsp) * nsp * kMaxRef * kMaxRef, 0.0), rvdw(nsp * nsp);
  std::vector<int> nref(nsp);
  for (int a = 0; a < nsp; ++a) {
    const int z = zlist[a];
    rcov[a] = covalent_radius(z);
    r4r2[a] = r4r2_value(z);
    nref[a] = tab.nref[z - 1];
    mref = std::max(mref, nref[a]);
    for (int r = 0; r < nref[a]; ++r) refcn[a * kMaxRef + r] = tab.refcn[z - 1][r];
    for (int c = 0; c < nsp; ++c) {
      const int zc = zlist[c];
      rvdw[a * nsp + c] = vdw_radius(zc, z);
      for (int ri = 0; ri < nref[a]; ++ri)
        for (int rj = 0; rj < tab.nref[zc - 1]; ++rj)
          c6ref[(size_t(a) * nsp + c) * kMaxRef * kMaxRef + ri * kMaxRef + rj] = reference_c6(ri, rj, z, zc);
    }
  }
  std::vector<SysGeom> geom(nsys);
  std::vector<SysParam> par(nsys);
  std::vector<double4> img;
  long long c6_total = 0;
  bool any_atm = false;
  for (int k = 0, a = 0; k < nsys; ++k) {
    const SystemDesc& sd = systems[k];
    SysGeom& g = geom[k];
    g.atom_begin = a;
    g.natoms = sd.natoms;
    a += sd.natoms;
    par[k] = sd.param;
    auto add = [&](double cut, int& begin, int& count) {
      std::vector<double> tr = lattice_points(sd.periodic, sd.lattice, cut);
      begin = static_cast<int>(img.size());
      count = static_cast<int>(tr.size() / 3);
      for (int t = 0; t < count; ++t) img.push_back(make_double4(tr[3 * t], tr[3 * t + 1], tr[3 * t + 2], 0.0));
    };
    add(sd.param.cut_cn, g.img_cn_begin, g.img_cn_count);
    add(sd.param.cut2, g.img2_begin, g.img2_count);
    add(sd.param.cut3, g.img3_begin, g.img3_count);
    g.c6_begin = c6_total;
    c6_total += static_cast<long long>(sd.natoms) * sd.natoms;
    if (std::fabs(sd.param.s9) >= kEps) any_atm = true;
  }
  std::vector<char>& sb = staging;
  sb.clear();
  std::vector<int> ghost;
  if (ghost_host) ghost.assign(ghost_host, ghost_host + natoms);
  const size_t o_ghost = stage(sb, ghost), o_sp = stage(sb, sp), o_sys = stage(sb, sysidx), o_geom = stage(sb, geom),
               o_par = stage(sb, par), o_img = stage(sb, img), o_rcov = stage(sb, rcov), o_r4 = stage(sb, r4r2),
               o_nref = stage(sb, nref), o_refcn = stage(sb, refcn), o_c6r = stage(sb, c6ref),
               o_rvdw = stage(sb, rvdw);
  char* dmeta = static_cast<char*>(meta.get(sb.size()));
  char* host = pinned.get(sb.size());
  std::memcpy(host, sb.data(), sb.size());
  hip_check(hipMemcpyAsync(dmeta, host, sb.size(), hipMemcpyHostToDevice, stream), "upload");
  pinned.record(stream);

  DevBatch& db = pb.db;
  db.natoms = natoms; db.nsys = nsys; db.nsp = nsp; db.mref = mref;
  db.pos = pos.as<double4>(natoms);
  db.sp = reinterpret_cast<const int*>(dmeta + o_sp);
  db.sys = reinterpret_cast<const int*>(dmeta + o_sys);
  db.geom = reinterpret_cast<const SysGeom*>(dmeta + o_geom);
  db.par = reinterpret_cast<const SysParam*>(dmeta + o_par);
  db.img = reinterpret_cast<const double4*>(dmeta + o_img);
  db.rcov = reinterpret_cast<const double*>(dmeta + o_rcov);
  db.r4r2 = reinterpret_cast<const double*>(dmeta + o_r4);
  db.nref = reinterpret_cast<const int*>(dmeta + o_nref);
  db.refcn = reinterpret_cast<const double*>(dmeta + o_refcn);
  db.c6ref = reinterpret_cast<const double*>(dmeta + o_c6r);
  db.rvdw = reinterpret_cast<const double*>(dmeta + o_rvdw);
  pb.natoms = natoms; pb.nsys = nsys; pb.nsp = nsp; pb.mref = mref;
  pb.c6_total = c6_total;
  pb.any_atm = any_atm;
  pb.ghost = ghost_host ? reinterpret_cast<const int*>(dmeta + o_ghost) : nullptr;
  timed("pack", stream, [&] {
    hipLaunchKernelGGL(pack_positions, dim3(blocks_for(natoms, kBlock)), dim3(kBlock), 0, stream, positions_dev,
                       natoms, const_cast<double4*>(db.pos));
  });
  return pb;
}

// Three-body neighbour lists: count, scan, one read of the total, fill.
long long Engine::Impl::build_atm_lists(const Prepared& pb, hipStream_t stream, long long** off_out,
                                        AtmEntry** list_out) {
  const int natoms = pb.natoms;
  const int wave_grid = blocks_for(natoms, kWavesPerBlock);
  long long* c = cnt.as<long long>(natoms);
  long long* o = off.as<long long>(natoms + 1);
  timed("atm_list", stream, [&] {
    hipLaunchKernelGGL(atm_list_kernel<false>, dim3(wave_grid), dim3(kBlock), 0, stream, pb.db, c,
                       (const long long*)nullptr, (AtmEntry*)nullptr);
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, stream, c, natoms, o);
  });
  long long total = 0;
  hip_check(hipMemcpyAsync(&total, o + natoms, sizeof total, hipMemcpyDeviceToHost, stream), "list size");
  hip_check(hipStreamSynchronize(stream), "list size sync");
  AtmEntry* l = list.as<AtmEntry>(total);
  timed("atm_list", stream, [&] {
    hipLaunchKernelGGL(atm_list_kernel<true>, dim3(wave_grid), dim3(kBlock), 0, stream, pb.db, c,
                       (const long long*)o, l);
  });
  *off_out = o;
  *list_out = l;
  return total;
}

void Engine::run(const std::vector<SystemDesc>& systems, const int* numbers, const double* positions_dev,
                 int precision, const BatchOutput& out, hipStream_t stream, const BatchDebug* debug,
                 RunStats* stats, const int* ghost_host) {
  if (systems.empty()) return;
  Impl& m = *impl_;
  const Prepared pb = m.prepare(systems, numbers, positions_dev, stream, ghost_host);
  const DevBatch& db = pb.db;
  const int natoms = pb.natoms;
  const bool grad = out.gradient || out.sigma;
  double* cn = debug && debug->cn ? debug->cn : m.cn.as<double>(natoms);
  double* gw = m.gw.as<double>(size_t(natoms) * pb.mref);
  double* dgw = grad ? m.dgw.as<double>(size_t(natoms) * pb.mref) : nullptr;
  double* c6 = m.c6.as<double>(pb.c6_total);
  double* dc6 = grad ? m.dc6.as<double>(pb.c6_total) : nullptr;
  double* dc6t = grad ? m.dc6t.as<double>(pb.c6_total) : nullptr;
  double* e_atom = m.e_atom.as<double>(natoms);
  double* dedcn = debug && debug->dedcn ? debug->dedcn : m.dedcn.as<double>(natoms);
  double* gout = out.gradient ? out.gradient : m.grad.as<double>(size_t(natoms) * 3);
  double* sig = m.sig.as<double>(size_t(natoms) * 6);
  const int wave_grid = blocks_for(natoms, kWavesPerBlock);
  const bool mixed = precision == kMixed;
  const dim3 blk(kBlock);

  m.timed("cn", stream, [&] {
    if (mixed) hipLaunchKernelGGL(cn_kernel<float>, dim3(wave_grid), blk, 0, stream, db, cn);
    else hipLaunchKernelGGL(cn_kernel<double>, dim3(wave_grid), blk, 0, stream, db, cn);
  });
  m.timed("weights", stream, [&] {
    hipLaunchKernelGGL(weights_kernel, dim3(blocks_for(natoms, kBlock)), blk, 0, stream, db, cn, gw, dgw);
  });
  m.timed("c6", stream, [&] {
    hipLaunchKernelGGL(c6_kernel, dim3(natoms), blk, 0, stream, db, gw, dgw, pb.ghost, c6, dc6, dc6t);
  });
  m.timed("disp2", stream, [&] {
    if (grad && mixed)
      hipLaunchKernelGGL((disp2_kernel<float, true>), dim3(wave_grid), blk, 0, stream, db, c6, dc6, e_atom, dedcn,
                         gout, sig);
    else if (grad)
      hipLaunchKernelGGL((disp2_kernel<double, true>), dim3(wave_grid), blk, 0, stream, db, c6, dc6, e_atom, dedcn,
                         gout, sig);
    else if (mixed)
      hipLaunchKernelGGL((disp2_kernel<float, false>), dim3(wave_grid), blk, 0, stream, db, c6, dc6, e_atom, dedcn,
                         gout, sig);
    else
      hipLaunchKernelGGL((disp2_kernel<double, false>), dim3(wave_grid), blk, 0, stream, db, c6, dc6, e_atom, dedcn,
                         gout, sig);
  });
  if (pb.any_atm) {
    long long* off = nullptr;
    AtmEntry* list = nullptr;
    const long long total = m.build_atm_lists(pb, stream, &off, &list);
    if (stats) stats->atm_list_entries = total;
    m.timed("atm", stream, [&] {
      if (grad && mixed)
        hipLaunchKernelGGL((atm_kernel<float, true>), dim3(natoms), blk, 0, stream, db, off, list, c6, dc6, dc6t,
                           e_atom, dedcn, gout, sig);
      else if (grad)
        hipLaunchKernelGGL((atm_kernel<double, true>), dim3(natoms), blk, 0, stream, db, off, list, c6, dc6, dc6t,
                           e_atom, dedcn, gout, sig);
      else if (mixed)
        hipLaunchKernelGGL((atm_kernel<float, false>), dim3(natoms), blk, 0, stream, db, off, list, c6, dc6, dc6t,
                           e_atom, dedcn, gout, sig);
      else
        hipLaunchKernelGGL((atm_kernel<double, false>), dim3(natoms), blk, 0, stream, db, off, list, c6, dc6, dc6t,
                           e_atom, dedcn, gout, sig);
    });
  }
  if (grad) {
    m.timed("cn_chain", stream, [&] {
      if (mixed) hipLaunchKernelGGL(cn_chain_kernel<float>, dim3(wave_grid), blk, 0, stream, db, dedcn, gout, sig);
      else hipLaunchKernelGGL(cn_chain_kernel<double>, dim3(wave_grid), blk, 0, stream, db, dedcn, gout, sig);
    });
  }
  m.timed("reduce", stream, [&] {
    hipLaunchKernelGGL(reduce_kernel, dim3(pb.nsys), blk, 0, stream, db, e_atom, sig, out.energy,
                       grad ? out.sigma : nullptr);
  });
  hip_check(hipGetLastError(), "kernel launch");
  if (m.timer.enabled) m.timer.collect();
}

void Engine::count_terms(const std::vector<SystemDesc>& systems, const int* numbers, const double* positions_dev,
                         long long* counts_host, hipStream_t stream) {
  if (systems.empty()) return;
  Impl& m = *impl_;
  std::vector<SystemDesc> sys = systems;
  for (auto& s : sys)
    if (std::fabs(s.param.s9) < kEps) s.param.s9 = 1.0;  // triples are counted whatever s9 is
  const Prepared pb = m.prepare(sys, numbers, positions_dev, stream, nullptr);
  unsigned long long* dc = m.counts.as<unsigned long long>(size_t(pb.nsys) * 3);
  hip_check(hipMemsetAsync(dc, 0, sizeof(unsigned long long) * 3 * pb.nsys, stream), "hipMemsetAsync");
  hipLaunchKernelGGL(count_pairs_kernel, dim3(blocks_for(pb.natoms, kWavesPerBlock)), dim3(kBlock), 0, stream,
                     pb.db, dc);
  long long* off = nullptr;
  AtmEntry* list = nullptr;
  m.build_atm_lists(pb, stream, &off, &list);
  hipLaunchKernelGGL(count_triples_kernel, dim3(pb.natoms), dim3(kBlock), 0, stream, pb.db, off, list, dc);
  hip_check(hipGetLastError(), "kernel launch");
  std::vector<unsigned long long> h(3 * pb.nsys);
  hip_check(hipMemcpyAsync(h.data(), dc, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, stream),
            "download counts");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  for (size_t k = 0; k < h.size(); ++k) counts_host[k] = static_cast<long long>(h[k]);
}

}  // namespace d3
