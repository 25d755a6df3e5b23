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

#include <hipcub/device/device_radix_sort.hpp>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <map>
#include <mutex>
#include <string>

#include "d3_hess3.hpp"
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
  R pdmp;             // (r0_ij / r_ij)^(alp/3), r0_ij = 4/3 rvdw_ij, alp = alp6 + 2
  R sc6h;             // sqrt(C6_ij), halved for an image of atom i itself (j == i)
  R qi;               // (dC6_ij / dCN_i) / C6_ij
  R qx;               // (dC6_ij / dCN_j) / C6_ij
  int atom;           // j (global index)
  int rank;           // position of j in the spatial order of its system
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

// Calls image(t) for the images t of one set (kImgCn, kImg2, kImg3) that can
// bring atoms i and j within the set's cutoff, in the order of the list:
// |x_i - x_j - T_n| <= cut bounds each index n_d to within cut |b_d| / 2 pi
// of the fractional coordinate f_d of x_i - x_j, so only that sub-box of the
// index box is visited (3D-periodic systems); otherwise every image.  The
// caller still applies the exact distance test, so the accepted images and
// their order are those of the full list.
template <typename F>
__device__ __forceinline__ void for_images(const SysGeom& g, int set, int count, const double4& xi,
                                           const double4& xj, F&& image) {
  const int* box = g.box[set];
  if (box[0] < 0) {
    for (int t = 0; t < count; ++t) image(t);
    return;
  }
  int lo[3], hi[3];
  const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double f = g.frac[3 * d] * dx + g.frac[3 * d + 1] * dy + g.frac[3 * d + 2] * dz;
    lo[d] = max(-box[d], static_cast<int>(ceil(f - g.span[set][d] - 1e-6)));
    hi[d] = min(box[d], static_cast<int>(floor(f + g.span[set][d] + 1e-6)));
  }
  const int w1 = 2 * box[1] + 1, w2 = 2 * box[2] + 1;
  for (int ix = lo[0]; ix <= hi[0]; ++ix)
    for (int iy = lo[1]; iy <= hi[1]; ++iy) {
      const int base = ((ix + box[0]) * w1 + (iy + box[1])) * w2 + box[2];
      for (int iz = lo[2]; iz <= hi[2]; ++iz) image(base + iz);
    }
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

// Whether every position of tile t (64 consecutive positions of a system's
// spatial order) lies farther than sqrt(cut2) from x: the distance from x to
// the tile's bounding box exceeds the cutoff (with a margin against the
// rounding of the box distance), so every exact pair test with the tile
// fails and a pair loop may skip it.  Systems without lattice images only.
__device__ __forceinline__ bool tile_beyond(const DevBatch& b, int t, const double4& x, double cut2) {
  const double4 lo = b.tbox[2 * t], hi = b.tbox[2 * t + 1];
  const double dx = fmax(0.0, fmax(lo.x - x.x, x.x - hi.x));
  const double dy = fmax(0.0, fmax(lo.y - x.y, x.y - hi.y));
  const double dz = fmax(0.0, fmax(lo.z - x.z, x.z - hi.z));
  return dx * dx + dy * dy + dz * dz > cut2 * (1.0 + 1e-12) + 1e-12;
}

// Tiles t0 .. t0 + 63 of a system (bit k: tile t0 + k) that a pair loop of
// the position x with cutoff sqrt(cut2) has to visit: the box tests of 64
// tiles at once, one per lane, so that a loop over the tiles tests a bit
// instead of waiting for a box load per tile.  Every tile unless g.cull.
__device__ __forceinline__ unsigned long long near_tiles(const DevBatch& b, const SysGeom& g, int t0,
                                                         const double4& x, double cut2) {
  if (!g.cull) return ~0ull;
  const int t = t0 + wave_lane();
  bool near = false;
  if (t * kWave < g.natoms) near = !tile_beyond(b, g.tile_begin + t, x, cut2);
  return __ballot(near);
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

void hip_check_sort(hipError_t e) {
  if (e != hipSuccess) throw std::runtime_error(std::string("hipcub radix sort: ") + hipGetErrorString(e));
}

// --------------------------------------------------------------- packing
__global__ void pack_positions(const double* __restrict__ xyz, int n, double4* __restrict__ pos) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pos[i] = make_double4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.0);
}

// ----------------------------------------------------------- spatial order
// Molecules: Morton code of the 4-bohr cell of every atom (10 bits per axis,
// coordinates wrapped modulo 4096 bohr).  Systems with lattice images keep the
// input order (their lists hold several images of an atom, which a cell-based
// order does not keep together; measured slower).  Keyed by the system index
// so that one sort orders every system.
__device__ __forceinline__ unsigned spread_bits3(unsigned v) {
  v &= 1023u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ void spatial_key_kernel(DevBatch b, unsigned long long* __restrict__ keys, int* __restrict__ idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.natoms) return;
  const int s = b.sys[i];
  const SysGeom g = b.geom[s];
  unsigned code;
  if (g.img_cn_count > 1 || g.img2_count > 1 || g.img3_count > 1) {
    code = static_cast<unsigned>(i - g.atom_begin);
  } else {
    const double4 x = b.pos[i];
    const unsigned qx = static_cast<unsigned>(static_cast<int>(floor(x.x * 0.25)));
    const unsigned qy = static_cast<unsigned>(static_cast<int>(floor(x.y * 0.25)));
    const unsigned qz = static_cast<unsigned>(static_cast<int>(floor(x.z * 0.25)));
    code = spread_bits3(qx) | (spread_bits3(qy) << 1) | (spread_bits3(qz) << 2);
  }
  keys[i] = (static_cast<unsigned long long>(s) << 30) | code;
  idx[i] = i;
}

__global__ void spatial_rank_kernel(DevBatch b, const int* __restrict__ order, int* __restrict__ rank) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= b.natoms) return;
  const int a = order[p];
  rank[a] = p - b.geom[b.sys[a]].atom_begin;
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
  auto partner = [&](int j, const double4& xj) {
    const R rc = rci + R(b.rcov[b.sp[j]]);
    const bool swap = j > i;
    for_images(g, kImgCn, g.img_cn_count, xi, xj, [&](int t) {
      double u0, u1, u2;
      const double r2 = pair_vec(xi, xj, b.img[g.img_cn_begin + t], swap, u0, u1, u2);
      if (r2 > cut2 || r2 < kCnFloor) return;
      const R r1 = d_sqrt(R(r2));
      acc += double(d_rcp(R(1) + d_exp(-R(kKcn) * (rc / r1 - R(1)))));
    });
  };
  if (g.cull) {  // partners in spatial order, tiles beyond the cutoff skipped
    unsigned long long near = 0;
    for (int r0 = 0; r0 < g.natoms; r0 += kWave) {
      const int t = r0 / kWave;
      if ((t & (kWave - 1)) == 0) near = near_tiles(b, g, t, xi, cut2);
      if (!((near >> (t & (kWave - 1))) & 1ull) || r0 + lane >= g.natoms) continue;
      const int x = g.atom_begin + r0 + lane;
      partner(b.order[x], b.posr[x]);
    }
  } else {
    for (int j = g.atom_begin + lane; j < g.atom_begin + g.natoms; j += kWave) partner(j, b.pos[j]);
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
// term, atm.f90:667-668,811-816).  One block per row i of the system block;
// rows and columns are laid out in the spatial order of the atoms.
// The weights, dw/dCN, species and ghost flags of the atoms at position x of
// the batch's spatial order, reference-major: gwr[r][x], gwr[mref + r][x]
// (with dw/dCN) and spr[x] = species | kGhostBit for a ghost atom, so that
// the lanes of pair_kernel, which walk a row in that order, read consecutive
// addresses.
constexpr int kGhostBit = 1 << 20;
__global__ __launch_bounds__(kBlock) void rank_weights_kernel(DevBatch b, const double* __restrict__ gw,
                                                              const double* __restrict__ dgw,
                                                              double* __restrict__ gwr, int* __restrict__ spr) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= b.natoms) return;
  const int j = b.order[x];
  const int mref = b.mref, nat = b.natoms;
  for (int r = 0; r < mref; ++r) {
    gwr[static_cast<long long>(r) * nat + x] = gw[j * mref + r];
    if (dgw) gwr[static_cast<long long>(mref + r) * nat + x] = dgw[j * mref + r];
  }
  spr[x] = b.sp[j] | (b.ghost && b.ghost[j] ? kGhostBit : 0);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void pair_kernel(DevBatch b, const double* __restrict__ gw,
                                                      const double* __restrict__ dgw,
                                                      const double* __restrict__ gwr, const int* __restrict__ spr,
                                                      R4<R>* __restrict__ pair) {
  const int i = blockIdx.x;
  const int s = b.sys[i];
  const SysGeom g = b.geom[s];
  const double alp3 = (b.par[s].alp + 2.0) / 3.0;
  const int n = g.natoms;
  const int spi = b.sp[i];
  const int mref = b.mref;
  const int nri = b.nref[spi];
  double gi[kMaxRef], dgi[kMaxRef];
  for (int r = 0; r < kMaxRef; ++r) {
    gi[r] = r < nri ? gw[i * mref + r] : 0.0;
    dgi[r] = (r < nri && dgw) ? dgw[i * mref + r] : 0.0;
  }
  // per species s of the batch: h[s][r] = sum_ri w_i[ri] C6ref[ri][r] and the
  // same with dw_i/dCN_i, so that a record costs three nref_j-term dots; and
  // (4/3 rvdw_is)^(alp/3), in dynamic LDS ((2 kMaxRef + 1) nsp doubles, sized
  // at the launch).
  extern __shared__ double pair_lds[];
  double* __restrict__ h = pair_lds;                   // [nsp][kMaxRef]
  double* __restrict__ dh = pair_lds + b.nsp * kMaxRef;  // [nsp][kMaxRef]
  double* __restrict__ r0p_s = dh + b.nsp * kMaxRef;     // [nsp]
  for (int t = threadIdx.x; t < b.nsp * kMaxRef; t += blockDim.x) {
    const int sj = t / kMaxRef, rj = t % kMaxRef;
    const double* ref = b.c6ref + (spi * b.nsp + sj) * (kMaxRef * kMaxRef);
    double a = 0.0, d = 0.0;
    for (int ri = 0; ri < nri; ++ri) {
      a += gi[ri] * ref[ri * kMaxRef + rj];
      d += dgi[ri] * ref[ri * kMaxRef + rj];
    }
    h[t] = a;
    dh[t] = d;
  }
  for (int t = threadIdx.x; t < b.nsp; t += blockDim.x) r0p_s[t] = pow(kRs9 * b.rvdw[t * b.nsp + spi], alp3);
  __syncthreads();
  const bool gi_ghost = b.ghost && b.ghost[i];
  R4<R>* row = pair + g.c6_begin + static_cast<long long>(b.rank[i]) * n;
  const long long nat = b.natoms;
  for (int jr = threadIdx.x; jr < n; jr += blockDim.x) {
    const int x = g.atom_begin + jr;  // position of partner j in the spatial order
    const int code = spr[x];
    const int spj = code & (kGhostBit - 1);
    const int nrj = b.nref[spj];
    double cc = 0.0, di = 0.0, dj = 0.0;
    if (!(gi_ghost || (code & kGhostBit))) {  // model.f90:436-440
      for (int rj = 0; rj < nrj; ++rj) {
        const double w = gwr[rj * nat + x];
        cc += w * h[spj * kMaxRef + rj];
        if (dgw) {
          di += w * dh[spj * kMaxRef + rj];
          dj += gwr[(mref + rj) * nat + x] * h[spj * kMaxRef + rj];
        }
      }
    }
    const double ic = cc > 0.0 ? 1.0 / cc : 0.0;
    row[jr] = make_r4<R>(R(sqrt(cc)), R(di * ic), R(dj * ic), R(r0p_s[spj]));
  }
}

// ------------------------------------------------- two-body damping families
// Per-pair constants of the damping (what the reference hoists out of the
// image loop) and one pair term without the switching function:
// edisp0 (E = -C6 sw edisp0 / 2 per image), gdisp0 (dE/dr / r) and, for the
// Z damping whose damping depends on C6, dedc6 = d(C6 edisp0)/dC6.
//   rational   rational.f90:452-481   k0 = r0^6, k1 = r0^8, r0 = a1 sqrt(rrij) + a2
//   zero       zero.f90:484-512       k0 = rs6 R0, k1 = rs8 R0
//   mzero      mzero.f90:441-470      k0 = 1/(rs6 R0), k1 = 1/(rs8 R0), k2 = bet R0,
//                                     k3 = rs6 bet R0^2, k4 = rs8 bet R0^2
//   opt. power optimizedpower.f90     k0 = r0^bet r0^6, k1 = r0^bet r0^8
//   CSO        cso.f90                k0 = (a3 sqrt(rrij) + a4)^6, k1 = a2 sqrt(rrij)
//   Z          z.f90                  k0 = a1/(id_i + id_j) C6, k1 = k0 rrij
template <typename R>
struct DampPair {
  R rrij, k0, k1, k2, k3, k4;
};

template <typename R, int kKind>
__device__ __forceinline__ DampPair<R> damp_setup(const SysParam& p, R rrij, double rvdw, R c6ij, int idsum) {
  DampPair<R> d{rrij, R(0), R(0), R(0), R(0), R(0)};
  switch (kKind) {
    case kRational:
    case kOptimizedPower: {
      const R r0 = R(p.a1) * d_sqrt(rrij) + R(p.a2);
      const R r02 = r0 * r0;
      d.k0 = r02 * r02 * r02;
      d.k1 = d.k0 * r02;
      if (kKind == kOptimizedPower) {
        const R ab = d_pow(r0, R(p.bet));
        d.k0 *= ab;
        d.k1 *= ab;
      }
      break;
    }
    case kZero: {
      const R r0 = R(rvdw);
      d.k0 = R(p.a1) * r0;
      d.k1 = R(p.a2) * r0;
      break;
    }
    case kModifiedZero: {
      const R r0 = R(rvdw);
      d.k0 = d_rcp(R(p.a1) * r0);
      d.k1 = d_rcp(R(p.a2) * r0);
      d.k2 = R(p.bet) * r0;
      d.k3 = R(p.a1) * R(p.bet) * r0 * r0;
      d.k4 = R(p.a2) * R(p.bet) * r0 * r0;
      break;
    }
    case kCSO: {
      const R r0 = d_sqrt(rrij);
      const R d6 = R(p.a3) * r0 + R(p.a4);
      const R d3 = d6 * d6 * d6;
      d.k0 = d3 * d3;
      d.k1 = R(p.a2) * r0;
      break;
    }
    case kZDamping: {
      d.k0 = R(p.a1) / R(idsum) * c6ij;
      d.k1 = d.k0 * rrij;
      break;
    }
    default:
      break;
  }
  return d;
}

template <typename R, bool kGrad, int kKind>
__device__ __forceinline__ R damp_term(const SysParam& p, const DampPair<R>& d, R r2, R r1, R& gdisp0, R& dedc6) {
  const R s6 = R(p.s6), s8 = R(p.s8);
  R e0 = R(0);
  gdisp0 = R(0);
  dedc6 = R(0);
  switch (kKind) {
    case kRational: {
      const R r4 = r2 * r2;
      const R t6 = d_rcp(r4 * r2 + d.k0);
      const R t8 = d_rcp(r4 * r4 + d.k1);
      e0 = s6 * t6 + s8 * d.rrij * t8;
      if (kGrad) {
        const R d6 = R(-6) * r4 * t6 * t6;
        const R d8 = R(-8) * r4 * r2 * t8 * t8;
        gdisp0 = s6 * d6 + s8 * d.rrij * d8;
      }
      break;
    }
    case kZero: {
      const R alp6 = R(p.alp), alp8 = R(p.alp) + R(2);
      const R r6 = r2 * r2 * r2, r8 = r6 * r2;
      const R ir1 = d_rcp(r1);
      const R t6 = d_pow(d.k0 * ir1, alp6);
      const R t8 = d_pow(d.k1 * ir1, alp8);
      const R f6 = d_rcp(R(1) + R(6) * t6);
      const R f8 = d_rcp(R(1) + R(6) * t8);
      const R ir6 = d_rcp(r6), ir8 = d_rcp(r8);
      e0 = s6 * f6 * ir6 + s8 * d.rrij * f8 * ir8;
      if (kGrad) {
        const R ir2 = d_rcp(r2);
        const R d6 = R(-6) * f6 * ir2 + R(6) * alp6 * t6 * f6 * f6 * ir2;
        const R d8 = R(-8) * f8 * ir2 + R(6) * alp8 * t8 * f8 * f8 * ir2;
        gdisp0 = s6 * d6 * ir6 + s8 * d.rrij * d8 * ir8;
      }
      break;
    }
    case kModifiedZero: {
      const R alp6 = R(p.alp), alp8 = R(p.alp) + R(2);
      const R r6 = r2 * r2 * r2, r8 = r6 * r2;
      const R t6 = d_pow(r1 * d.k0 + d.k2, -alp6);
      const R t8 = d_pow(r1 * d.k1 + d.k2, -alp8);
      const R f6 = d_rcp(R(1) + R(6) * t6);
      const R f8 = d_rcp(R(1) + R(6) * t8);
      const R ir6 = d_rcp(r6), ir8 = d_rcp(r8);
      e0 = s6 * f6 * ir6 + s8 * d.rrij * f8 * ir8;
      if (kGrad) {
        const R ir2 = d_rcp(r2);
        const R d6 = R(-6) * f6 * ir2 + R(6) * alp6 * t6 * f6 * f6 * d_rcp(r2 + d.k3 * r1);
        const R d8 = R(-8) * f8 * ir2 + R(6) * alp8 * t8 * f8 * f8 * d_rcp(r2 + d.k4 * r1);
        gdisp0 = s6 * d6 * ir6 + s8 * d.rrij * d8 * ir8;
      }
      break;
    }
    case kOptimizedPower: {
      const R bet = R(p.bet);
      const R r4 = r2 * r2;
      const R rb = d_pow(r2, bet * R(0.5));
      const R t6 = rb * d_rcp(rb * r4 * r2 + d.k0);
      const R t8 = rb * d_rcp(rb * r4 * r4 + d.k1);
      e0 = s6 * t6 + s8 * d.rrij * t8;
      if (kGrad) {
        const R irb2 = d_rcp(rb * r2);
        const R d6 = R(-6) * r4 * t6 * t6 + bet * d.k0 * t6 * t6 * irb2;
        const R d8 = R(-8) * r4 * r2 * t8 * t8 + bet * d.k1 * t8 * t8 * irb2;
        gdisp0 = s6 * d6 + s8 * d.rrij * d8;
      }
      break;
    }
    case kCSO: {
      const R a1 = R(p.a1);
      const R sf = d_rcp(R(1) + d_exp(r1 - d.k1));
      const R sig = s6 + a1 * sf;
      const R r4 = r2 * r2;
      const R t6 = d_rcp(r4 * r2 + d.k0);
      e0 = sig * t6;
      if (kGrad) {
        const R dsig = -a1 * sf * (R(1) - sf) * d_rcp(r1);
        const R dt6 = R(-6) * r4 * t6 * t6;
        gdisp0 = dsig * t6 + sig * dt6;
      }
      break;
    }
    case kZDamping: {
      const R r4 = r2 * r2;
      const R t6 = d_rcp(r4 * r2 + d.k0);
      const R t8 = d_rcp(r4 * r4 + d.k1);
      e0 = s6 * t6 + s8 * d.rrij * t8;
      if (kGrad) {
        const R d6 = R(-6) * r4 * t6 * t6;
        const R d8 = R(-8) * r4 * r2 * t8 * t8;
        gdisp0 = s6 * d6 + s8 * d.rrij * d8;
        dedc6 = s6 * r4 * r2 * t6 * t6 + s8 * d.rrij * r4 * r4 * t8 * t8;
      }
      break;
    }
    default:
      break;
  }
  return e0;
}

// get_dispersion2 returns without any term when both scalings vanish
// (rational.f90:267 and the same line of every family; CSO tests s6 and a1).
__device__ __forceinline__ bool two_body_active(const SysParam& p) {
  if (p.kind == kCSO) return !(fabs(p.s6) < kEps && fabs(p.a1) < kEps);
  return !(fabs(p.s6) < kEps && fabs(p.s8) < kEps);
}

// --------------------------------------------------------------- two-body
// get_dispersion_derivs of every damping family (rational.f90:378-514,
// zero.f90:409-553, mzero.f90:361-508, optimizedpower.f90, cso.f90, z.f90),
// E+grad+sigma; get_dispersion_energy (rational.f90:282-374, ...) when kGrad
// is false.  dE/dCN_i gets -dC6/dCN_i edisp, for the Z damping
// dC6/dCN_i dEdc6 with dEdc6 = -sw d(C6 edisp0)/dC6 (z.f90).  One wavefront
// per atom i; atom i collects the ordered pairs (i, j, T) it takes part in.
// kPairs (energy only) instead evaluates the pairs j <= i of atom i once and
// writes their image sums, the pairwise two-body energy of
// rational.f90:623-705 / zero.f90:662-755, to both pair2 entries (i, j) and
// (j, i) of the system's n x n block (atom order, offset c6_begin), so the
// matrix is exactly symmetric like the reference's; e_atom is then only
// scratch.
template <typename R, bool kGrad, bool kPairs, int kKind>
__device__ __forceinline__ void disp2_atom(const DevBatch& b, const R4<R>* __restrict__ pair,
                                           double* __restrict__ e_atom, double* __restrict__ dedcn,
                                           double* __restrict__ grad, double* __restrict__ sig_atom,
                                           double* __restrict__ pair2, int i, const SysGeom& g,
                                           const SysParam& p) {
  const int lane = wave_lane();
  const int n = g.natoms;
  const double cut2 = p.cut2 * p.cut2;
  const double4 xi = b.pos[i];
  const int spi = b.sp[i];
  const R r4r2i = R(b.r4r2[spi]);
  double acc_e = 0.0, acc_d = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
  double sxx = 0.0, syy = 0.0, szz = 0.0, sxy = 0.0, sxz = 0.0, syz = 0.0;
  const bool active = two_body_active(p) && !p.ewald;  // reciprocal space: ewald kernels
  const R cutr = R(p.cut2), width = R(p.width2);
  const int idi = b.sid[i];
  const R4<R>* row = pair + g.c6_begin + static_cast<long long>(b.rank[i]) * n;
  if (active || kPairs) {  // the pairwise energies are written (as zeros) for s6 = s8 = 0 too
    unsigned long long near = 0;
    for (int r0 = 0; r0 < n; r0 += kWave) {
      // tiles beyond the cutoff skipped (the pairwise mode writes every pair)
      const int t = r0 / kWave;
      if (!kPairs && (t & (kWave - 1)) == 0) near = near_tiles(b, g, t, xi, cut2);
      if (!kPairs && !((near >> (t & (kWave - 1))) & 1ull)) continue;
      const int jr = r0 + lane;
      if (jr >= n) continue;
      const int j = b.order[g.atom_begin + jr];
      if (kPairs && j > i) continue;
      if (p.nparts > 1) {  // rational.f90:453 and the other families
        const int il = i - g.atom_begin, jl = j - g.atom_begin;
        if (!owns_pair(p, max(il, jl), min(il, jl))) continue;
      }
      double ej = 0.0;
      const double4 xj = b.posr[g.atom_begin + jr];  // = pos[j], read in spatial order
      const int spj = b.sp[j];
      const R rrij = R(3) * r4r2i * R(b.r4r2[spj]);
      const R4<R> q = row[jr];
      const R c6ij = q.x * q.x;
      const R dc6i = q.y * c6ij;  // dC6_ij/dCN_i
      const DampPair<R> dp = damp_setup<R, kKind>(p, rrij, b.rvdw[spj * b.nsp + spi], c6ij, idi + b.sid[j]);
      const bool swap = j > i;
      auto image = [&](int t) {
        double u0, u1, u2;
        const double r2d = pair_vec(xi, xj, b.img[g.img2_begin + t], swap, u0, u1, u2);
        if (r2d > cut2 || r2d < kEps) return;
        const R r2 = R(r2d);
        const R r1 = d_sqrt(r2);
        R sw, dswdr;
        smooth_cutoff(r1, cutr, width, sw, dswdr);
        R gdisp0, dedc6;
        const R edisp0 = damp_term<R, kGrad, kKind>(p, dp, r2, r1, gdisp0, dedc6);
        const R edisp = sw * edisp0;
        const double de = double(-c6ij * edisp * R(0.5));
        acc_e += de;
        if (kPairs) ej += de;
        if (kGrad) {
          const R gdisp = sw * gdisp0 + dswdr * edisp0 / r1;
          const R f = -c6ij * gdisp;  // dG = f * vec
          if (kKind == kZDamping) acc_d -= double(dc6i * sw * dedc6);
          else acc_d -= double(dc6i * edisp);
          const R v0 = R(u0), v1 = R(u1), v2 = R(u2);
          if (j != i) {
            g0 += double(f * v0); g1 += double(f * v1); g2 += double(f * v2);
          }
          const R h = R(0.5) * f;  // dS = 0.5 dG (x) vec
          sxx += double(h * v0 * v0); syy += double(h * v1 * v1); szz += double(h * v2 * v2);
          sxy += double(h * v0 * v1); sxz += double(h * v0 * v2); syz += double(h * v1 * v2);
        }
      };
      for_images(g, kImg2, g.img2_count, xi, xj, image);
      if (kPairs) {
        const long long il = i - g.atom_begin, jl = j - g.atom_begin;
        pair2[g.c6_begin + il * n + jl] = ej;
        pair2[g.c6_begin + jl * n + il] = ej;
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


// The image loop is specialised per damping family (the family is a
// per-system parameter, uniform within a wavefront).
template <typename R, bool kGrad, bool kPairs>
// Mixed: at most 168 VGPRs, so three wavefronts per SIMD (it otherwise
// takes 175 and two).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 3 : 1)))
void disp2_kernel(DevBatch b, const R4<R>* __restrict__ pair,
                                                       double* __restrict__ e_atom, double* __restrict__ dedcn,
                                                       double* __restrict__ grad, double* __restrict__ sig_atom,
                                                       double* __restrict__ pair2) {
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= b.natoms) return;
  const int s = b.sys[i];
  const SysGeom g = b.geom[s];
  const SysParam p = b.par[s];
  switch (p.kind) {
    case kRational: disp2_atom<R, kGrad, kPairs, kRational>(b, pair, e_atom, dedcn, grad, sig_atom, pair2, i, g, p); break;
    case kZero: disp2_atom<R, kGrad, kPairs, kZero>(b, pair, e_atom, dedcn, grad, sig_atom, pair2, i, g, p); break;
    case kModifiedZero:
      disp2_atom<R, kGrad, kPairs, kModifiedZero>(b, pair, e_atom, dedcn, grad, sig_atom, pair2, i, g, p);
      break;
    case kOptimizedPower:
      disp2_atom<R, kGrad, kPairs, kOptimizedPower>(b, pair, e_atom, dedcn, grad, sig_atom, pair2, i, g, p);
      break;
    case kCSO: disp2_atom<R, kGrad, kPairs, kCSO>(b, pair, e_atom, dedcn, grad, sig_atom, pair2, i, g, p); break;
    default: disp2_atom<R, kGrad, kPairs, kZDamping>(b, pair, e_atom, dedcn, grad, sig_atom, pair2, i, g, p); break;
  }
}

// ------------------------------------------------------ three-body lists
// L(i) = {(j, T): j <= i, eps <= |x_j + T - x_i|^2 <= cut3^2, C6_ij > 0}
// (atm.f90:706-716; pairs with a vanishing C6 -- ghost atoms -- contribute
// nothing to either branch of the reference, atm.f90:725-726), with the
// partners j in the spatial order of the system and images T in order.  The triples of atom i are the unordered pairs {a, b} of distinct
// slots of L(i) with eps <= |v_b - v_a|^2 <= cut3^2 (atm.f90:717-742): a pair
// of different atoms is the reference's triple with the larger index as j, a
// pair of two images of the same atom stands for the two triples (j, Tj, j,
// Tk) and (j, Tk, j, Tj) of the reference, which have equal contributions.
// The count pass also clears the atom's three-body sums acc (spatial order;
// null when the gradient is not wanted) for the three-body kernel that
// follows it in every evaluation.
template <bool kFill, typename R>
__global__ __launch_bounds__(kBlock) void atm_list_kernel(DevBatch b, const R4<R>* __restrict__ pair,
                                                          long long* __restrict__ count,
                                                          const long long* __restrict__ offset,
                                                          Slot<R>* __restrict__ slots, long long capacity,
                                                          double4* __restrict__ acc) {
  const int lane = wave_lane();
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= b.natoms) return;
  if (kFill && offset[b.natoms] > capacity) return;  // the lists do not fit: no three-body terms
  if (!kFill && acc && lane == 0) acc[b.geom[b.sys[i]].atom_begin + b.rank[i]] = make_double4(0.0, 0.0, 0.0, 0.0);
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
  const int n = g.natoms;
  const R4<R>* row = pair ? pair + g.c6_begin + static_cast<long long>(b.rank[i]) * n : nullptr;
  Slot<R>* out = kFill ? slots + offset[i] : nullptr;
  long long total = 0;
  unsigned long long near = 0;
  for (int r0 = 0; r0 < n; r0 += kWave) {  // partners in spatial order
    const int t = r0 / kWave;
    if ((t & (kWave - 1)) == 0) near = near_tiles(b, g, t, xi, cut2);
    if (!((near >> (t & (kWave - 1))) & 1ull)) continue;  // no slot in this tile
    const int r = r0 + lane;
    const int j = r < n ? b.order[g.atom_begin + r] : i + 1;
    int cnt = 0;
    bool valid = j <= i;
    R4<R> q{};
    if (valid && row) {
      q = row[r];
      valid = !(double(q.x) * double(q.x) < kEps);
    }
    const double4 xj = valid ? b.posr[g.atom_begin + r] : xi;  // = pos[j], read in spatial order
    if (valid) {
      for_images(g, kImg3, g.img3_count, xi, xj, [&](int t) {
        double v0, v1, v2;
        const double r2 = atm_vec(xj, b.img[g.img3_begin + t], xi, v0, v1, v2);
        if (r2 > cut2 || r2 < kEps) return;
        ++cnt;
      });
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
      for_images(g, kImg3, g.img3_count, xi, xj, [&](int t) {
        double v0, v1, v2;
        const double r2 = atm_vec(xj, b.img[g.img3_begin + t], xi, v0, v1, v2);
        if (r2 > cut2 || r2 < kEps) return;
        Slot<R> e;
        e.vx = v0; e.vy = v1; e.vz = v2;
        e.r2 = R(r2);
        e.pdmp = q.w * d_pow(R(r2), malp6);
        e.sc6h = j == i ? R(0.5) * q.x : q.x;
        e.qi = q.y;
        e.qx = q.z;
        e.atom = j;
        e.rank = r;
        out[w++] = e;
      });
    }
    total += wtot;
  }
  if (!kFill && lane == 0) count[i] = total;
}

// The list offsets and the work-item offsets of n atoms in one block:
// offset[0..n] = exclusive scan of the list sizes count[k], item_off[0..n] =
// exclusive scan of their tile counts ceil(count[k] / 64).  Each thread sums
// a chunk of consecutive atoms; the chunk sums are scanned across the
// wavefronts (wavefront scans, then a scan of the wavefront totals).
constexpr int kScanBlock = 1024;
__global__ __launch_bounds__(kScanBlock) void list_scan_kernel(const long long* __restrict__ count, int n,
                                                                long long* __restrict__ offset,
                                                                long long* __restrict__ item_off) {
  __shared__ long long wtot[2][kScanBlock / kWave];
  const int tid = threadIdx.x, lane = wave_lane(), w = tid >> 6;
  const int chunk = (n + kScanBlock - 1) / kScanBlock;
  const int lo = min(n, tid * chunk), hi = min(n, lo + chunk);
  long long s0 = 0, s1 = 0;
  for (int k = lo; k < hi; ++k) {
    const long long c = count[k];
    s0 += c;
    s1 += (c + kTile - 1) / kTile;
  }
  long long i0 = s0, i1 = s1;  // inclusive scans within the wavefront
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const long long y0 = __shfl_up(i0, off, kWave), y1 = __shfl_up(i1, off, kWave);
    if (lane >= off) {
      i0 += y0;
      i1 += y1;
    }
  }
  if (lane == kWave - 1) {
    wtot[0][w] = i0;
    wtot[1][w] = i1;
  }
  __syncthreads();
  if (w == 0) {  // exclusive scan of the wavefront totals
    constexpr int kNw = kScanBlock / kWave;
    long long t0 = lane < kNw ? wtot[0][lane] : 0, t1 = lane < kNw ? wtot[1][lane] : 0;
    const long long e0 = t0, e1 = t1;
#pragma unroll
    for (int off = 1; off < kNw; off <<= 1) {
      const long long y0 = __shfl_up(t0, off, kWave), y1 = __shfl_up(t1, off, kWave);
      if (lane >= off) {
        t0 += y0;
        t1 += y1;
      }
    }
    if (lane < kNw) {
      wtot[0][lane] = t0 - e0;
      wtot[1][lane] = t1 - e1;
    }
    if (lane == kNw - 1) {
      offset[n] = t0;
      item_off[n] = t1;
    }
  }
  __syncthreads();
  long long r0 = wtot[0][w] + i0 - s0, r1 = wtot[1][w] + i1 - s1;
  for (int k = lo; k < hi; ++k) {
    const long long c = count[k];
    offset[k] = r0;
    item_off[k] = r1;
    r0 += c;
    r1 += (c + kTile - 1) / kTile;
  }
}

// --------------------------------------------------------- three-body term
constexpr int kQueue = 128;  // entries of the accepted-pair ring of a wavefront

// Slot data of one 64-slot tile in LDS.
template <typename R>
struct TileSlots {
  R4<R> g0[kTile];                          // vx, vy, vz, r2
  R4<R> g1[kTile];                          // pdmp, sc6h, qi, qx
  int rank[kTile];                          // storage index of the slot's atom (0 past the list's end)
};
// Accepted slot pair (a, b) of a tile pair, written at test step st:
// lo = a | (a + st) << 8 (b = (a + st) mod 64), C = r_jk^2.  The term phase
// finds the rest through the two slots: the (j, k) pair record rank_j n +
// rank_k and the flag of a triple whose two slots are images of atom i.
template <typename R>
struct PairEntry {
  unsigned lo;
  R C;
};
template <>
struct PairEntry<double> {
  unsigned lo, pad;
  double C;
};

// Per-slot sums of a tile pair: G = sum dE/dv_s = sum (f_is + f_jk) v_s -
// f_jk v_partner (three components), D = sum dE (qx_s + q_partner-side).
constexpr int kAcc = 4;

// Loads slot idx of the list of atom i into position l of a tile and its
// FP64 vector into (vx, vy, vz); returns the slot's atom (-1 past the end of
// the list).  Without a tile (T == nullptr) only the vector is loaded.
template <typename R>
__device__ __forceinline__ int load_tile(TileSlots<R>* T, const Slot<R>* __restrict__ slots, long long base, int idx,
                                         int m, int l, double& vx, double& vy, double& vz) {
  if (idx < m) {
    const Slot<R> s = slots[base + idx];
    vx = s.vx; vy = s.vy; vz = s.vz;
    if (T) {
      T->g0[l] = make_r4<R>(R(s.vx), R(s.vy), R(s.vz), s.r2);
      T->g1[l] = make_r4<R>(s.pdmp, s.sc6h, s.qi, s.qx);
      T->rank[l] = s.rank;
    }
    return s.atom;
  }
  // fails every distance test
  vx = 1e30; vy = 1e30; vz = 1e30;
  if (T) {
    T->g0[l] = make_r4<R>(R(0), R(0), R(0), R(0));
    T->g1[l] = make_r4<R>(R(0), R(0), R(0), R(0));
    T->rank[l] = 0;
  }
  return -1;
}

// Segmented sums over the lanes of a wavefront: lane l holds the values of
// slot l of a tile and its atom k (k < 0: no slot).  A list holds the images
// of an atom one after another, so the slots of one atom are a run of lanes;
// on return the first lane of every run (head) holds the run's sums.  Every
// lane of the wavefront takes part.
template <int N>
__device__ __forceinline__ bool run_sums(double (&v)[N], int k) {
  const int l = wave_lane();
  const int knext = __shfl_down(k, 1, kWave);
  const unsigned long long last = __ballot(l == kWave - 1 || knext != k);
  if (~last != 0ull) {  // some run is longer than one slot
    const int e = l + __builtin_ctzll(last >> l);  // last lane of this lane's run
    for (int off = 1; off < kWave; off <<= 1) {
      double t[N];
#pragma unroll
      for (int c = 0; c < N; ++c) t[c] = __shfl_down(v[c], off, kWave);
      if (l + off <= e) {
#pragma unroll
        for (int c = 0; c < N; ++c) v[c] += t[c];
      }
    }
  }
  const int kprev = __shfl_up(k, 1, kWave);
  return l == 0 || kprev != k;
}

// Adds the slot sums of one tile (lane l: slot l with vector v, sums G and
// D, atom k at position r of its system's spatial order; k < 0 for no slot)
// to the three-body sums of the slots' atoms, acc[r] = (G, D) of the system
// (turned into gradient and dE/dCN by atm_acc_kernel), and to the i-side
// sums:  grad_i -= G,  sigma += sym(v G^T).  (The virial of a triple
// (i, a, b) with pair forces f_X v_X is f_A v_a v_a^T + f_B v_b v_b^T +
// f_C v_ab v_ab^T = v_a G_a^T + v_b G_b^T, v relative to atom i and
// G_s = dE/dv_s, so the slot sums carry it.)  The runs of one atom's images
// are summed first, so one atomic per atom and value; the sums are indexed
// by the spatial order, so the atomics of a tile touch few cache lines.
// Every lane of the wavefront calls it.
__device__ __forceinline__ void flush_tile(double Gx, double Gy, double Gz, double D, double vx, double vy,
                                           double vz, int k, int r, double4* __restrict__ acc, double* gi,
                                           double* sg) {
  if (k >= 0) {
    gi[0] -= Gx; gi[1] -= Gy; gi[2] -= Gz;
    sg[0] += vx * Gx;
    sg[1] += vy * Gy;
    sg[2] += vz * Gz;
    sg[3] += 0.5 * (vx * Gy + vy * Gx);
    sg[4] += 0.5 * (vx * Gz + vz * Gx);
    sg[5] += 0.5 * (vy * Gz + vz * Gy);
  }
  double v[4] = {Gx, Gy, Gz, D};
  const bool head = run_sums(v, k);
  if (!head || k < 0 || (v[0] == 0.0 && v[1] == 0.0 && v[2] == 0.0 && v[3] == 0.0)) return;
  atomicAdd(&acc[r].x, v[0]);
  atomicAdd(&acc[r].y, v[1]);
  atomicAdd(&acc[r].z, v[2]);
  atomicAdd(&acc[r].w, v[3]);
}

// grad += G and dE/dCN -= D / 2 of every atom from the three-body sums acc
// (spatial order, thread per position).
__global__ void atm_acc_kernel(DevBatch b, const double4* __restrict__ acc, double* __restrict__ grad,
                               double* __restrict__ dedcn) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= b.natoms) return;
  const int a = b.order[x];
  const double4 s = acc[x];
  grad[3 * a] += s.x;
  grad[3 * a + 1] += s.y;
  grad[3 * a + 2] += s.z;
  dedcn[a] -= 0.5 * s.w;
}

// Three-body hot path: x**y for x > 0, 1/x and 1/sqrt(x) for normal x > 0.
// FP32 uses the hardware approximations (v_exp_f32, v_log_f32, v_rcp_f32,
// v_rsq_f32, about 1 ulp); FP64 the device-library functions, except 1/x:
// the hardware reciprocal refined by two Newton steps (about 1 ulp, not the
// correctly rounded division).
template <typename R> __device__ __forceinline__ R pow_pos(R x, R y);
template <> __device__ __forceinline__ double pow_pos<double>(double x, double y) { return exp2(y * log2(x)); }
template <> __device__ __forceinline__ float pow_pos<float>(float x, float y) {
  return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
// x**(-8/3) for normal x > 0, the damping power of the default alp = 14
// (alp + 2 = 16, (r0/r)^((alp + 2)/3) with x = r^2): the FP32 hardware
// estimate scaled into range in FP64, then two Newton steps on
// t^6 x^16 = 1, t <- t + t (1 - t^6 x^16) / 6 (quadratic convergence from
// ~2e-7: about one ulp).  The general power costs the device library's
// log2 and exp2 in FP64.
__device__ __forceinline__ double pow_m8_3(double x) {
  const float e = (-8.0f / 3.0f) * __builtin_amdgcn_logf(static_cast<float>(x));
  const float k = __builtin_rintf(e);
  double t = ldexp(static_cast<double>(__builtin_amdgcn_exp2f(e - k)), static_cast<int>(k));
  const double x2 = x * x, x4 = x2 * x2, x8 = x4 * x4, x16 = x8 * x8;
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double t2 = t * t, t6 = t2 * t2 * t2;
    const double r = fma(-t6, x16, 1.0);
    t = fma(t * r, 1.0 / 6.0, t);
  }
  return t;
}
template <typename R> __device__ __forceinline__ R rcp_fast(R x);
template <> __device__ __forceinline__ double rcp_fast<double>(double x) {  // v_rcp_f64, two Newton steps
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  return fma(r, fma(-x, r, 1.0), r);
}
template <> __device__ __forceinline__ float rcp_fast<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <typename R> __device__ __forceinline__ R rsqrt_fast(R x);
template <> __device__ __forceinline__ double rsqrt_fast<double>(double x) { return rsqrt(x); }
template <> __device__ __forceinline__ float rsqrt_fast<float>(float x) { return __builtin_amdgcn_rsqf(x); }

// atm.f90:625-825 get_atm_dispersion_derivs (kGrad) and atm.f90:500-621,
// with rs9 = 4/3 and alp + 2 (rational.f90:563, zero.f90:602).
//
// Work item = (atom i, slot block P of L(i)); the four wavefronts of the
// block take the slot blocks Q >= P in turn, the A tile (block P) is shared.
// A tile pair runs in two phases that alternate through a ring of kQueue
// entries per wavefront:
//  * test: at step s lane l tests a = l against b = (l + s) mod 64 with the
//    reference's FP64 r_jk^2 (atm.f90:723-724) and appends accepted pairs
//    (ballot + mbcnt compaction) as (a, b, r_jk^2): the b-slot vectors
//    rotate across the lanes, the rest of the b-slot stays in LDS;
//  * terms: each lane takes one queued pair, so every lane evaluates a
//    triple (the pair record found from the two slots' ranks); its per-slot
//    sums go into double-precision LDS sums with atomics (eight per triple).
// A diagonal tile (Q == P) runs the steps 1..32 (32 only for l < 32) so that
// every unordered pair is visited once.  The sums of the b-slots are flushed
// at the end of each tile pair, those of the a-slots at the end of the item.
//
// With A = r_ij^2, B = r_ik^2, C = r_jk^2, S = ABC, t1 = A+B-C, t2 = A-B+C,
// t3 = -A+B+C, P = t1 t2 t3 the angular factor is
//   ang = S^(-5/2) (3/8 P + S),
// and 2X d(ang)/dX = S^(-5/2) (3/4 X dP/dX - 15/8 P - 3S) for X = A, B, C
// with dP/dA = t2t3 + t1t3 - t1t2 (and cyclic), the polynomials of
// atm.f90:754-760 in factored form.  The radial factor of pair X is
// f_X = -c9 fdmp S^(-5/2) (3/4 X dP/dX + K1) / X with the damping part
// K1 = -15/8 P - 3S + 2 alp pp fdmp (3/8 P + S); dG_X = f_X v_X.
// The per-slot sums G, D (see kAcc) give the gradient and virial when the
// slot is flushed (flush_tile); the weight of a triple, triple_scale times 2
// for two images of one atom (atm.f90:954-980), is 1/2 per slot that is an
// image of atom i (folded into sc6h) and 4/3 more when both are.
// kSmooth adds the switching function of a nonzero width3 (cutoff.f90:73-99,
// atm.f90:720-790).
//
// Work items are taken in reverse order (the longest slot lists first):
// block k takes item total-1-k, or, with a work counter (persistent launch),
// the items total-1-c for the counter values c it draws until none is left.
//
// kPairs (energy only) distributes every triple term over the pairs of its
// atoms, the pairwise three-body energy of atm.f90:829-950: a sixth of the
// term, t, goes to each ordered pair, i.e. t to each of the three unordered
// pairs {i, j}, {i, k}, {j, k} in both orders.  The pair sums are collected
// in pair3s, an n x n block per system in storage order (rank_x, rank_y),
// each unordered pair in one of its two orders: {j, k} at the queue entry's
// pair-record index, {i, s} per slot s at (rank_i, rank_s) when the slot is
// flushed.  The pair energy of (x, y) is then pair3s(rx, ry) + pair3s(ry, rx)
// (pair3_finalize_kernel).
constexpr int kAccPairs = 1;  // per-slot sums of the pairwise mode: sum t

// Lane l receives the value of lane (l + 1) mod 64 (DPP wave_rol:1; every
// lane has a source, so the move's old value is never used).
__device__ __forceinline__ int wave_rotate1(int x) { return __builtin_amdgcn_update_dpp(x, x, 0x134, 0xF, 0xF, false); }
__device__ __forceinline__ double wave_rotate1(double v) {
  return __hiloint2double(wave_rotate1(__double2hiint(v)), wave_rotate1(__double2loint(v)));
}

template <typename R, bool kGrad, bool kSmooth, int kW, bool kPairs, bool kPart = false>
__device__ __forceinline__ void atm_body(DevBatch b, const long long* __restrict__ item_off,
                                         const long long* __restrict__ offset, const Slot<R>* __restrict__ slots,
                                         const R4<R>* __restrict__ pair, double* __restrict__ e_atom,
                                         double* __restrict__ dedcn, double* __restrict__ grad,
                                         double* __restrict__ sig_atom, double4* __restrict__ slot_acc,
                                         unsigned long long* __restrict__ work_counter, double* __restrict__ pair3s) {
  static_assert(!(kGrad && kPairs), "the pairwise mode is energy only");
  constexpr int kNAcc = kGrad ? kAcc : (kPairs ? kAccPairs : 0);
  constexpr int kAccLen = kNAcc > 0 ? kNAcc * kTile : 1;
  __shared__ TileSlots<R> tileA;
  __shared__ TileSlots<R> tileB[kW];
  // per-slot sums in double (FP32 LDS atomics, ds_add_f32, measured five
  // times slower than FP64 ones in this kernel): the A tile's for the whole
  // item, each wavefront's B tile for one tile pair
  __shared__ double accA[kAccLen];
  __shared__ double accBw[kW][kAccLen];
  __shared__ PairEntry<R> ring[kW][kQueue];
  __shared__ double red[kW];
  __shared__ long long next_item;
  // the atoms of the A tile and of each wavefront's B tile (work partition only)
  __shared__ int part_atom[kPart ? (kW + 1) * kTile : 1];
  const long long total = item_off[b.natoms];
  long long k_item = blockIdx.x;
  if (work_counter) {
    if (threadIdx.x == 0) next_item = static_cast<long long>(atomicAdd(work_counter, 1ull));
    __syncthreads();
    k_item = next_item;
  }
  for (; k_item < total;) {
  const long long item = total - 1 - k_item;
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
  const R alp2 = R(2) * alp;
  const R malp6 = -alp / R(6);
  const bool alp14 = sizeof(R) == 8 && p.alp == 14.0;  // FP64 with the default alp: pow_m8_3
  const R ns9 = -R(p.s9);
  const R cutr = R(p.cut3), width = R(p.width3);
  const R4<R>* __restrict__ prow = pair + g.c6_begin;  // n x n pair records of the system
  // wavefront index as a uniform value, so the tile loop, the ring counters
  // and the LDS tile addresses live in scalar registers
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = wave_lane();

  // A tile (shared)
  double* __restrict__ accB = accBw[w];
  int jat = -1;  // atom of a-slot l
  double ax, ay, az;  // FP64 vector of a-slot l (every wavefront holds them)
  if (w == 0) {
    jat = load_tile(&tileA, slots, L0, P * kTile + l, m, l, ax, ay, az);
    if (kPart) part_atom[l] = jat < 0 ? i : jat;
    for (int c = 0; c < kNAcc; ++c) accA[c * kTile + l] = 0.0;
  } else {
    load_tile<R>(nullptr, slots, L0, P * kTile + l, m, l, ax, ay, az);
  }
  for (int c = 0; c < kNAcc; ++c) accB[c * kTile + l] = 0.0;
  __syncthreads();
  // Slots past the end of the list hold the coordinates 1e30, so every test
  // with one of them fails (r^2 ~ 1e60, or 0 < eps for two of them) and the
  // test needs no validity flag.  Step 32 of a diagonal tile pairs a with
  // a + 32 and a + 32 with a: only lanes l < 32 take it (cutoff -1 otherwise),
  // in a step of its own after the loop over the others.
  const double cut2_half = l < kTile / 2 ? cut2 : -1.0;
  const int arank = tileA.rank[l];
  const int ri = b.rank[i];  // storage index of atom i: slots with this rank are its images
  const unsigned lane_tag = static_cast<unsigned>(l) * 257u;  // entry word a | a << 8
  const long long irow = kPairs ? static_cast<long long>(ri) * n : 0;  // pairs {i, s} in pair3s

  double ei = 0.0, di = 0.0;
  double gi[3] = {0.0, 0.0, 0.0};
  double sg[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  PairEntry<R>* __restrict__ q_ring = ring[w];

  for (int Q = P + w; Q < nb; Q += kW) {
    const bool diag = Q == P;
    TileSlots<R>& TB = diag ? tileA : tileB[w];
    double* __restrict__ accb = diag ? accA : accB;  // b-slots of a diagonal tile are a-slots
    int katl = -1;                          // atom of b-slot l of this tile
    double vbx = 0.0, vby = 0.0, vbz = 0.0;  // and its FP64 vector
    if (!diag) katl = load_tile(&TB, slots, L0, Q * kTile + l, m, l, vbx, vby, vbz);
    if (kPart && !diag) part_atom[(w + 1) * kTile + l] = katl < 0 ? i : katl;
    const int* __restrict__ batom = part_atom + (kPart && !diag ? (w + 1) * kTile : 0);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const int s0 = diag ? 1 : 0, s1 = diag ? kTile / 2 : kTile - 1;
    int st = s0, head = 0, tail = 0;
    // the b-slot of lane l at step st is (l + st) mod 64: its vector rotates
    // by one lane after every step
    double bx = vbx, by = vby, bz = vbz;
    if (diag) {
      bx = wave_rotate1(ax);
      by = wave_rotate1(ay);
      bz = wave_rotate1(az);
    }
    // one test step: lane l tests a = l against its b-slot with the cutoff
    // thr, appends an accepted pair to the ring, and the b-slots rotate
    auto test_step = [&](double thr) {
      const int bl = (l + st) & (kTile - 1);
      double v0, v1, v2;
      const double Cd = diff_norm2(bx, by, bz, ax, ay, az, v0, v1, v2);
      const bool in_c = Cd <= thr, in_e = Cd >= kEps;
      bool ok = in_c & in_e;
      // the ballot from the two compares' lane masks (no VALU copy of ok)
      unsigned long long bal = __builtin_amdgcn_ballot_w64(in_c) & __builtin_amdgcn_ballot_w64(in_e);
      if (kPart) {  // atm.f90:709: the triple belongs to the pair (i, larger partner)
        ok = ok && owns_pair(p, i - g.atom_begin, max(part_atom[l], batom[bl]) - g.atom_begin);
        bal = __ballot(ok);
      }
      if (ok) {
        const unsigned pos = __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(bal >> 32),
                                                       __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(bal),
                                                                                 static_cast<unsigned>(tail)));
        PairEntry<R> e;
        e.lo = lane_tag + (static_cast<unsigned>(st) << 8);
        e.C = R(Cd);
        q_ring[pos & (kQueue - 1)] = e;
      }
      tail += __popcll(bal);
      bx = wave_rotate1(bx);
      by = wave_rotate1(by);
      bz = wave_rotate1(bz);
      ++st;
    };
    // the steps before the half step of a diagonal tile, or all 64
    const int s_full = diag ? kTile / 2 - 1 : kTile - 1;
    while (true) {
      // test phase: append the accepted pairs of the next steps
      while (st <= s_full && tail - head <= kQueue - kWave) test_step(cut2);
      if (diag && st == kTile / 2 && tail - head <= kQueue - kWave) test_step(cut2_half);
      const bool done = st > s1;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      __builtin_amdgcn_wave_barrier();
      // term phase: one queued pair per lane, 64 entries at a time
      for (; tail - head >= kWave || (done && tail > head); head += kWave) {
        if (head + l < tail) {
          const PairEntry<R> e = q_ring[(head + l) & (kQueue - 1)];
          const int a = e.lo & 0xff, bb = (e.lo >> 8) & (kTile - 1);
          const int ka = tileA.rank[a], kb = TB.rank[bb];
          const unsigned long long rec =
              static_cast<unsigned long long>(static_cast<unsigned>(ka)) * static_cast<unsigned>(n) +
              static_cast<unsigned>(kb);
          const R4<R> q = prow[rec];
          const R C = e.C;
          const R4<R> a0 = tileA.g0[a], a1 = tileA.g1[a];
          const R4<R> b0 = TB.g0[bb], b1 = TB.g1[bb];
          const R A = a0.w, B = b0.w;
          const R AB = A * B;
          const R S = AB * C;
          const R ir1 = rsqrt_fast(S);
          const R ir2 = ir1 * ir1;
          const R ir5 = ir2 * ir2 * ir1;
          const R dab = A - B;
          const R t1 = (A + B) - C, t2 = dab + C, t3 = C - dab;
          const R u = t2 * t3, v = t1 * t3, wv = t1 * t2;
          const R Pt = t1 * u;
          const R core = R(0.375) * Pt + S;
          const R ang = ir5 * core;
          const R pp = a1.x * b1.x * (q.w * (alp14 ? R(pow_m8_3(double(C))) : pow_pos(C, malp6)));
          const R fd = rcp_fast(R(1) + R(6) * pp);
          R c9 = ns9 * a1.y * b1.y * q.x;
          if (ka == ri && kb == ri) c9 *= R(4.0 / 3.0);  // j == k == i
          const R cwf = c9 * fd;
          R dE = ang * cwf;
          R sw = R(1), swa = R(1), dswa = R(0), swb = R(1), dswb = R(0), swc = R(1), dswc = R(0);
          if (kSmooth) {
            const R ra = d_sqrt(A), rb = d_sqrt(B), rc = d_sqrt(C);
            smooth_cutoff(ra, cutr, width, swa, dswa);
            smooth_cutoff(rb, cutr, width, swb, dswb);
            smooth_cutoff(rc, cutr, width, swc, dswc);
            dswa /= ra; dswb /= rb; dswc /= rc;
            sw = swa * swb * swc;
          }
          if (kGrad) {
            const R Y = alp2 * pp * fd;
            const R sgm = u + v + wv;
            const R K1 = R(-1.875) * Pt - R(3) * S + Y * core;
            const R hA = R(0.75) * A * (sgm - R(2) * wv) + K1;
            const R hB = R(0.75) * B * (sgm - R(2) * v) + K1;
            const R hC = R(0.75) * C * (sgm - R(2) * u) + K1;
            const R om = -cwf * ir5 * ir2;  // 1/X = (S/X) / S
            R fA = om * hA * (B * C);
            R fB = om * hB * (A * C);
            R fC = om * hC * AB;
            if (kSmooth) {  // atm.f90:754-790, switching-function terms
              fA = sw * fA - dE * dswa * swb * swc;
              fB = sw * fB - dE * dswb * swa * swc;
              fC = sw * fC - dE * dswc * swa * swb;
              dE *= sw;
            }
            di += double(dE * (a1.z + b1.z));
            // dE/dv_a = fA v_a + fC (v_a - v_b), dE/dv_b = fB v_b + fC (v_b - v_a)
            const R cx = fC * (a0.x - b0.x), cy = fC * (a0.y - b0.y), cz = fC * (a0.z - b0.z);
            atomicAdd(&accA[0 * kTile + a], double(fA * a0.x + cx));
            atomicAdd(&accA[1 * kTile + a], double(fA * a0.y + cy));
            atomicAdd(&accA[2 * kTile + a], double(fA * a0.z + cz));
            atomicAdd(&accA[3 * kTile + a], double(dE * (a1.w + q.y)));
            atomicAdd(&accb[0 * kTile + bb], double(fB * b0.x - cx));
            atomicAdd(&accb[1 * kTile + bb], double(fB * b0.y - cy));
            atomicAdd(&accb[2 * kTile + bb], double(fB * b0.z - cz));
            atomicAdd(&accb[3 * kTile + bb], double(dE * (b1.w + q.z)));
          } else if (kSmooth) {
            dE *= sw;
          }
          ei -= double(dE);
          if (kPairs) {
            const double t = double(dE) * (-1.0 / 6.0);
            atomicAdd(&accA[a], t);
            atomicAdd(&accb[bb], t);
            atomicAdd(pair3s + g.c6_begin + rec, t);
          }
        }
      }
      // the next test phase may overwrite the ring entries read above
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      __builtin_amdgcn_wave_barrier();
      if (done) break;
    }
    if (kGrad && !diag) {  // flush the b-slots of this tile pair
      const double Gx = accB[0 * kTile + l], Gy = accB[1 * kTile + l], Gz = accB[2 * kTile + l];
      const double D = accB[3 * kTile + l];
      for (int c = 0; c < kAcc; ++c) accB[c * kTile + l] = 0.0;
      flush_tile(Gx, Gy, Gz, D, vbx, vby, vbz, katl, TB.rank[l], slot_acc + g.atom_begin, gi, sg);
    }
    if (kPairs && !diag) {  // pair {i, b-slot atom}
      double t[1] = {accB[l]};
      accB[l] = 0.0;
      if (run_sums(t, katl) && katl >= 0 && t[0] != 0.0) atomicAdd(pair3s + g.c6_begin + irow + TB.rank[l], t[0]);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  if (kGrad && w == 0)
    flush_tile(accA[0 * kTile + l], accA[1 * kTile + l], accA[2 * kTile + l], accA[3 * kTile + l], ax, ay, az,
               jat, arank, slot_acc + g.atom_begin, gi, sg);
  if (kPairs && w == 0) {  // pair {i, a-slot atom}
    double t[1] = {accA[l]};
    if (run_sums(t, jat) && jat >= 0 && t[0] != 0.0) atomicAdd(pair3s + g.c6_begin + irow + arank, t[0]);
  }
  block_add(e_atom + i, ei, red);
  if (kGrad) {
    block_add(grad + 3 * i, gi[0], red);
    block_add(grad + 3 * i + 1, gi[1], red);
    block_add(grad + 3 * i + 2, gi[2], red);
    block_add(dedcn + i, -0.5 * di, red);
    for (int c = 0; c < 6; ++c) block_add(sig_atom + 6 * i + c, sg[c], red);
  }
  if (!work_counter) break;
  // block_add ended with a barrier: every wavefront is done with the LDS
  if (threadIdx.x == 0) next_item = static_cast<long long>(atomicAdd(work_counter, 1ull));
  __syncthreads();
  k_item = next_item;
  __syncthreads();  // next_item is rewritten only after every thread has read it
  }
}

// At least three wavefronts per SIMD in FP64 (the E+grad instantiation
// otherwise takes 184 VGPRs, two wavefronts, and runs 14 % slower; with the
// hint it takes 168 and spills a few dozen bytes) and four in FP32.
template <typename R, bool kGrad, bool kSmooth, int kW, bool kPart>
__global__ __launch_bounds__(kW * kWave) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 3))) void atm_kernel(DevBatch b, const long long* __restrict__ item_off,
                                                         const long long* __restrict__ offset,
                                                         const Slot<R>* __restrict__ slots,
                                                         const R4<R>* __restrict__ pair, double* __restrict__ e_atom,
                                                         double* __restrict__ dedcn, double* __restrict__ grad,
                                                         double* __restrict__ sig_atom,
                                                         double4* __restrict__ slot_acc,
                                                         unsigned long long* __restrict__ work_counter) {
  atm_body<R, kGrad, kSmooth, kW, false, kPart>(b, item_off, offset, slots, pair, e_atom, dedcn, grad, sig_atom,
                                                slot_acc, work_counter, nullptr);
}

// Pairwise three-body energies (one block of four wavefronts per work item).
template <typename R, bool kSmooth>
__global__ __launch_bounds__(4 * kWave) void atm_pairs_kernel(DevBatch b, const long long* __restrict__ item_off,
                                                              const long long* __restrict__ offset,
                                                              const Slot<R>* __restrict__ slots,
                                                              const R4<R>* __restrict__ pair,
                                                              double* __restrict__ e_atom,
                                                              double* __restrict__ pair3s) {
  atm_body<R, false, kSmooth, 4, true>(b, item_off, offset, slots, pair, e_atom, nullptr, nullptr, nullptr, nullptr,
                                       nullptr, pair3s);
}

// Pair energies in atom order from the storage-order sums of atm_pairs_kernel:
// out[c6_begin + x * n + y] = pair3s(rx, ry) + pair3s(ry, rx), one wavefront
// per atom x (row x of its system's block).
__global__ __launch_bounds__(kBlock) void pair3_finalize_kernel(DevBatch b, const double* __restrict__ pair3s,
                                                                double* __restrict__ out) {
  const int x = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (x >= b.natoms) return;
  const SysGeom g = b.geom[b.sys[x]];
  const int n = g.natoms;
  const long long rx = b.rank[x];
  const double* __restrict__ S = pair3s + g.c6_begin;
  double* __restrict__ row = out + g.c6_begin + static_cast<long long>(x - g.atom_begin) * n;
  for (int y = wave_lane(); y < n; y += kWave) {
    const long long ry = b.rank[g.atom_begin + y];
    row[y] = S[rx * n + ry] + S[ry * n + rx];
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
  auto partner = [&](int j, const double4& xj) {
    const R rc = rci + R(b.rcov[b.sp[j]]);
    const bool swap = j > i;
    // the reference adds ds*(dEdcn_i + dEdcn_j) once per unordered pair and
    // ds*dEdcn_i once per self-image (ncoord.f90:214-216)
    const R de = R(dei + dedcn[j]);
    const R wsig = (j == i) ? R(dei) : R(0.5) * de;
    for_images(g, kImgCn, g.img_cn_count, xi, xj, [&](int t) {
      double u0, u1, u2;
      const double r2d = pair_vec(xi, xj, b.img[g.img_cn_begin + t], swap, u0, u1, u2);
      if (r2d > cut2 || r2d < kCnFloor) return;
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
    });
  };
  if (g.cull) {  // partners in spatial order, tiles beyond the cutoff skipped (as cn_kernel)
    unsigned long long near = 0;
    for (int r0 = 0; r0 < g.natoms; r0 += kWave) {
      const int t = r0 / kWave;
      if ((t & (kWave - 1)) == 0) near = near_tiles(b, g, t, xi, cut2);
      if (!((near >> (t & (kWave - 1))) & 1ull) || r0 + lane >= g.natoms) continue;
      const int x = g.atom_begin + r0 + lane;
      partner(b.order[x], b.posr[x]);
    }
  } else {
    for (int j = g.atom_begin + lane; j < g.atom_begin + g.natoms; j += kWave) partner(j, b.pos[j]);
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

// ----------------------------------------------------------------- Hessian
// disp.f90:236-308 get_dispersion_hessian, FP64.  Per system (atom order
// within the system, offsets from c6_begin = sum of n^2 of the systems
// before):
//   c6, dc6, d2c6, d2c6ij  n x n at c6_begin       [i][j]: C6_ij and its derivatives
//                                                  d/dCN_i, d2/dCN_i^2, d2/dCN_i dCN_j
//   dr, dEdcndr            3n x n at 3 c6_begin    [3i+c][k]: dCN_k/dR_ic, d2E/dR_ic dCN_k
//   dEdcndcn               n x n at c6_begin       d2E/dCN_k dCN_l
//   tmp                    n x 3n at 3 c6_begin    dEdcndcn dr^T
//   hess                   3n x 3n at 9 c6_begin   [3i+a][3j+b]
// The two-body, CN and dCN/dR kernels own the rows of their atom (ordered
// pairs (i, j, T) with u = R_i - R_j - T; every pair term is even in u or
// changes sign with it), so they need no atomics; the three-body kernel adds
// with atomics.

// model.f90:201-287 weight_references with first and second derivatives.
__global__ void hess_weights_kernel(DevBatch b, const double* __restrict__ cn, double* __restrict__ gw,
                                    double* __restrict__ gwd, double* __restrict__ gwd2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= b.natoms) return;
  const double wf2 = 8.0;
  const int s = b.sp[i];
  const int nr = b.nref[s];
  const double* rcn = b.refcn + s * kMaxRef;
  const double c = cn[i];
  double gwt[kMaxRef];
  double norm = 0.0, dnorm = 0.0, d2norm = 0.0, maxcn = -1.0e300;
  for (int r = 0; r < nr; ++r) {
    const double d = c - rcn[r];
    const double w = exp(-4.0 * (d * d));
    const double dcn = rcn[r] - c;
    gwt[r] = w;
    norm += w;
    dnorm += wf2 * dcn * w;
    d2norm += (wf2 * wf2 * dcn * dcn - wf2) * w;
    maxcn = fmax(maxcn, rcn[r]);
  }
  norm = 1.0 / norm;
  for (int r = 0; r < b.mref; ++r) {
    double gwk = 0.0, dgwk = 0.0, d2gwk = 0.0;
    if (r < nr) {
      const double expw = gwt[r], dcn = rcn[r] - c;
      const double expd = wf2 * dcn * expw, expdd = (wf2 * wf2 * dcn * dcn - wf2) * expw;
      gwk = expw * norm;
      if (isnan(gwk) || isinf(gwk)) gwk = (maxcn == rcn[r]) ? 1.0 : 0.0;
      dgwk = expd * norm - expw * dnorm * (norm * norm);
      if (isnan(dgwk) || isinf(dgwk)) dgwk = 0.0;
      d2gwk = expdd * norm - 2.0 * expd * dnorm * (norm * norm) - expw * d2norm * (norm * norm) +
              2.0 * expw * dnorm * dnorm * (norm * norm * norm);
      if (isnan(d2gwk) || isinf(d2gwk)) d2gwk = 0.0;
    }
    gw[i * b.mref + r] = gwk;
    gwd[i * b.mref + r] = dgwk;
    gwd2[i * b.mref + r] = d2gwk;
  }
}

// model.f90:384-428 get_atomic_c6, second-derivative branch: block per atom i,
// threads over the atoms j of its system.
__global__ __launch_bounds__(kBlock) void hess_c6_kernel(DevBatch b, const double* __restrict__ gw,
                                                         const double* __restrict__ gwd,
                                                         const double* __restrict__ gwd2, double* __restrict__ c6,
                                                         double* __restrict__ dc6, double* __restrict__ d2c6,
                                                         double* __restrict__ d2c6ij) {
  const int i = blockIdx.x;
  const SysGeom g = b.geom[b.sys[i]];
  const int n = g.natoms, il = i - g.atom_begin;
  const int spi = b.sp[i], nri = b.nref[spi], mref = b.mref;
  const bool gi = b.ghost && b.ghost[i];
  for (int jl = threadIdx.x; jl < n; jl += blockDim.x) {
    const int j = g.atom_begin + jl;
    const int spj = b.sp[j], nrj = b.nref[spj];
    double a = 0.0, di = 0.0, ddi = 0.0, mix = 0.0;
    if (!(gi || (b.ghost && b.ghost[j]))) {
      const double* ref = b.c6ref + (spi * b.nsp + spj) * (kMaxRef * kMaxRef);
      for (int ri = 0; ri < nri; ++ri)
        for (int rj = 0; rj < nrj; ++rj) {
          const double r = ref[ri * kMaxRef + rj];
          const double wj = gw[j * mref + rj];
          a += gw[i * mref + ri] * wj * r;
          di += gwd[i * mref + ri] * wj * r;
          ddi += gwd2[i * mref + ri] * wj * r;
          mix += gwd[i * mref + ri] * gwd[j * mref + rj] * r;
        }
    }
    const long long o = g.c6_begin + static_cast<long long>(il) * n + jl;
    c6[o] = a;
    dc6[o] = di;
    d2c6[o] = ddi;
    d2c6ij[o] = mix;
  }
}

// dCN_k/dR_ic of the exponential count of ncoord.f90:118-133 (wavefront per
// atom i, lanes over j != i): the pair (i, j, T) adds g = dcf/dr u/r to the
// columns i and j of the rows of i.
__global__ __launch_bounds__(kBlock) void hess_dcndr_kernel(DevBatch b, double* __restrict__ dr) {
  const int lane = wave_lane();
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= b.natoms) return;
  const SysGeom g = b.geom[b.sys[i]];
  const int n = g.natoms, il = i - g.atom_begin;
  const double cut = b.par[b.sys[i]].cut_cn, cut2 = cut * cut;
  const double4 xi = b.pos[i];
  const double rci = b.rcov[b.sp[i]];
  double* __restrict__ rows = dr + 3 * g.c6_begin + static_cast<long long>(3 * il) * n;
  double si0 = 0.0, si1 = 0.0, si2 = 0.0;
  for (int jl = lane; jl < n; jl += kWave) {
    const int j = g.atom_begin + jl;
    if (j == i) continue;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    const double4 xj = b.pos[j];
    const double rc = rci + b.rcov[b.sp[j]];
    for (int t = 0; t < g.img_cn_count; ++t) {
      double u0, u1, u2;
      const double r2 = pair_vec(xi, xj, b.img[g.img_cn_begin + t], j > i, u0, u1, u2);
      if (r2 > cut2 || r2 < kCnFloor) continue;
      const double r1 = sqrt(r2);
      const double cf = 1.0 / (1.0 + exp(-kKcn * (rc / r1 - 1.0)));
      const double f = -cf * (1.0 - cf) * kKcn * rc / r2 / r1;
      s0 += f * u0; s1 += f * u1; s2 += f * u2;
    }
    rows[jl] = s0;
    rows[n + jl] = s1;
    rows[2 * n + jl] = s2;
    si0 += s0; si1 += s1; si2 += s2;
  }
  si0 = wave_sum(si0); si1 = wave_sum(si1); si2 = wave_sum(si2);
  if (lane == 0) {
    rows[il] = si0;
    rows[n + il] = si1;
    rows[2 * n + il] = si2;
  }
}

// get_damping_kernel of every family (rational.f90, zero.f90, mzero.f90,
// optimizedpower.f90, cso.f90, z.f90): e(u, C6) for u = r^2 and e_u, e_c,
// e_uu, e_uc, e_cc.
__device__ void damping_kernel_u(const SysParam& p, double rrij, double rvdw, int idsum, double u, double c6,
                                 double& e, double& er, double& ec, double& err, double& erc, double& ecc) {
  e = er = ec = err = erc = ecc = 0.0;
  if (p.kind != kCSO && fabs(p.s6) < kEps && fabs(p.s8) < kEps) return;
  double pv = 0.0, pd = 0.0, pdd = 0.0;
  switch (p.kind) {
    case kRational: {
      const double r0 = p.a1 * sqrt(rrij) + p.a2, r02 = r0 * r0, r06 = r02 * r02 * r02, r08 = r06 * r02;
      const double t6 = 1.0 / (u * u * u + r06), t8 = 1.0 / (u * u * u * u + r08);
      const double t6d = -3.0 * u * u * t6 * t6, t8d = -4.0 * u * u * u * t8 * t8;
      const double t6dd = -6.0 * u * t6 * t6 + 18.0 * pow(u, 4) * pow(t6, 3);
      const double t8dd = -12.0 * u * u * t8 * t8 + 32.0 * pow(u, 6) * pow(t8, 3);
      pv = p.s6 * t6 + p.s8 * rrij * t8;
      pd = p.s6 * t6d + p.s8 * rrij * t8d;
      pdd = p.s6 * t6dd + p.s8 * rrij * t8dd;
      break;
    }
    case kZero:
    case kModifiedZero: {
      const double alp6 = p.alp, alp8 = p.alp + 2.0;
      double t6, t8, t6d, t8d, t6dd, t8dd;
      if (p.kind == kZero) {
        const double a6 = 0.5 * alp6, a8 = 0.5 * alp8;
        t6 = pow(p.a1 * rvdw / sqrt(u), alp6);
        t8 = pow(p.a2 * rvdw / sqrt(u), alp8);
        t6d = -a6 * t6 / u;
        t8d = -a8 * t8 / u;
        t6dd = a6 * (a6 + 1.0) * t6 / (u * u);
        t8dd = a8 * (a8 + 1.0) * t8 / (u * u);
      } else {
        const double r1 = sqrt(u), betr0 = p.bet * rvdw;
        const double w6 = r1 / (p.a1 * rvdw) + betr0, w8 = r1 / (p.a2 * rvdw) + betr0;
        const double w6d = 1.0 / (p.a1 * rvdw * 2.0 * r1), w8d = 1.0 / (p.a2 * rvdw * 2.0 * r1);
        const double w6dd = -1.0 / (p.a1 * rvdw * 4.0 * u * r1), w8dd = -1.0 / (p.a2 * rvdw * 4.0 * u * r1);
        t6 = pow(w6, -alp6);
        t8 = pow(w8, -alp8);
        t6d = -alp6 * pow(w6, -alp6 - 1.0) * w6d;
        t8d = -alp8 * pow(w8, -alp8 - 1.0) * w8d;
        t6dd = alp6 * (alp6 + 1.0) * pow(w6, -alp6 - 2.0) * w6d * w6d - alp6 * pow(w6, -alp6 - 1.0) * w6dd;
        t8dd = alp8 * (alp8 + 1.0) * pow(w8, -alp8 - 2.0) * w8d * w8d - alp8 * pow(w8, -alp8 - 1.0) * w8dd;
      }
      const double f6 = 1.0 / (1.0 + 6.0 * t6), f8 = 1.0 / (1.0 + 6.0 * t8);
      const double f6d = -6.0 * t6d * f6 * f6, f8d = -6.0 * t8d * f8 * f8;
      const double f6dd = -6.0 * t6dd * f6 * f6 + 72.0 * t6d * t6d * pow(f6, 3);
      const double f8dd = -6.0 * t8dd * f8 * f8 + 72.0 * t8d * t8d * pow(f8, 3);
      const double c6t = f6 / pow(u, 3), c6td = f6d / pow(u, 3) - 3.0 * f6 / pow(u, 4);
      const double c6tdd = f6dd / pow(u, 3) - 6.0 * f6d / pow(u, 4) + 12.0 * f6 / pow(u, 5);
      const double c8t = f8 / pow(u, 4), c8td = f8d / pow(u, 4) - 4.0 * f8 / pow(u, 5);
      const double c8tdd = f8dd / pow(u, 4) - 8.0 * f8d / pow(u, 5) + 20.0 * f8 / pow(u, 6);
      pv = p.s6 * c6t + p.s8 * rrij * c8t;
      pd = p.s6 * c6td + p.s8 * rrij * c8td;
      pdd = p.s6 * c6tdd + p.s8 * rrij * c8tdd;
      break;
    }
    case kOptimizedPower: {
      const double r0 = p.a1 * sqrt(rrij) + p.a2, r02 = r0 * r0, r06 = r02 * r02 * r02, r08 = r06 * r02;
      const double ab = pow(r0, p.bet), abr06 = ab * r06, abr08 = ab * r08, q = 0.5 * p.bet;
      const double d6 = pow(u, 3) + abr06 * pow(u, -q), d8 = pow(u, 4) + abr08 * pow(u, -q);
      const double d6d = 3.0 * u * u - q * abr06 * pow(u, -q - 1.0);
      const double d8d = 4.0 * pow(u, 3) - q * abr08 * pow(u, -q - 1.0);
      const double d6dd = 6.0 * u + q * (q + 1.0) * abr06 * pow(u, -q - 2.0);
      const double d8dd = 12.0 * u * u + q * (q + 1.0) * abr08 * pow(u, -q - 2.0);
      const double t6 = 1.0 / d6, t8 = 1.0 / d8;
      const double t6d = -d6d * t6 * t6, t8d = -d8d * t8 * t8;
      const double t6dd = -d6dd * t6 * t6 + 2.0 * d6d * d6d * pow(t6, 3);
      const double t8dd = -d8dd * t8 * t8 + 2.0 * d8d * d8d * pow(t8, 3);
      pv = p.s6 * t6 + p.s8 * rrij * t8;
      pd = p.s6 * t6d + p.s8 * rrij * t8d;
      pdd = p.s6 * t6dd + p.s8 * rrij * t8dd;
      break;
    }
    case kCSO: {
      const double r1 = sqrt(u), r0 = sqrt(rrij);
      const double dd6 = p.a3 * r0 + p.a4, dd6_6 = pow(dd6, 6);
      const double sf = 1.0 / (1.0 + exp(r1 - p.a2 * r0));
      const double sfd = -sf * (1.0 - sf), sfdd = sf * (1.0 - sf) * (1.0 - 2.0 * sf);
      const double sig = p.s6 + p.a1 * sf, sigd = p.a1 * sfd, sigdd = p.a1 * sfdd;
      const double sigu = 0.5 * sigd / r1, siguu = 0.25 * sigdd / u - 0.25 * sigd / (u * r1);
      const double t6 = 1.0 / (u * u * u + dd6_6);
      const double t6d = -3.0 * u * u * t6 * t6;
      const double t6dd = -6.0 * u * t6 * t6 + 18.0 * pow(u, 4) * pow(t6, 3);
      pv = sig * t6;
      pd = sigu * t6 + sig * t6d;
      pdd = siguu * t6 + 2.0 * sigu * t6d + sig * t6dd;
      break;
    }
    default: {  // Z damping
      const double r0 = p.a1 / static_cast<double>(idsum), b8 = r0 * rrij;
      const double t6 = 1.0 / (u * u * u + r0 * c6), t8 = 1.0 / (u * u * u * u + b8 * c6);
      const double t6u = -3.0 * u * u * t6 * t6, t8u = -4.0 * pow(u, 3) * t8 * t8;
      const double t6uu = -6.0 * u * t6 * t6 + 18.0 * pow(u, 4) * pow(t6, 3);
      const double t8uu = -12.0 * u * u * t8 * t8 + 32.0 * pow(u, 6) * pow(t8, 3);
      const double t6c = -r0 * t6 * t6, t8c = -b8 * t8 * t8;
      const double t6uc = 6.0 * u * u * r0 * pow(t6, 3), t8uc = 8.0 * pow(u, 3) * b8 * pow(t8, 3);
      const double t6cc = 2.0 * r0 * r0 * pow(t6, 3), t8cc = 2.0 * b8 * b8 * pow(t8, 3);
      const double pz = p.s6 * t6 + p.s8 * rrij * t8, pu = p.s6 * t6u + p.s8 * rrij * t8u;
      const double puu = p.s6 * t6uu + p.s8 * rrij * t8uu, pc = p.s6 * t6c + p.s8 * rrij * t8c;
      const double puc = p.s6 * t6uc + p.s8 * rrij * t8uc, pcc = p.s6 * t6cc + p.s8 * rrij * t8cc;
      e = -c6 * pz;
      er = -c6 * pu;
      ec = -pz - c6 * pc;
      err = -c6 * puu;
      erc = -pu - c6 * puc;
      ecc = -2.0 * pc - c6 * pcc;
      return;
    }
  }
  e = -c6 * pv;
  er = -c6 * pd;
  ec = -pv;
  err = -c6 * pdd;
  erc = -pd;
}

// damping.f90:393-539 get_dispersion2_hessian (wavefront per atom i, lanes
// over the atoms j of its system): the rows of atom i of the Hessian at
// fixed CN, of dEdcndr, dEdcndcn and dEdcn.  Adds to zeroed arrays.
__global__ __launch_bounds__(kBlock) void hess2_kernel(DevBatch b, const double* __restrict__ c6,
                                                       const double* __restrict__ dc6,
                                                       const double* __restrict__ d2c6,
                                                       const double* __restrict__ d2c6ij, double* __restrict__ hess,
                                                       double* __restrict__ dedcn, double* __restrict__ dEdcndr,
                                                       double* __restrict__ dEdcndcn) {
  const int lane = wave_lane();
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= b.natoms) return;
  const int s = b.sys[i];
  const SysGeom g = b.geom[s];
  const SysParam p = b.par[s];
  const int n = g.natoms, il = i - g.atom_begin, ndim = 3 * n;
  const double cut2 = p.cut2 * p.cut2;
  const double4 xi = b.pos[i];
  const int spi = b.sp[i];
  const long long c0 = g.c6_begin, ri = c0 + static_cast<long long>(il) * n;
  double* __restrict__ hrow = hess + 9 * c0 + static_cast<long long>(3 * il) * ndim;
  double* __restrict__ drow = dEdcndr + 3 * c0 + static_cast<long long>(3 * il) * n;
  double hii[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double dii[3] = {0.0, 0.0, 0.0}, e_cn = 0.0, e_cc = 0.0;
  for (int jl = lane; jl < n; jl += kWave) {
    if (!owns_pair(p, max(il, jl), min(il, jl))) continue;  // damping.f90:452
    const int j = g.atom_begin + jl;
    const int spj = b.sp[j];
    const double rrij = 3.0 * b.r4r2[spi] * b.r4r2[spj];
    const double c6ij = c6[ri + jl];
    const bool self = j == i;
    // the reference's cnd / cnd2 of the pair, i-side entries
    const double cni = self ? 2.0 * dc6[ri + il] : dc6[ri + jl];
    const double cnj = self ? 0.0 : dc6[c0 + static_cast<long long>(jl) * n + il];
    const double cnii = self ? 2.0 * d2c6[ri + il] + 2.0 * d2c6ij[ri + il] : d2c6[ri + jl];
    const double cnij = self ? 0.0 : d2c6ij[ri + jl];
    const double fac = self ? 0.5 : 1.0;
    double hb[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double dij[3] = {0.0, 0.0, 0.0}, ccij = 0.0;
    for (int t = 0; t < g.img2_count; ++t) {
      double v0, v1, v2;
      const double r2 = pair_vec(xi, b.pos[j], b.img[g.img2_begin + t], j > i, v0, v1, v2);
      if (r2 > cut2 || r2 < kEps) continue;
      double e0, e0r, e0c, e0rr, e0rc, e0cc, sw, swr, swrr;
      damping_kernel_u(p, rrij, b.rvdw[spj * b.nsp + spi], b.sid[i] + b.sid[j], r2, c6ij, e0, e0r, e0c, e0rr,
                       e0rc, e0cc);
      switch_r2(r2, p.cut2, p.width2, sw, swr, swrr);
      const double fr = fac * (swr * e0 + sw * e0r);
      const double frr = fac * (swrr * e0 + 2.0 * swr * e0r + sw * e0rr);
      const double fc = fac * sw * e0c;
      const double frc = fac * (swr * e0c + sw * e0rc);
      const double fcc = fac * sw * e0cc;
      e_cn += fc * cni;
      e_cc += fcc * cni * cni + fc * cnii;
      if (!self) {
        ccij += fcc * cni * cnj + fc * cnij;
        const double v[3] = {v0, v1, v2};
        for (int a = 0; a < 3; ++a) {
          for (int c = 0; c < 3; ++c) hb[3 * a + c] += 4.0 * frr * v[a] * v[c];
          hb[4 * a] += 2.0 * fr;
          dii[a] += frc * 2.0 * v[a] * cni;
          dij[a] += frc * 2.0 * v[a] * cnj;
        }
      }
    }
    if (!self) {
      for (int a = 0; a < 3; ++a) {
        for (int c = 0; c < 3; ++c) {
          hrow[static_cast<long long>(a) * ndim + 3 * jl + c] -= hb[3 * a + c];
          hii[3 * a + c] += hb[3 * a + c];
        }
        drow[static_cast<long long>(a) * n + jl] += dij[a];
      }
      dEdcndcn[ri + jl] += ccij;
    }
  }
  for (int k = 0; k < 9; ++k) hii[k] = wave_sum(hii[k]);
  for (int a = 0; a < 3; ++a) dii[a] = wave_sum(dii[a]);
  e_cn = wave_sum(e_cn);
  e_cc = wave_sum(e_cc);
  if (lane == 0) {
    for (int a = 0; a < 3; ++a) {
      for (int c = 0; c < 3; ++c) hrow[static_cast<long long>(a) * ndim + 3 * il + c] += hii[3 * a + c];
      drow[static_cast<long long>(a) * n + il] += dii[a];
    }
    dedcn[i] += e_cn;
    dEdcndcn[ri + il] += e_cc;
  }
}

// atm.f90:38-429 get_atm_dispersion_hessian: wavefront per (atom i, slot a
// of L(i)), lanes over the slots b after a, i.e. over the triple terms of the
// three-body lists (weights as in atm_kernel: 1/2 per slot that is an image
// of atom i, 4/3 more when both are).  Each term is K(u1, u2, u3) with
// u1 = |v_a|^2, u2 = |v_b|^2, u3 = |v_b - v_a|^2; its gradient and Hessian in
// u are contracted with du/dR.  Entries between the atoms of i and a (the
// Hessian blocks, dEdcndr, dEdcndcn and dEdcn entries every term of the
// wavefront shares) are summed in registers and added once per wavefront;
// the entries of the atom of b are added with atomics.
// Accumulation of hess3_triple for one wavefront: the entries among triple
// positions 0 and 1 (atom i and the atom of slot a, the same for every lane)
// in registers, added once per wavefront; the others with FP64 atomics.
struct Hess3DeviceSink {
  double hf_[2][2][9], df_[2][3][2], cf_[2][2], ef_[2];
  double* hess;      // the system's 3n x 3n block
  double* dedcn;     // the system's atoms
  double* dEdcndr;   // the system's 3n x n block
  double* dEdcndcn;  // the system's n x n block
  int n, ndim;
  __device__ __forceinline__ void hf(int x, int y, int k, double v) { hf_[x][y][k] += v; }
  __device__ __forceinline__ void h(int row, int col, double v) {
    atomicAdd(hess + static_cast<long long>(row) * ndim + col, v);
  }
  __device__ __forceinline__ void cf(int x, int y, double v) { cf_[x][y] += v; }
  __device__ __forceinline__ void c(int r, int col, double v) {
    atomicAdd(dEdcndcn + static_cast<long long>(r) * n + col, v);
  }
  __device__ __forceinline__ void df(int x, int c, int y, double v) { df_[x][c][y] += v; }
  __device__ __forceinline__ void d(int row, int atom, double v) {
    atomicAdd(dEdcndr + static_cast<long long>(row) * n + atom, v);
  }
  __device__ __forceinline__ void ef(int x, double v) { ef_[x] += v; }
  __device__ __forceinline__ void e(int atom, double v) { atomicAdd(dedcn + atom, v); }
};

__global__ __launch_bounds__(kBlock) void hess3_kernel(DevBatch b, const long long* __restrict__ offset,
                                                       const Slot<double>* __restrict__ slots,
                                                       const double* __restrict__ c6,
                                                       const double* __restrict__ dc6,
                                                       const double* __restrict__ d2c6,
                                                       const double* __restrict__ d2c6ij, double* __restrict__ hess,
                                                       double* __restrict__ dedcn, double* __restrict__ dEdcndr,
                                                       double* __restrict__ dEdcndcn) {
  const int lane = wave_lane();
  const long long t = static_cast<long long>(blockIdx.x) * kWavesPerBlock + (threadIdx.x >> 6);
  const long long total = offset[b.natoms];
  if (t >= total) return;
  int lo = 0, hi = b.natoms;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offset[mid] <= t) lo = mid;
    else hi = mid;
  }
  const int i = lo;
  const SysGeom g = b.geom[b.sys[i]];
  const int n = g.natoms;
  const long long c0 = g.c6_begin;
  const Hess3System sys{b.par[b.sys[i]], n, c6 + c0, dc6 + c0, d2c6 + c0, d2c6ij + c0, b.rvdw, b.nsp};
  Hess3DeviceSink sink;
#pragma unroll
  for (int x = 0; x < 2; ++x) {
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int c = 0; c < 9; ++c) sink.hf_[x][y][c] = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) sink.df_[x][c][0] = sink.df_[x][c][1] = 0.0;
    sink.cf_[x][0] = sink.cf_[x][1] = 0.0;
    sink.ef_[x] = 0.0;
  }
  sink.hess = hess + 9 * c0;
  sink.dedcn = dedcn + g.atom_begin;
  sink.dEdcndr = dEdcndr + 3 * c0;
  sink.dEdcndcn = dEdcndcn + c0;
  sink.n = n;
  sink.ndim = 3 * n;
  const Slot<double> sa = slots[t];
  const long long end = offset[i + 1];
  const int at0 = i - g.atom_begin, at1 = sa.atom - g.atom_begin;
  const int sp0 = b.sp[i], sp1 = b.sp[sa.atom];
  const double va[3] = {sa.vx, sa.vy, sa.vz};
  for (long long bi = t + 1 + lane; bi < end; bi += kWave) {
    const Slot<double> sb = slots[bi];
    const double vb[3] = {sb.vx, sb.vy, sb.vz};
    hess3_triple(sys, at0, at1, sb.atom - g.atom_begin, sp0, sp1, b.sp[sb.atom], sa.atom == i, sb.atom == i, va,
                 sa.r2, vb, sb.r2, sink);
  }
  // the entries between positions 0 and 1, once per wavefront (both refer
  // to atom i when slot a is an image of it)
  const int fat[2] = {at0, at1};
  const int ndim = 3 * n;
#pragma unroll
  for (int x = 0; x < 2; ++x) {
#pragma unroll
    for (int y = 0; y < 2; ++y) {
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        const double v = wave_sum(sink.hf_[x][y][c]);
        if (lane == 0 && v != 0.0)
          atomicAdd(hess + 9 * c0 + static_cast<long long>(3 * fat[x] + c / 3) * ndim + 3 * fat[y] + c % 3, v);
      }
      const double v = wave_sum(sink.cf_[x][y]);
      if (lane == 0 && v != 0.0) atomicAdd(dEdcndcn + c0 + static_cast<long long>(fat[x]) * n + fat[y], v);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double w = wave_sum(sink.df_[x][c][y]);
        if (lane == 0 && w != 0.0) atomicAdd(dEdcndr + 3 * c0 + static_cast<long long>(3 * fat[x] + c) * n + fat[y], w);
      }
    }
    const double v = wave_sum(sink.ef_[x]);
    if (lane == 0 && v != 0.0) atomicAdd(dedcn + g.atom_begin + fat[x], v);
  }
}

// ncoord.f90:236-306 add_coordination_number_hessian (wavefront per atom i,
// lanes over j != i): dE/dCN times the second derivatives of the CN.
__global__ __launch_bounds__(kBlock) void hess_cn_kernel(DevBatch b, const double* __restrict__ dedcn,
                                                         double* __restrict__ hess) {
  const int lane = wave_lane();
  const int i = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (i >= b.natoms) return;
  const SysGeom g = b.geom[b.sys[i]];
  const int n = g.natoms, il = i - g.atom_begin, ndim = 3 * n;
  const double cut = b.par[b.sys[i]].cut_cn, cut2 = cut * cut;
  const double4 xi = b.pos[i];
  const double rci = b.rcov[b.sp[i]];
  double* __restrict__ hrow = hess + 9 * g.c6_begin + static_cast<long long>(3 * il) * ndim;
  double hii[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int jl = lane; jl < n; jl += kWave) {
    const int j = g.atom_begin + jl;
    if (j == i) continue;
    const double4 xj = b.pos[j];
    const double rc = rci + b.rcov[b.sp[j]];
    const double dEij = dedcn[i] + dedcn[j];
    double hb[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int t = 0; t < g.img_cn_count; ++t) {
      double v0, v1, v2;
      const double r2 = pair_vec(xi, xj, b.img[g.img_cn_begin + t], j > i, v0, v1, v2);
      if (r2 > cut2 || r2 < kCnFloor) continue;
      const double r1 = sqrt(r2);
      const double cf = 1.0 / (1.0 + exp(-kKcn * (rc / r1 - 1.0)));
      const double tmp = cf * (1.0 - cf);
      const double dcf = -tmp * kKcn * rc / r2;
      const double d2cf = tmp * (1.0 - 2.0 * cf) * (kKcn * rc) * (kKcn * rc) / (r2 * r2) +
                          2.0 * tmp * kKcn * rc / (r2 * r1);
      const double v[3] = {v0, v1, v2};
      for (int a = 0; a < 3; ++a) {
        for (int c = 0; c < 3; ++c) hb[3 * a + c] += dEij * (d2cf * v[a] * v[c] / r2 - dcf * v[a] * v[c] / (r2 * r1));
        hb[4 * a] += dEij * dcf / r1;
      }
    }
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        hrow[static_cast<long long>(a) * ndim + 3 * jl + c] -= hb[3 * a + c];
        hii[3 * a + c] += hb[3 * a + c];
      }
  }
  for (int k = 0; k < 9; ++k) hii[k] = wave_sum(hii[k]);
  if (lane == 0)
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) hrow[static_cast<long long>(a) * ndim + 3 * il + c] += hii[3 * a + c];
}

// tmp[k][y] = sum_l dEdcndcn[k][l] dr[y][l] (n x 3n per system; grid.z = system).
constexpr int kHT = 16;  // tile edge of the dense products
__global__ __launch_bounds__(kHT * kHT) void hess_tmp_kernel(DevBatch b, const double* __restrict__ dEdcndcn,
                                                             const double* __restrict__ dr, double* __restrict__ tmp) {
  const SysGeom g = b.geom[blockIdx.z];
  const int n = g.natoms, ndim = 3 * n;
  const int k = blockIdx.y * kHT + threadIdx.y, y = blockIdx.x * kHT + threadIdx.x;
  __shared__ double A[kHT][kHT + 1], B[kHT][kHT + 1];
  const double* C = dEdcndcn + g.c6_begin;
  const double* D = dr + 3 * g.c6_begin;
  double acc = 0.0;
  for (int l0 = 0; l0 < n; l0 += kHT) {
    const int la = l0 + threadIdx.x;
    A[threadIdx.y][threadIdx.x] = (k < n && la < n) ? C[static_cast<long long>(k) * n + la] : 0.0;
    const int yb = blockIdx.x * kHT + threadIdx.y;
    B[threadIdx.y][threadIdx.x] = (yb < ndim && la < n) ? D[static_cast<long long>(yb) * n + la] : 0.0;
    __syncthreads();
    for (int q = 0; q < kHT; ++q) acc += A[threadIdx.y][q] * B[threadIdx.x][q];
    __syncthreads();
  }
  if (k < n && y < ndim) tmp[3 * g.c6_begin + static_cast<long long>(k) * ndim + y] = acc;
}

// hess[x][y] += sum_k dEdcndr[x][k] dr[y][k] + dr[x][k] dEdcndr[y][k] + dr[x][k] tmp[k][y]
// (disp.f90:303-306; grid.z = system).
__global__ __launch_bounds__(kHT * kHT) void hess_assemble_kernel(DevBatch b, const double* __restrict__ dEdcndr,
                                                                  const double* __restrict__ dr,
                                                                  const double* __restrict__ tmp,
                                                                  double* __restrict__ hess) {
  const SysGeom g = b.geom[blockIdx.z];
  const int n = g.natoms, ndim = 3 * n;
  const int x = blockIdx.y * kHT + threadIdx.y, y = blockIdx.x * kHT + threadIdx.x;
  __shared__ double Ax[kHT][kHT + 1], Dx[kHT][kHT + 1], Ay[kHT][kHT + 1], Dy[kHT][kHT + 1], Tk[kHT][kHT + 1];
  const double* A = dEdcndr + 3 * g.c6_begin;
  const double* D = dr + 3 * g.c6_begin;
  const double* T = tmp + 3 * g.c6_begin;
  double acc = 0.0;
  for (int k0 = 0; k0 < n; k0 += kHT) {
    const int kk = k0 + threadIdx.x;
    const int xr = blockIdx.y * kHT + threadIdx.y, yr = blockIdx.x * kHT + threadIdx.y;
    const bool kok = kk < n;
    Ax[threadIdx.y][threadIdx.x] = (xr < ndim && kok) ? A[static_cast<long long>(xr) * n + kk] : 0.0;
    Dx[threadIdx.y][threadIdx.x] = (xr < ndim && kok) ? D[static_cast<long long>(xr) * n + kk] : 0.0;
    Ay[threadIdx.y][threadIdx.x] = (yr < ndim && kok) ? A[static_cast<long long>(yr) * n + kk] : 0.0;
    Dy[threadIdx.y][threadIdx.x] = (yr < ndim && kok) ? D[static_cast<long long>(yr) * n + kk] : 0.0;
    const int kt = k0 + threadIdx.y, yt = blockIdx.x * kHT + threadIdx.x;
    Tk[threadIdx.y][threadIdx.x] = (kt < n && yt < ndim) ? T[static_cast<long long>(kt) * ndim + yt] : 0.0;
    __syncthreads();
    for (int q = 0; q < kHT; ++q)
      acc += Ax[threadIdx.y][q] * Dy[threadIdx.x][q] + Dx[threadIdx.y][q] * Ay[threadIdx.x][q] +
             Dx[threadIdx.y][q] * Tk[q][threadIdx.x];
    __syncthreads();
  }
  if (x < ndim && y < ndim) hess[9 * g.c6_begin + static_cast<long long>(x) * ndim + y] += acc;
}


// ---------------------------------------------------- reciprocal space
// fourier/ewald.f90:47-252, the two-body energy of a 3D-periodic system as a
// sum over reciprocal vectors k of the structure factors of the low-rank C6
// factors, S_la(k) = sum_{i of species a} c6l_il exp(i k.x_i).  Of each +-k
// pair one vector is visited (S(-k) = conj S(k), identical contributions)
// with weight 2, k = 0 with weight 1.  Vectors k = h b1 + k b2 + l b3 come
// in tiles of kWave consecutive l on one line (h, k); lane = vector, and
// the flat vector index is tile * kWave + lane (lanes past a tile's count
// are padding with weight 0).  The phase exp(i k.x) is the product of
// per-atom tables exp(i n b_d.x), so the sums over vectors and atoms
// evaluate no sine or cosine.

// kernel.f90:83-129 get_fourier_transform: 4 pi / k int r sin(kr) f(r) dr of
// f = pref r^m / (r^alpha + R^alpha) by the residues at the alpha/2 poles
// of the upper half plane, and its derivative in k.
__device__ void fourier_transform(const FourierTerm& t, double k, double& phi, double& dphi) {
  phi = 0.0;
  dphi = 0.0;
  if (fabs(t.pref) <= 0.0) return;
  const int nexp = t.m - t.alpha + 2;
  const double pre = 4.0 * M_PI * M_PI / t.alpha * t.pref * pow(t.radius, nexp);
  if (k <= 0.0) {
    phi = pre * t.radius / sin(M_PI * (t.m + 3) / double(t.alpha));
    return;
  }
  const double kr = k * t.radius;
  double acc = 0.0, dacc = 0.0;
  for (int p = 0; p < t.alpha / 2; ++p) {
    const double th = M_PI * (2 * p + 1) / double(t.alpha);
    double st, ct, sa, ca;
    sincos(th, &st, &ct);
    const double arg = 0.5 * M_PI + nexp * th + kr * ct;
    const double damp = exp(-kr * st);
    sincos(arg, &sa, &ca);
    acc += sa * damp;
    dacc += (ca * ct - sa * st) * damp;  // cos(arg + th)
  }
  phi = pre * acc / k;
  dphi = (pre * t.radius * dacc - phi) / k;
}

// The vector of lane `lane` of a tile, its weight (0 for padding lanes) and
// its entries of the phase tables (h, k wave-uniform; l clamped to the tile
// for padding lanes).
struct TileVec {
  double kx, ky, kz, w;
  int e1, e2, e3;
};
__device__ __forceinline__ TileVec tile_vector(const EwaldBatch& e, const int4 tile, int lane) {
  TileVec v;
  const bool valid = lane < tile.w;
  const int l = tile.z + (valid ? lane : 0);
  v.kx = e.rec[0] * tile.x + e.rec[3] * tile.y + e.rec[6] * l;
  v.ky = e.rec[1] * tile.x + e.rec[4] * tile.y + e.rec[7] * l;
  v.kz = e.rec[2] * tile.x + e.rec[5] * tile.y + e.rec[8] * l;
  const bool origin = tile.x == 0 && tile.y == 0 && l == 0;
  if (e.nparts <= 1) {
    v.w = !valid ? 0.0 : origin ? 1.0 : 2.0;
  } else if (!valid) {
    v.w = 0.0;
  } else {  // ewald.f90:161: the vector and its negative by their box indices
    const long long w1 = 2 * e.rep1 + 1, w2 = 2 * e.rep2 + 1, nbox = (2 * e.rep0 + 1) * w1 * w2;
    const long long t = ((tile.x + e.rep0) * w1 + (tile.y + e.rep1)) * w2 + (l + e.rep2);
    v.w = (t % e.nparts == e.part ? 1.0 : 0.0) + (!origin && (nbox - 1 - t) % e.nparts == e.part ? 1.0 : 0.0);
  }
  v.e1 = e.toff0 + e.rep0 + tile.x;
  v.e2 = e.toff1 + e.rep1 + tile.y;
  v.e3 = e.toff2 + e.rep2 + l;
  return v;
}

// exp(i k.x) of atom row E for the vector v: E[h] E[k] (uniform) E[l].
__device__ __forceinline__ void tile_phase(const double2* __restrict__ E, const TileVec& v, double& c, double& s) {
  const double2 p1 = E[v.e1], p2 = E[v.e2], p3 = E[v.e3];
  const double pr = p1.x * p2.x - p1.y * p2.y, pi = p1.x * p2.y + p1.y * p2.x;
  c = pr * p3.x - pi * p3.y;
  s = pr * p3.y + pi * p3.x;
}

// Phase tables, one row per atom in species order (row r for atom
// sorder[r]): etab[r][toff_d + rep_d + n] = exp(i n b_d.x), |n| <= rep_d.
__global__ __launch_bounds__(kBlock) void ewald_table_kernel(DevBatch b, EwaldBatch e, double2* __restrict__ etab) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<long long>(e.natoms) * e.tstride) return;
  const int row = static_cast<int>(t / e.tstride), q = static_cast<int>(t % e.tstride);
  const int d = q < e.toff1 ? 0 : q < e.toff2 ? 1 : 2;
  const int n = q - (d == 0 ? e.toff0 + e.rep0 : d == 1 ? e.toff1 + e.rep1 : e.toff2 + e.rep2);
  const double bx = d == 0 ? e.rec[0] : d == 1 ? e.rec[3] : e.rec[6];
  const double by = d == 0 ? e.rec[1] : d == 1 ? e.rec[4] : e.rec[7];
  const double bz = d == 0 ? e.rec[2] : d == 1 ? e.rec[5] : e.rec[8];
  const double4 x = b.pos[e.sorder[row]];
  double s, c;
  sincos(n * (bx * x.x + by * x.y + bz * x.z), &s, &c);
  etab[t] = make_double2(c, s);
}

// decomposition.f90:202-242 get_weights: c6l_il = sum_r vec(sp_i, r, l) w_ir
// (and dc6l with dw/dCN); zero rows for ghost atoms.  Stored by the atom's
// position r in species order: c6l[r][l] (rows of rank_pad, zero padded)
// for the structure factors, lambda_l c6l and lambda_l dc6l as lc6l[l][r]
// and ldc6l[l][r] (rows of natoms + kEwaldGroup) for the per-atom pass.  The
// self-interaction of the reciprocal sum (ewald.f90:131-147), 1/2 sum_l
// lambda_l c6l_il^2 phi_aa(0) and its CN derivative, is added to the atom's
// energy here.  One thread per atom of the system.
__global__ __launch_bounds__(kBlock) void lowrank_kernel(DevBatch b, EwaldBatch e, const double* __restrict__ gw,
                                                         const double* __restrict__ dgw, double* __restrict__ c6l,
                                                         double* __restrict__ lc6l, double* __restrict__ ldc6l,
                                                         double* __restrict__ e_atom, double* __restrict__ dedcn) {
  const int il = blockIdx.x * blockDim.x + threadIdx.x;
  if (il >= e.natoms) return;
  const int i = e.atom_begin + il;
  const int row = e.spos[il];
  const long long ld = e.natoms + kEwaldGroup;
  const int sp = b.sp[i];
  const int nr = b.nref[sp];
  const bool ghost = b.ghost && b.ghost[i];
  double es = 0.0, ds = 0.0;
  for (int l = 0; l < e.rank_pad; ++l) {
    double a = 0.0, d = 0.0;
    if (!ghost && l < e.rank) {
      for (int r = 0; r < nr; ++r) {
        const double v = e.vec[(sp * kMaxRef + r) * e.rank + l];
        a += v * gw[i * b.mref + r];
        if (dgw) d += v * dgw[i * b.mref + r];
      }
    }
    c6l[static_cast<long long>(row) * e.rank_pad + l] = a;
    if (l >= e.rank) continue;
    lc6l[l * ld + row] = e.lambda[l] * a;
    if (dgw) ldc6l[l * ld + row] = e.lambda[l] * d;
    // partitioned: the mesh keeps the terms l of its part (spme.f90:166), the
    // direct sum the whole self-interaction on part 0 (ewald.f90:127)
    if (e.nparts > 1 && (e.mesh ? l % e.nparts != e.part : e.part != 0)) continue;
    es += 0.5 * e.lambda[l] * a * a * e.pval[sp];
    ds += e.lambda[l] * a * d * e.pval[sp];
  }
  e_atom[i] += es;
  if (dgw) dedcn[i] += ds;
}

// Partial structure factors: one wavefront per (tile, chunk of atoms of one
// species), lane = vector; the chunk's atoms are wave-uniform, so their
// factors and the h, k table entries come through scalar loads.  Factors
// l0 .. l0+lcount of the partial sum over the chunk go to
// part[chunk][l][vector] (no atomics; ewald_f sums the chunks of a species);
// kL = lcount rounded up to 8, the factor rows are zero padded to it.
template <int kL>
__global__ __launch_bounds__(kBlock) void ewald_sf_kernel(EwaldBatch e, const double* __restrict__ c6l,
                                                          const double2* __restrict__ etab, int l0, int lcount,
                                                          double2* __restrict__ part) {
  // wave-uniform: the tile, the chunk and its atoms' data come through scalar loads
  const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  if (item >= e.ntiles * e.nchunks) return;
  const int chunk = item / e.ntiles, t = item % e.ntiles;
  const int lane = wave_lane();
  const TileVec v = tile_vector(e, e.tiles[t], lane);
  const int4 c = e.chunks[chunk];
  double sr[kL], si[kL];
#pragma unroll
  for (int l = 0; l < kL; ++l) {
    sr[l] = 0.0;
    si[l] = 0.0;
  }
  for (int q = c.y; q < c.z; ++q) {
    double co, s;
    tile_phase(etab + static_cast<long long>(q) * e.tstride, v, co, s);
    const double* f = c6l + static_cast<long long>(q) * e.rank_pad + l0;  // l0 + kL <= rank_pad
#pragma unroll
    for (int l = 0; l < kL; ++l) {
      const double w = f[l];
      sr[l] += w * co;
      si[l] += w * s;
    }
  }
  const long long k = static_cast<long long>(t) * kWave + lane;
  const long long nk = static_cast<long long>(e.ntiles) * kWave;
  const bool valid = v.w > 0.0;
#pragma unroll
  for (int l = 0; l < kL; ++l)
    if (l < lcount)
      part[(static_cast<long long>(chunk) * e.rank + l0 + l) * nk + k] =
          valid ? make_double2(sr[l], si[l]) : make_double2(0.0, 0.0);
}

// Per vector k (one lane each): S_al from the chunk partials, the transform
// phihat_ab(k) of the pair potential and F_al = sum_b phihat_ab S_bl
// (written for ewald_atom); and the virial of the vector (ewald.f90:221-233):
// 1/2 qq k k^T / (|k| V) with qq = sum_l lambda_l sum_ab dphihat_ab
// Re(S_al conj S_bl), minus its energy E_k on the diagonal (ewald.f90:
// 248-250).  The virial is summed over the wavefront and added to the
// per-atom virial of the system's first atom.
// kMany (more than kMaxEwaldSpecies species): the structure factors S_al of
// the lane's vector go to Sbuf [nsp][rank][nk] (each lane reads back only its
// own entries) instead of registers.
template <bool kMany>
__global__ __launch_bounds__(kBlock) void ewald_f_kernel(EwaldBatch e, const double2* __restrict__ part,
                                                         double2* __restrict__ phi, double2* __restrict__ F,
                                                         double2* __restrict__ Sbuf, double* __restrict__ sig_atom,
                                                         int grad) {
  const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long nk = static_cast<long long>(e.ntiles) * kWave;
  const int t = static_cast<int>(k / kWave);
  const int nsp = e.nsp;
  double sg[6] = {0, 0, 0, 0, 0, 0};
  if (t < e.ntiles) {
    const TileVec v = tile_vector(e, e.tiles[t], static_cast<int>(k % kWave));
    const double knorm = sqrt(v.kx * v.kx + v.ky * v.ky + v.kz * v.kz);
    for (int a = 0; a < nsp; ++a)
      for (int c = 0; c <= a; ++c) {
        double p = 0.0, dp = 0.0;
        for (int q = 0; q < e.nterm[a * nsp + c]; ++q) {
          double pv, dpv;
          fourier_transform(e.terms[2 * (a * nsp + c) + q], knorm, pv, dpv);
          p += pv;
          dp += dpv;
        }
        phi[static_cast<long long>(a * (a + 1) / 2 + c) * nk + k] = make_double2(p, dp);
      }
    double ek = 0.0, qq = 0.0;
    for (int l = 0; l < e.rank && kMany; ++l) {
      // S_al of this lane's vector, then F_al = sum_c phihat_ac S_cl
      for (int a = 0; a < nsp; ++a) {
        double sr = 0.0, si = 0.0;
        for (int ch = e.chunk_seg[a]; ch < e.chunk_seg[a + 1]; ++ch) {
          const double2 pv = part[(static_cast<long long>(ch) * e.rank + l) * nk + k];
          sr += pv.x;
          si += pv.y;
        }
        Sbuf[(static_cast<long long>(a) * e.rank + l) * nk + k] = make_double2(sr, si);
      }
      double el = 0.0, ql = 0.0;
      for (int a = 0; a < nsp; ++a) {
        const double2 sa = Sbuf[(static_cast<long long>(a) * e.rank + l) * nk + k];
        double f_re = 0.0, f_im = 0.0;
        for (int c = 0; c < nsp; ++c) {
          const double2 sc = Sbuf[(static_cast<long long>(c) * e.rank + l) * nk + k];
          const int pr = a >= c ? a * (a + 1) / 2 + c : c * (c + 1) / 2 + a;
          const double2 pv = phi[static_cast<long long>(pr) * nk + k];
          f_re += pv.x * sc.x;
          f_im += pv.x * sc.y;
          ql += pv.y * (sa.x * sc.x + sa.y * sc.y);
        }
        F[(static_cast<long long>(a) * e.rank + l) * nk + k] = make_double2(f_re, f_im);
        el += sa.x * f_re + sa.y * f_im;
      }
      ek += e.lambda[l] * el;
      qq += e.lambda[l] * ql;
    }
    for (int l = 0; l < e.rank && !kMany; ++l) {
      double s_re[kMaxEwaldSpecies], s_im[kMaxEwaldSpecies];
#pragma unroll
      for (int a = 0; a < kMaxEwaldSpecies; ++a) {
        s_re[a] = 0.0;
        s_im[a] = 0.0;
        if (a < nsp) {
          for (int ch = e.chunk_seg[a]; ch < e.chunk_seg[a + 1]; ++ch) {
            const double2 pv = part[(static_cast<long long>(ch) * e.rank + l) * nk + k];
            s_re[a] += pv.x;
            s_im[a] += pv.y;
          }
        }
      }
      double el = 0.0, ql = 0.0;
      // a runtime loop, as in spme_conv_kernel (register pressure)
#pragma unroll 1
      for (int a = 0; a < nsp; ++a) {
        double sa_re = 0.0, sa_im = 0.0;
#pragma unroll
        for (int c = 0; c < kMaxEwaldSpecies; ++c)
          if (c == a) {
            sa_re = s_re[c];
            sa_im = s_im[c];
          }
        double f_re = 0.0, f_im = 0.0;
#pragma unroll
        for (int c = 0; c < kMaxEwaldSpecies; ++c) {
          if (c >= nsp) continue;
          const int pr = a >= c ? a * (a + 1) / 2 + c : c * (c + 1) / 2 + a;
          const double2 pv = phi[static_cast<long long>(pr) * nk + k];
          f_re += pv.x * s_re[c];
          f_im += pv.x * s_im[c];
          ql += pv.y * (sa_re * s_re[c] + sa_im * s_im[c]);
        }
        F[(static_cast<long long>(a) * e.rank + l) * nk + k] = make_double2(f_re, f_im);
        el += sa_re * f_re + sa_im * f_im;
      }
      ek += e.lambda[l] * el;
      qq += e.lambda[l] * ql;
    }
    if (grad && v.w > 0.0) {
      // E_k = -1/(2V) sum_l lambda_l sum_a Re(S_al conj F_al), weighted
      const double erec = -0.5 * v.w * ek * e.inv_vol;
      const double h = knorm > 0.0 ? 0.5 * v.w * qq * e.inv_vol / knorm : 0.0;
      sg[0] = h * v.kx * v.kx - erec;
      sg[1] = h * v.ky * v.ky - erec;
      sg[2] = h * v.kz * v.kz - erec;
      sg[3] = h * v.kx * v.ky;
      sg[4] = h * v.kx * v.kz;
      sg[5] = h * v.ky * v.kz;
    }
  }
  if (grad) {
#pragma unroll
    for (int c = 0; c < 6; ++c) sg[c] = wave_sum(sg[c]);
    if (wave_lane() == 0)
      for (int c = 0; c < 6; ++c) atomicAdd(&sig_atom[6 * e.atom_begin + c], sg[c]);
  }
}

// Per group of up to kEwaldGroup atoms of one species and a range of tiles
// (one wavefront, lane = vector): A_i = sum_l lambda_l c6l_il conj F_al(k)
// and B_i the same with dc6l, each F load shared by the group; then
// (ewald.f90:193-212) E_i -= 1/(2V) Re(e^{ik.x} A_i), dE/dCN_i -= 1/V
// Re(e^{ik.x} B_i), grad_i += 1/V Im(e^{ik.x} A_i) k, weighted; the
// wavefront's sums are added to the atoms' accumulators.
template <bool kGrad>
__global__ __launch_bounds__(kBlock) void ewald_atom_kernel(EwaldBatch e, const double* __restrict__ lc6l,
                                                            const double* __restrict__ ldc6l,
                                                            const double2* __restrict__ etab,
                                                            const double2* __restrict__ F, int ksplit,
                                                            double* __restrict__ e_atom, double* __restrict__ dedcn,
                                                            double* __restrict__ grad) {
  // wave-uniform: the group, its atoms' factors and the tiles come through scalar loads
  const int item = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  if (item >= e.ngroups * ksplit) return;
  // neighbouring wavefronts: neighbouring groups (mostly one species) on the
  // same tile range, so their F loads meet in the caches
  const int g = item % e.ngroups, part = item / e.ngroups;
  const int4 grp = e.groups[g];
  const int cnt = grp.z - grp.y;
  const int lane = wave_lane();
  const long long ld = e.natoms + kEwaldGroup;
  const double* gl = lc6l + grp.y;   // gl[l * ld + j]: lambda_l c6l of the group's atom j
  const double* gd = kGrad ? ldc6l + grp.y : nullptr;
  const long long nk = static_cast<long long>(e.ntiles) * kWave;
  const double2* Fa = F + static_cast<long long>(grp.x) * e.rank * nk;
  const int t0 = static_cast<int>(static_cast<long long>(e.ntiles) * part / ksplit);
  const int t1 = static_cast<int>(static_cast<long long>(e.ntiles) * (part + 1) / ksplit);
  double ae[kEwaldGroup], ad[kEwaldGroup], g0[kEwaldGroup], g1[kEwaldGroup], g2[kEwaldGroup];
#pragma unroll
  for (int j = 0; j < kEwaldGroup; ++j) ae[j] = ad[j] = g0[j] = g1[j] = g2[j] = 0.0;
  for (int t = t0; t < t1; ++t) {
    const TileVec v = tile_vector(e, e.tiles[t], lane);
    const long long k = static_cast<long long>(t) * kWave + lane;
    double a_re[kEwaldGroup], a_im[kEwaldGroup], b_re[kEwaldGroup], b_im[kEwaldGroup];
#pragma unroll
    for (int j = 0; j < kEwaldGroup; ++j) a_re[j] = a_im[j] = b_re[j] = b_im[j] = 0.0;
    // four factors per step, their F loads issued together
    for (int l0 = 0; l0 < e.rank; l0 += 4) {
      double2 fv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        fv[u] = l0 + u < e.rank ? Fa[static_cast<long long>(l0 + u) * nk + k] : make_double2(0.0, 0.0);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (l0 + u >= e.rank) break;
        const int l = l0 + u;
        // all kEwaldGroup columns (the padding past the group is discarded):
        // adjacent loads the compiler merges
#pragma unroll
        for (int j = 0; j < kEwaldGroup; ++j) {
          const double cl = gl[l * ld + j];
          a_re[j] += cl * fv[u].x;
          a_im[j] += cl * fv[u].y;
          if (kGrad) {
            const double dl = gd[l * ld + j];
            b_re[j] += dl * fv[u].x;
            b_im[j] += dl * fv[u].y;
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kEwaldGroup; ++j) {
      if (j < cnt) {
        double c, s;
        tile_phase(etab + static_cast<long long>(grp.y + j) * e.tstride, v, c, s);
        // e^{ik.x} conj(F) = (c + i s)(F_re - i F_im)
        ae[j] += v.w * (c * a_re[j] + s * a_im[j]);
        if (kGrad) {
          ad[j] += v.w * (c * b_re[j] + s * b_im[j]);
          const double zim = v.w * (s * a_re[j] - c * a_im[j]);
          g0[j] += zim * v.kx;
          g1[j] += zim * v.ky;
          g2[j] += zim * v.kz;
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kEwaldGroup; ++j) {
    if (j < cnt) {
      const double se = wave_sum(ae[j]);
      double sd = 0.0, s0 = 0.0, s1 = 0.0, s2 = 0.0;
      if (kGrad) {
        sd = wave_sum(ad[j]);
        s0 = wave_sum(g0[j]);
        s1 = wave_sum(g1[j]);
        s2 = wave_sum(g2[j]);
      }
      if (lane == 0) {
        const int i = e.sorder[grp.y + j];
        atomicAdd(&e_atom[i], -0.5 * se * e.inv_vol);
        if (kGrad) {
          atomicAdd(&dedcn[i], -sd * e.inv_vol);
          atomicAdd(&grad[3 * i], s0 * e.inv_vol);
          atomicAdd(&grad[3 * i + 1], s1 * e.inv_vol);
          atomicAdd(&grad[3 * i + 2], s2 * e.inv_vol);
        }
      }
    }
  }
}

// ------------------------------------------------------- particle mesh
// fourier/spme.f90:78-328: the low-rank factors spread onto one mesh per
// species with order-6 cardinal B-splines, transformed (real to half-complex,
// hipFFT), contracted with the tabulated kernel, transformed back and
// gathered onto the atoms, term by term of the expansion.  Mesh points are
// stored [g3][g2][g1] (g1 fastest), half-complex ones [g3][g2][h1], h1 <= m0/2.

__device__ __forceinline__ int mesh_wrap(int i, int m) { return i - m * ((2 * i) / m); }  // spme.f90:393-395

// spme.f90:457-499 fill_bspline (order 6) at offset w in [0, 1).
__device__ void bspline6(double w, double* th, double* dth) {
  for (int k = 0; k < kSplineOrder; ++k) th[k] = 0.0;
  th[0] = 1.0 - w;
  th[1] = w;
  for (int korder = 3; korder <= kSplineOrder - 1; ++korder) {
    const double div = 1.0 / (korder - 1);
    th[korder - 1] = div * w * th[korder - 2];
    for (int j = 1; j <= korder - 2; ++j)
      th[korder - j - 1] = div * ((w + j) * th[korder - j - 2] + (korder - j - w) * th[korder - j - 1]);
    th[0] = div * (1.0 - w) * th[0];
  }
  dth[0] = -th[0];
  for (int j = 2; j <= kSplineOrder; ++j) dth[j - 1] = th[j - 2] - th[j - 1];
  const double div = 1.0 / (kSplineOrder - 1);
  th[kSplineOrder - 1] = div * w * th[kSplineOrder - 2];
  for (int j = 1; j <= kSplineOrder - 2; ++j)
    th[kSplineOrder - j - 1] =
        div * ((w + j) * th[kSplineOrder - j - 2] + (kSplineOrder - j - w) * th[kSplineOrder - j - 1]);
  th[0] = div * (1.0 - w) * th[0];
}

// spme.f90:417-453 get_spline_weights: per atom of the system and direction
// the first mesh point (base + 1) and the weights and their derivatives.
__global__ __launch_bounds__(kBlock) void spme_spline_kernel(DevBatch b, EwaldBatch e, SpmeBatch s,
                                                             int* __restrict__ base, double* __restrict__ theta,
                                                             double* __restrict__ dtheta) {
  const int il = blockIdx.x * blockDim.x + threadIdx.x;
  if (il >= e.natoms) return;
  const double4 x = b.pos[e.atom_begin + il];
  const int md[3] = {s.m0, s.m1, s.m2};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double f = s.inv[3 * d] * x.x + s.inv[3 * d + 1] * x.y + s.inv[3 * d + 2] * x.z;
    const double u = md[d] * (f - floor(f));
    base[3 * il + d] = static_cast<int>(floor(u)) - kSplineOrder;
    bspline6(u - floor(u), theta + (3 * il + d) * kSplineOrder, dtheta + (3 * il + d) * kSplineOrder);
  }
}

// spme.f90:346-413 get_mesh_kernel with spme.f90:503-529 (Euler factors),
// one thread per half-complex point m: the kernel of every species pair,
// K = sfac (phihat(k_m) + phihat(k_-m)) / 2 (the real part of the
// reference's back transform is the transform of the kernel symmetrised over
// +-m; k_-m differs from -k_m on the Nyquist planes); with the gradient, the
// virial weights sfac dphihat at m and, when -m is not stored, at -m.
__global__ __launch_bounds__(kBlock) void spme_table_kernel(EwaldBatch e, SpmeBatch s, double* __restrict__ K,
                                                            double* __restrict__ Dm, double* __restrict__ Dmir) {
  const long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= s.nch) return;
  const int h1 = static_cast<int>(p % s.mh), g2 = static_cast<int>((p / s.mh) % s.m1);
  const int g3 = static_cast<int>(p / (static_cast<long long>(s.mh) * s.m1));
  const int r1 = (s.m0 - h1) % s.m0, r2 = (s.m1 - g2) % s.m1, r3 = (s.m2 - g3) % s.m2;
  const bool stored_mirror = r1 == h1 || (2 * h1 == s.m0);  // -m is a stored point (or m itself)
  double km[3], kr[3];
  const int j1 = mesh_wrap(h1, s.m0), j2 = mesh_wrap(g2, s.m1), j3 = mesh_wrap(g3, s.m2);
  const int q1 = mesh_wrap(r1, s.m0), q2 = mesh_wrap(r2, s.m1), q3 = mesh_wrap(r3, s.m2);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    km[c] = s.rec[c] * j1 + s.rec[3 + c] * j2 + s.rec[6 + c] * j3;
    kr[c] = s.rec[c] * q1 + s.rec[3 + c] * q2 + s.rec[6 + c] * q3;
  }
  const double nm = sqrt(km[0] * km[0] + km[1] * km[1] + km[2] * km[2]);
  const double nr = sqrt(kr[0] * kr[0] + kr[1] * kr[1] + kr[2] * kr[2]);
  const double sfac = s.bfac[h1] * s.bfac[s.mmax + g2] * s.bfac[2 * s.mmax + g3];
  const int nsp = e.nsp;
  for (int a = 0; a < nsp; ++a)
    for (int c = 0; c <= a; ++c) {
      double pm = 0.0, dpm = 0.0, pr = 0.0, dpr = 0.0;
      for (int t = 0; t < e.nterm[a * nsp + c]; ++t) {
        double pv, dpv;
        fourier_transform(e.terms[2 * (a * nsp + c) + t], nm, pv, dpv);
        pm += pv;
        dpm += dpv;
        if (nr != nm) {
          fourier_transform(e.terms[2 * (a * nsp + c) + t], nr, pv, dpv);
          pr += pv;
          dpr += dpv;
        }
      }
      if (nr == nm) {
        pr = pm;
        dpr = dpm;
      }
      const long long o = static_cast<long long>(a * (a + 1) / 2 + c) * s.nch + p;
      K[o] = sfac * 0.5 * (pm + pr);
      if (Dm) {
        Dm[o] = nm > 0.0 ? sfac * dpm : 0.0;
        Dmir[o] = (!stored_mirror && nr > 0.0) ? sfac * dpr : 0.0;
      }
    }
}

// Spreading of term l (spme.f90:201-218): one block per atom, one thread per
// each of its 6^3 mesh points.
__global__ __launch_bounds__(kBlock) void spme_spread_kernel(DevBatch b, EwaldBatch e, SpmeBatch s,
                                                             const double* __restrict__ c6l,
                                                             const int* __restrict__ base,
                                                             const double* __restrict__ theta, int l,
                                                             double* __restrict__ Q) {
  const int il = blockIdx.x;
  const int j = threadIdx.x;
  if (j >= kSplineOrder * kSplineOrder * kSplineOrder) return;
  const int j1 = j % kSplineOrder, j2 = (j / kSplineOrder) % kSplineOrder, j3 = j / (kSplineOrder * kSplineOrder);
  const int g1 = ((base[3 * il] + j1 + 1) % s.m0 + s.m0) % s.m0;
  const int g2 = ((base[3 * il + 1] + j2 + 1) % s.m1 + s.m1) % s.m1;
  const int g3 = ((base[3 * il + 2] + j3 + 1) % s.m2 + s.m2) % s.m2;
  const double* th = theta + 3 * il * kSplineOrder;
  const double w = th[j1] * th[kSplineOrder + j2] * th[2 * kSplineOrder + j3];
  const double f = c6l[static_cast<long long>(e.spos[il]) * e.rank_pad + l];
  const int a = b.sp[e.atom_begin + il];
  atomicAdd(&Q[static_cast<long long>(a) * s.ng + (static_cast<long long>(g3) * s.m1 + g2) * s.m0 + g1], f * w);
}

// Contraction with the kernel (spme.f90:225-242) per half-complex point and
// the virial of term l (spme.f90:284-314): qq at m (and at -m where it is not
// stored) times k k^T / (2 |k| V).  Grid-stride over the points; the virial
// is summed over the block and added to the per-atom virial of the system's
// first atom once per block.
// kMany (more than kMaxEwaldSpecies species): the transforms Q_c are read
// from Qh in the inner loop instead of registers.
template <bool kMany>
__global__ __launch_bounds__(kBlock) void spme_conv_kernel(EwaldBatch e, SpmeBatch s, const double2* __restrict__ Qh,
                                                           const double* __restrict__ K, const double* __restrict__ Dm,
                                                           const double* __restrict__ Dmir, int l,
                                                           double2* __restrict__ Th, double* __restrict__ sig_atom) {
  const int nsp = e.nsp;
  double sg[6] = {0, 0, 0, 0, 0, 0};
  for (long long p = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; p < s.nch;
       p += static_cast<long long>(gridDim.x) * blockDim.x) {
    double qm = 0.0, qx = 0.0;
    if (kMany) {
      for (int a = 0; a < nsp; ++a) {
        const double2 qa = Qh[static_cast<long long>(a) * s.nch + p];
        double tr = 0.0, ti = 0.0;
        for (int c = 0; c < nsp; ++c) {
          const double2 qc = Qh[static_cast<long long>(c) * s.nch + p];
          const long long o = static_cast<long long>(a >= c ? a * (a + 1) / 2 + c : c * (c + 1) / 2 + a) * s.nch + p;
          const double kk = K[o];
          tr += kk * qc.x;
          ti += kk * qc.y;
          if (Dm && c <= a) {  // each pair once
            const double re = (c == a ? 1.0 : 2.0) * (qa.x * qc.x + qa.y * qc.y);
            qm += Dm[o] * re;
            qx += Dmir[o] * re;
          }
        }
        Th[static_cast<long long>(a) * s.nch + p] = make_double2(tr, ti);
      }
    }
    double qr[kMaxEwaldSpecies], qi[kMaxEwaldSpecies];
#pragma unroll
    for (int a = 0; a < kMaxEwaldSpecies; ++a) {
      const double2 v = (!kMany && a < nsp) ? Qh[static_cast<long long>(a) * s.nch + p] : make_double2(0.0, 0.0);
      qr[a] = v.x;
      qi[a] = v.y;
    }
    // a runtime loop (a fully unrolled pair loop hoists every table load
    // into registers and runs at one wavefront per SIMD)
#pragma unroll 1
    for (int a = 0; a < nsp && !kMany; ++a) {
      double ar = 0.0, ai = 0.0;  // Q_a, selected without indexing the register array
#pragma unroll
      for (int c = 0; c < kMaxEwaldSpecies; ++c)
        if (c == a) {
          ar = qr[c];
          ai = qi[c];
        }
      double tr = 0.0, ti = 0.0;
#pragma unroll
      for (int c = 0; c < kMaxEwaldSpecies; ++c) {
        if (c >= nsp) continue;
        const long long o = static_cast<long long>(a >= c ? a * (a + 1) / 2 + c : c * (c + 1) / 2 + a) * s.nch + p;
        const double kk = K[o];
        tr += kk * qr[c];
        ti += kk * qi[c];
        if (Dm && c <= a) {  // each pair once
          const double re = (c == a ? 1.0 : 2.0) * (ar * qr[c] + ai * qi[c]);
          qm += Dm[o] * re;
          qx += Dmir[o] * re;
        }
      }
      Th[static_cast<long long>(a) * s.nch + p] = make_double2(tr, ti);
    }
    if (Dm) {
      const int h1 = static_cast<int>(p % s.mh);
      const int rest = static_cast<int>(p / s.mh);
      const int g2 = rest % s.m1, g3 = rest / s.m1;
      const int r1 = (s.m0 - h1) % s.m0, r2 = (s.m1 - g2) % s.m1, r3 = (s.m2 - g3) % s.m2;
      const int jj[2][3] = {{mesh_wrap(h1, s.m0), mesh_wrap(g2, s.m1), mesh_wrap(g3, s.m2)},
                            {mesh_wrap(r1, s.m0), mesh_wrap(r2, s.m1), mesh_wrap(r3, s.m2)}};
      const double qq[2] = {e.lambda[l] * qm, e.lambda[l] * qx};
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        double k[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) k[c] = s.rec[c] * jj[q][0] + s.rec[3 + c] * jj[q][1] + s.rec[6 + c] * jj[q][2];
        const double kn = sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
        if (kn <= 0.0) continue;
        const double h = 0.5 * qq[q] * s.inv_vol / kn;
        sg[0] += h * k[0] * k[0];
        sg[1] += h * k[1] * k[1];
        sg[2] += h * k[2] * k[2];
        sg[3] += h * k[0] * k[1];
        sg[4] += h * k[0] * k[2];
        sg[5] += h * k[1] * k[2];
      }
    }
  }
  if (Dm) {
    __shared__ double red[kWavesPerBlock][6];
#pragma unroll
    for (int c = 0; c < 6; ++c) sg[c] = wave_sum(sg[c]);
    const int wv = threadIdx.x >> 6;
    if (wave_lane() == 0)
      for (int c = 0; c < 6; ++c) red[wv][c] = sg[c];
    __syncthreads();
    if (threadIdx.x < 6) {
      double v = 0.0;
      for (int q = 0; q < kWavesPerBlock; ++q) v += red[q][threadIdx.x];
      atomicAdd(&sig_atom[6 * e.atom_begin + threadIdx.x], v);
    }
  }
}

// Gathering of term l (spme.f90:246-282): one wavefront per atom, lanes over
// its 6^3 mesh points of its species' back-transformed mesh; energy, dE/dCN,
// gradient, and the energy on the virial diagonal (spme.f90:319-322).
template <bool kGrad>
__global__ __launch_bounds__(kBlock) void spme_gather_kernel(DevBatch b, EwaldBatch e, SpmeBatch s,
                                                             const double* __restrict__ T,
                                                             const int* __restrict__ base,
                                                             const double* __restrict__ theta,
                                                             const double* __restrict__ dtheta,
                                                             const double* __restrict__ lc6l,
                                                             const double* __restrict__ ldc6l, int l,
                                                             double* __restrict__ e_atom, double* __restrict__ dedcn,
                                                             double* __restrict__ grad, double* __restrict__ sig_atom) {
  const int il = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  if (il >= e.natoms) return;
  const int i = e.atom_begin + il;
  const int a = b.sp[i];
  const double* Ta = T + static_cast<long long>(a) * s.ng;
  const double* th = theta + 3 * il * kSplineOrder;
  const double* dth = dtheta + 3 * il * kSplineOrder;
  double w = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
  for (int j = wave_lane(); j < kSplineOrder * kSplineOrder * kSplineOrder; j += kWave) {
    const int j1 = j % kSplineOrder, j2 = (j / kSplineOrder) % kSplineOrder, j3 = j / (kSplineOrder * kSplineOrder);
    const int g1 = ((base[3 * il] + j1 + 1) % s.m0 + s.m0) % s.m0;
    const int g2 = ((base[3 * il + 1] + j2 + 1) % s.m1 + s.m1) % s.m1;
    const int g3 = ((base[3 * il + 2] + j3 + 1) % s.m2 + s.m2) % s.m2;
    const double sf = Ta[(static_cast<long long>(g3) * s.m1 + g2) * s.m0 + g1];
    const double t1 = th[j1], t2 = th[kSplineOrder + j2], t3 = th[2 * kSplineOrder + j3];
    w += t1 * t2 * t3 * sf;
    if (kGrad) {
      const double u1 = dth[j1] * t2 * t3, u2 = t1 * dth[kSplineOrder + j2] * t3;
      const double u3 = t1 * t2 * dth[2 * kSplineOrder + j3];
      d0 += (s.dudr[0] * u1 + s.dudr[3] * u2 + s.dudr[6] * u3) * sf;
      d1 += (s.dudr[1] * u1 + s.dudr[4] * u2 + s.dudr[7] * u3) * sf;
      d2 += (s.dudr[2] * u1 + s.dudr[5] * u2 + s.dudr[8] * u3) * sf;
    }
  }
  w = wave_sum(w);
  if (kGrad) {
    d0 = wave_sum(d0);
    d1 = wave_sum(d1);
    d2 = wave_sum(d2);
  }
  if (wave_lane() == 0) {
    const long long ld = e.natoms + kEwaldGroup;
    const int row = e.spos[il];
    const double escale = -0.5 * s.inv_vol;
    const double lc = lc6l[l * ld + row];
    const double en = escale * lc * w;
    e_atom[i] += en;
    if (kGrad) {
      dedcn[i] += 2.0 * escale * ldc6l[l * ld + row] * w;
      grad[3 * i] += 2.0 * escale * lc * d0;
      grad[3 * i + 1] += 2.0 * escale * lc * d1;
      grad[3 * i + 2] += 2.0 * escale * lc * d2;
      for (int c = 0; c < 3; ++c) atomicAdd(&sig_atom[6 * e.atom_begin + c], -en);
    }
  }
}

}  // namespace

// ---------------------------------------------------------------- launchers
size_t pair_record_bytes(bool mixed) { return mixed ? sizeof(float4) : sizeof(double4); }
size_t slot_record_bytes(bool mixed) { return mixed ? sizeof(Slot<float>) : sizeof(Slot<double>); }

void launch_pack(const double* xyz, int n, double4* pos, hipStream_t st) {
  hipLaunchKernelGGL(pack_positions, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, xyz, n, pos);
}

namespace {
size_t align256(size_t x) { return (x + 255) & ~size_t(255); }
size_t radix_sort_bytes(int n) {
  size_t bytes = 0;
  hip_check_sort(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, static_cast<unsigned long long*>(nullptr),
                                                    static_cast<unsigned long long*>(nullptr),
                                                    static_cast<int*>(nullptr), static_cast<int*>(nullptr), n));
  return bytes;
}
}  // namespace

size_t spatial_order_temp_bytes(int n) {
  return 2 * align256(sizeof(unsigned long long) * n) + align256(sizeof(int) * n) + radix_sort_bytes(n);
}

void launch_spatial_order(const DevBatch& b, void* temp, size_t temp_bytes, int* rank, int* order, hipStream_t st) {
  const int n = b.natoms;
  char* t = static_cast<char*>(temp);
  auto* keys_in = reinterpret_cast<unsigned long long*>(t);
  t += align256(sizeof(unsigned long long) * n);
  auto* keys_out = reinterpret_cast<unsigned long long*>(t);
  t += align256(sizeof(unsigned long long) * n);
  int* idx = reinterpret_cast<int*>(t);
  t += align256(sizeof(int) * n);
  size_t sort_bytes = temp_bytes - static_cast<size_t>(t - static_cast<char*>(temp));
  int sys_bits = 0;
  while ((1LL << sys_bits) < b.nsys) ++sys_bits;
  hipLaunchKernelGGL(spatial_key_kernel, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, b, keys_in, idx);
  hip_check_sort(hipcub::DeviceRadixSort::SortPairs(t, sort_bytes, keys_in, keys_out, idx, order, n, 0,
                                                    30 + sys_bits, st));
  hipLaunchKernelGGL(spatial_rank_kernel, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, st, b, order, rank);
}

void launch_cn(const DevBatch& b, bool mixed, double* cn, hipStream_t st) {
  const dim3 grid(blocks_for(b.natoms, kWavesPerBlock)), blk(kBlock);
  if (mixed) hipLaunchKernelGGL(cn_kernel<float>, grid, blk, 0, st, b, cn);
  else hipLaunchKernelGGL(cn_kernel<double>, grid, blk, 0, st, b, cn);
}

void launch_weights(const DevBatch& b, const double* cn, double* gw, double* dgw, hipStream_t st) {
  hipLaunchKernelGGL(weights_kernel, dim3(blocks_for(b.natoms, kBlock)), dim3(kBlock), 0, st, b, cn, gw, dgw);
}

namespace {

// The positions in spatial order, posr[x] = pos[order[x]], and the bounding
// box of every 64-position tile of that order (wavefront per tile; the tiles
// cover every position).
__global__ __launch_bounds__(kBlock) void tile_box_kernel(DevBatch b, double4* __restrict__ posr,
                                                          double4* __restrict__ tbox) {
  const int lane = wave_lane();
  const int t = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (t >= b.ntiles) return;
  const int2 tile = b.tiles[t];
  const int x = tile.x + min(lane, tile.y - 1);  // lanes past the tile repeat its last position
  const double4 p = b.pos[b.order[x]];
  if (lane < tile.y) posr[x] = p;
  double lo[3] = {p.x, p.y, p.z}, hi[3] = {p.x, p.y, p.z};
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      lo[c] = fmin(lo[c], __shfl_xor(lo[c], off, kWave));
      hi[c] = fmax(hi[c], __shfl_xor(hi[c], off, kWave));
    }
  if (lane == 0) {
    tbox[2 * t] = make_double4(lo[0], lo[1], lo[2], 0.0);
    tbox[2 * t + 1] = make_double4(hi[0], hi[1], hi[2], 0.0);
  }
}
}  // namespace

// Work area of the allocation-free entry point.  In host code this name is a
// stand-in, not a device address: only the address from hipGetSymbolAddress
// (noalloc_area) may be passed to kernels or copies.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

void* noalloc_area(int device, hipError_t* hip_err) {
  static std::mutex mu;
  static std::map<int, void*> addr;  // per device: each GPU holds its own copy of the array
  std::lock_guard<std::mutex> lock(mu);
  *hip_err = hipSuccess;
  auto it = addr.find(device);
  if (it != addr.end()) return it->second;
  void* p = nullptr;
  *hip_err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (*hip_err != hipSuccess) return nullptr;
  addr.emplace(device, p);
  return p;
}

void launch_rank_pos(const DevBatch& b, hipStream_t st) {
  if (b.ntiles > 0)
    hipLaunchKernelGGL(tile_box_kernel, dim3(blocks_for(b.ntiles, kWavesPerBlock)), dim3(kBlock), 0, st, b,
                       const_cast<double4*>(b.posr), const_cast<double4*>(b.tbox));
}

void launch_pair(const DevBatch& b, bool mixed, const double* gw, const double* dgw, double* gwr, int* spr,
                 void* pair, hipStream_t st) {
  hipLaunchKernelGGL(rank_weights_kernel, dim3(blocks_for(b.natoms, kBlock)), dim3(kBlock), 0, st, b, gw, dgw, gwr,
                     spr);
  const size_t lds = sizeof(double) * static_cast<size_t>(b.nsp) * (2 * kMaxRef + 1);  // h, dh, r0p
  if (mixed)
    hipLaunchKernelGGL(pair_kernel<float>, dim3(b.natoms), dim3(kBlock), lds, st, b, gw, dgw, gwr, spr,
                       static_cast<float4*>(pair));
  else
    hipLaunchKernelGGL(pair_kernel<double>, dim3(b.natoms), dim3(kBlock), lds, st, b, gw, dgw, gwr, spr,
                       static_cast<double4*>(pair));
}

template <typename R, bool kGrad, bool kPairs>
void launch_disp2_t(const DevBatch& b, const DevWork& w, double* pair2, hipStream_t st) {
  const dim3 grid(blocks_for(b.natoms, kWavesPerBlock)), blk(kBlock);
  hipLaunchKernelGGL((disp2_kernel<R, kGrad, kPairs>), grid, blk, 0, st, b, static_cast<const R4<R>*>(w.pair),
                     w.e_atom, w.dedcn, w.grad, w.sig_atom, pair2);
}

void launch_disp2(const DevBatch& b, bool mixed, bool grad, const DevWork& w, double* pair2, hipStream_t st) {
  if (pair2) {
    if (mixed) launch_disp2_t<float, false, true>(b, w, pair2, st);
    else launch_disp2_t<double, false, true>(b, w, pair2, st);
  } else if (mixed) {
    if (grad) launch_disp2_t<float, true, false>(b, w, nullptr, st);
    else launch_disp2_t<float, false, false>(b, w, nullptr, st);
  } else {
    if (grad) launch_disp2_t<double, true, false>(b, w, nullptr, st);
    else launch_disp2_t<double, false, false>(b, w, nullptr, st);
  }
}

void launch_atm_count(const DevBatch& b, const void* pair, bool mixed, long long* count, double4* acc,
                      hipStream_t st) {
  const dim3 grid(blocks_for(b.natoms, kWavesPerBlock)), blk(kBlock);
  if (mixed)
    hipLaunchKernelGGL((atm_list_kernel<false, float>), grid, blk, 0, st, b, static_cast<const float4*>(pair), count,
                       (const long long*)nullptr, (Slot<float>*)nullptr, 0LL, acc);
  else
    hipLaunchKernelGGL((atm_list_kernel<false, double>), grid, blk, 0, st, b, static_cast<const double4*>(pair),
                       count, (const long long*)nullptr, (Slot<double>*)nullptr, 0LL, acc);
}

void launch_list_scan(const long long* count, int n, long long* offset, long long* item_off, hipStream_t st) {
  hipLaunchKernelGGL(list_scan_kernel, dim3(1), dim3(kScanBlock), 0, st, count, n, offset, item_off);
}

void launch_atm_fill(const DevBatch& b, const void* pair, bool mixed, const long long* offset, void* slots,
                     long long capacity, hipStream_t st) {
  const dim3 grid(blocks_for(b.natoms, kWavesPerBlock)), blk(kBlock);
  if (mixed)
    hipLaunchKernelGGL((atm_list_kernel<true, float>), grid, blk, 0, st, b, static_cast<const float4*>(pair),
                       (long long*)nullptr, offset, static_cast<Slot<float>*>(slots), capacity, (double4*)nullptr);
  else
    hipLaunchKernelGGL((atm_list_kernel<true, double>), grid, blk, 0, st, b, static_cast<const double4*>(pair),
                       (long long*)nullptr, offset, static_cast<Slot<double>*>(slots), capacity, (double4*)nullptr);
}

namespace {
// One thread per atom.  When the lists hold more than `capacity` slots no
// list is filled: the flag is raised, the work-item range is emptied, and the
// outputs of every atom of a system with a three-body term become NaN (its
// energy, gradient, dE/dCN -- which carries the NaN through the chain rule --
// and virial), so a replayed evaluation cannot return two-body-only results
// as if they were complete.  Systems without a three-body term keep theirs.
__global__ void atm_guard_kernel(DevBatch b, const long long* __restrict__ offset, long long capacity,
                                 long long* __restrict__ item_off, int* __restrict__ overflow,
                                 double* __restrict__ e_atom, double* __restrict__ dedcn, double* __restrict__ grad,
                                 double* __restrict__ sig_atom) {
  if (offset[b.natoms] <= capacity) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    *overflow = 1;
    item_off[b.natoms] = 0;
  }
  if (i >= b.natoms || fabs(b.par[b.sys[i]].s9) < kEps) return;
  const double nan = __builtin_nan("");
  e_atom[i] = nan;
  if (dedcn) dedcn[i] = nan;
  if (grad) grad[3 * i] = grad[3 * i + 1] = grad[3 * i + 2] = nan;
  if (sig_atom)
    for (int c = 0; c < 6; ++c) sig_atom[6 * i + c] = nan;
}

}  // namespace

void launch_atm_guard(const DevBatch& b, const long long* offset, long long capacity, long long* item_off,
                      int* overflow, const DevWork& w, bool grad, hipStream_t st) {
  hipLaunchKernelGGL(atm_guard_kernel, dim3(blocks_for(std::max(b.natoms, 1), kBlock)), dim3(kBlock), 0, st, b,
                     offset, capacity, item_off, overflow, w.e_atom, grad ? w.dedcn : nullptr,
                     grad ? w.grad : nullptr, grad ? w.sig_atom : nullptr);
}


namespace {
__global__ void zero_counter(unsigned long long* c) { *c = 0ull; }

template <typename R, bool kGrad, bool kSmooth, int kW, bool kPart>
void launch_atm_t(const DevBatch& b, const long long* item_off, long long max_items, const long long* offset,
                  const void* slots, const DevWork& w, const AtmShape& shape, unsigned long long* counter,
                  hipStream_t st) {
  long long grid = max_items;
  unsigned long long* wc = nullptr;
  if (shape.blocks_per_cu > 0) {
    grid = std::min<long long>(max_items, static_cast<long long>(shape.blocks_per_cu) * shape.num_cus);
    wc = counter;
    hipLaunchKernelGGL(zero_counter, dim3(1), dim3(1), 0, st, wc);
  }
  // w.atm_acc was cleared by the list count pass of this evaluation
  if (kGrad && !w.atm_acc) throw std::invalid_argument("three-body gradient launch without its per-atom sums");
  hipLaunchKernelGGL((atm_kernel<R, kGrad, kSmooth, kW, kPart>), dim3(static_cast<unsigned>(grid)), dim3(kW * kWave), 0,
                     st, b, item_off, offset, static_cast<const Slot<R>*>(slots),
                     static_cast<const R4<R>*>(w.pair), w.e_atom, w.dedcn, w.grad, w.sig_atom, w.atm_acc, wc);
  if (kGrad)
    hipLaunchKernelGGL(atm_acc_kernel, dim3(blocks_for(b.natoms, kBlock)), dim3(kBlock), 0, st, b, w.atm_acc,
                       w.grad, w.dedcn);
}

template <int kW, bool kPart>
void launch_atm_w(const DevBatch& b, bool mixed, bool grad, bool smooth, const long long* item_off,
                  long long max_items, const long long* offset, const void* slots, const DevWork& w,
                  const AtmShape& shape, unsigned long long* counter, hipStream_t st) {
  using L = void (*)(const DevBatch&, const long long*, long long, const long long*, const void*, const DevWork&,
                     const AtmShape&, unsigned long long*, hipStream_t);
  static const L table[8] = {
      launch_atm_t<double, false, false, kW, kPart>, launch_atm_t<double, false, true, kW, kPart>,
      launch_atm_t<double, true, false, kW, kPart>,  launch_atm_t<double, true, true, kW, kPart>,
      launch_atm_t<float, false, false, kW, kPart>,  launch_atm_t<float, false, true, kW, kPart>,
      launch_atm_t<float, true, false, kW, kPart>,   launch_atm_t<float, true, true, kW, kPart>};
  table[(mixed ? 4 : 0) + (grad ? 2 : 0) + (smooth ? 1 : 0)](b, item_off, max_items, offset, slots, w, shape,
                                                             counter, st);
}
}  // namespace

void launch_atm(const DevBatch& b, bool mixed, bool grad, bool smooth, bool partitioned, const AtmShape& shape,
                const long long* item_off, long long max_items, const long long* offset, const void* slots,
                const DevWork& w, unsigned long long* counter, hipStream_t st) {
  if (max_items <= 0) return;
  if (shape.blocks_per_cu > 0 && (!counter || shape.num_cus <= 0))
    throw std::invalid_argument("persistent three-body launch needs a work counter and the CU count");
  if (partitioned)  // its own instantiation (four wavefronts), so the others keep their registers
    launch_atm_w<4, true>(b, mixed, grad, smooth, item_off, max_items, offset, slots, w, shape, counter, st);
  else if (shape.waves == 2)
    launch_atm_w<2, false>(b, mixed, grad, smooth, item_off, max_items, offset, slots, w, shape, counter, st);
  else if (shape.waves == 4)
    launch_atm_w<4, false>(b, mixed, grad, smooth, item_off, max_items, offset, slots, w, shape, counter, st);
  else if (shape.waves == 1)
    launch_atm_w<1, false>(b, mixed, grad, smooth, item_off, max_items, offset, slots, w, shape, counter, st);
  else
    throw std::invalid_argument("three-body launch: wavefronts per block must be 1, 2 or 4");
}

namespace {
template <typename R, bool kSmooth>
void launch_atm_pairs_t(const DevBatch& b, const long long* item_off, long long max_items, const long long* offset,
                        const void* slots, const DevWork& w, double* pair3s, hipStream_t st) {
  hipLaunchKernelGGL((atm_pairs_kernel<R, kSmooth>), dim3(static_cast<unsigned>(max_items)), dim3(4 * kWave), 0, st,
                     b, item_off, offset, static_cast<const Slot<R>*>(slots), static_cast<const R4<R>*>(w.pair),
                     w.e_atom, pair3s);
}
}  // namespace

void launch_atm_pairs(const DevBatch& b, bool mixed, bool smooth, const long long* item_off, long long max_items,
                      const long long* offset, const void* slots, const DevWork& w, double* pair3s, hipStream_t st) {
  if (max_items <= 0) return;
  if (mixed) {
    if (smooth) launch_atm_pairs_t<float, true>(b, item_off, max_items, offset, slots, w, pair3s, st);
    else launch_atm_pairs_t<float, false>(b, item_off, max_items, offset, slots, w, pair3s, st);
  } else {
    if (smooth) launch_atm_pairs_t<double, true>(b, item_off, max_items, offset, slots, w, pair3s, st);
    else launch_atm_pairs_t<double, false>(b, item_off, max_items, offset, slots, w, pair3s, st);
  }
}

void launch_pair3_finalize(const DevBatch& b, const double* pair3s, double* out, hipStream_t st) {
  hipLaunchKernelGGL(pair3_finalize_kernel, dim3(blocks_for(b.natoms, kWavesPerBlock)), dim3(kBlock), 0, st, b,
                     pair3s, out);
}

void launch_lowrank(const DevBatch& b, const EwaldBatch& e, bool grad, const DevWork& w, const EwaldWork& ew,
                    hipStream_t st) {
  if (ew.etab)
    hipLaunchKernelGGL(ewald_table_kernel, dim3(blocks_for(static_cast<long long>(e.natoms) * e.tstride, kBlock)),
                       dim3(kBlock), 0, st, b, e, ew.etab);
  hipLaunchKernelGGL(lowrank_kernel, dim3(blocks_for(e.natoms, kBlock)), dim3(kBlock), 0, st, b, e, w.gw,
                     grad ? w.dgw : nullptr, ew.c6l, ew.lc6l, ew.ldc6l, w.e_atom, w.dedcn);
}

void launch_spme_setup(const DevBatch& b, const EwaldBatch& e, const SpmeBatch& s, bool grad, bool table,
                       const SpmeWork& sw, hipStream_t st) {
  hipLaunchKernelGGL(spme_spline_kernel, dim3(blocks_for(e.natoms, kBlock)), dim3(kBlock), 0, st, b, e, s, sw.base,
                     sw.theta, sw.dtheta);
  if (table)
    hipLaunchKernelGGL(spme_table_kernel, dim3(blocks_for(s.nch, kBlock)), dim3(kBlock), 0, st, e, s, sw.K,
                       grad ? sw.Dm : nullptr, grad ? sw.Dmir : nullptr);
}

void launch_spme_spread(const DevBatch& b, const EwaldBatch& e, const SpmeBatch& s, const EwaldWork& ew,
                        const SpmeWork& sw, int l, hipStream_t st) {
  hipLaunchKernelGGL(spme_spread_kernel, dim3(e.natoms), dim3(kBlock), 0, st, b, e, s, ew.c6l, sw.base, sw.theta, l,
                     sw.Q);
}

void launch_spme_conv(const EwaldBatch& e, const SpmeBatch& s, bool grad, const DevWork& w, const SpmeWork& sw, int l,
                      hipStream_t st) {
  const int blocks = static_cast<int>(std::min<long long>(blocks_for(s.nch, kBlock), 4096));
  if (e.nsp > kMaxEwaldSpecies)
    hipLaunchKernelGGL(spme_conv_kernel<true>, dim3(blocks), dim3(kBlock), 0, st, e, s, sw.Qh, sw.K,
                       grad ? sw.Dm : nullptr, grad ? sw.Dmir : nullptr, l, sw.Th, w.sig_atom);
  else
    hipLaunchKernelGGL(spme_conv_kernel<false>, dim3(blocks), dim3(kBlock), 0, st, e, s, sw.Qh, sw.K,
                       grad ? sw.Dm : nullptr, grad ? sw.Dmir : nullptr, l, sw.Th, w.sig_atom);
}

void launch_spme_gather(const DevBatch& b, const EwaldBatch& e, const SpmeBatch& s, bool grad, const DevWork& w,
                        const EwaldWork& ew, const SpmeWork& sw, int l, hipStream_t st) {
  const dim3 grid(blocks_for(e.natoms, kWavesPerBlock));
  if (grad)
    hipLaunchKernelGGL(spme_gather_kernel<true>, grid, dim3(kBlock), 0, st, b, e, s, sw.T, sw.base, sw.theta,
                       sw.dtheta, ew.lc6l, ew.ldc6l, l, w.e_atom, w.dedcn, w.grad, w.sig_atom);
  else
    hipLaunchKernelGGL(spme_gather_kernel<false>, grid, dim3(kBlock), 0, st, b, e, s, sw.T, sw.base, sw.theta,
                       sw.dtheta, ew.lc6l, ew.ldc6l, l, w.e_atom, w.dedcn, w.grad, w.sig_atom);
}

void launch_ewald_sf(const EwaldBatch& e, const EwaldWork& ew, hipStream_t st) {
  const dim3 grid(blocks_for(static_cast<long long>(e.ntiles) * e.nchunks, kWavesPerBlock));
  for (int l0 = 0; l0 < e.rank; l0 += 32) {
    const int lc = std::min(32, e.rank - l0);
    if (lc <= 8)
      hipLaunchKernelGGL(ewald_sf_kernel<8>, grid, dim3(kBlock), 0, st, e, ew.c6l, ew.etab, l0, lc, ew.part);
    else if (lc <= 16)
      hipLaunchKernelGGL(ewald_sf_kernel<16>, grid, dim3(kBlock), 0, st, e, ew.c6l, ew.etab, l0, lc, ew.part);
    else if (lc <= 24)
      hipLaunchKernelGGL(ewald_sf_kernel<24>, grid, dim3(kBlock), 0, st, e, ew.c6l, ew.etab, l0, lc, ew.part);
    else
      hipLaunchKernelGGL(ewald_sf_kernel<32>, grid, dim3(kBlock), 0, st, e, ew.c6l, ew.etab, l0, lc, ew.part);
  }
}

void launch_ewald_f(const EwaldBatch& e, bool grad, const DevWork& w, const EwaldWork& ew, hipStream_t st) {
  const dim3 grid(blocks_for(static_cast<long long>(e.ntiles) * kWave, kBlock));
  if (e.nsp > kMaxEwaldSpecies)
    hipLaunchKernelGGL(ewald_f_kernel<true>, grid, dim3(kBlock), 0, st, e, ew.part, ew.phi, ew.F, ew.S, w.sig_atom,
                       grad ? 1 : 0);
  else
    hipLaunchKernelGGL(ewald_f_kernel<false>, grid, dim3(kBlock), 0, st, e, ew.part, ew.phi, ew.F, ew.S, w.sig_atom,
                       grad ? 1 : 0);
}

int ewald_atom_split(const EwaldBatch& e) {
  // some 16k wavefronts: each group's tiles split into ksplit ranges
  return std::max(1, std::min(e.ntiles, (16384 + e.ngroups - 1) / std::max(e.ngroups, 1)));
}

void launch_ewald_atom(const EwaldBatch& e, bool grad, const DevWork& w, const EwaldWork& ew, hipStream_t st) {
  const int ksplit = ewald_atom_split(e);
  const dim3 grid(blocks_for(static_cast<long long>(e.ngroups) * ksplit, kWavesPerBlock));
  if (grad)
    hipLaunchKernelGGL(ewald_atom_kernel<true>, grid, dim3(kBlock), 0, st, e, ew.lc6l, ew.ldc6l, ew.etab, ew.F,
                       ksplit, w.e_atom, w.dedcn, w.grad);
  else
    hipLaunchKernelGGL(ewald_atom_kernel<false>, grid, dim3(kBlock), 0, st, e, ew.lc6l, ew.ldc6l, ew.etab, ew.F,
                       ksplit, w.e_atom, w.dedcn, w.grad);
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

void launch_hessian(const DevBatch& b, const HessWork& h, const long long* offset, const void* slots,
                    long long total_slots, int max_natoms, hipStream_t st) {
  const dim3 wgrid(blocks_for(b.natoms, kWavesPerBlock)), blk(kBlock);
  hipLaunchKernelGGL(hess_weights_kernel, dim3(blocks_for(b.natoms, kBlock)), dim3(kBlock), 0, st, b, h.cn, h.gw,
                     h.gwd, h.gwd2);
  hipLaunchKernelGGL(hess_c6_kernel, dim3(b.natoms), dim3(kBlock), 0, st, b, h.gw, h.gwd, h.gwd2, h.c6, h.dc6,
                     h.d2c6, h.d2c6ij);
  hipLaunchKernelGGL(hess_dcndr_kernel, wgrid, blk, 0, st, b, h.dr);
  hipLaunchKernelGGL(hess2_kernel, wgrid, blk, 0, st, b, h.c6, h.dc6, h.d2c6, h.d2c6ij, h.hess, h.dedcn, h.dEdcndr,
                     h.dEdcndcn);
  if (total_slots > 0)
    hipLaunchKernelGGL(hess3_kernel, dim3(blocks_for(total_slots, kWavesPerBlock)), dim3(kBlock), 0, st, b, offset,
                       static_cast<const Slot<double>*>(slots), h.c6, h.dc6, h.d2c6, h.d2c6ij, h.hess, h.dedcn,
                       h.dEdcndr, h.dEdcndcn);
  hipLaunchKernelGGL(hess_cn_kernel, wgrid, blk, 0, st, b, h.dedcn, h.hess);
  const int ndim = 3 * max_natoms;
  hipLaunchKernelGGL(hess_tmp_kernel, dim3(blocks_for(ndim, kHT), blocks_for(max_natoms, kHT), b.nsys),
                     dim3(kHT, kHT), 0, st, b, h.dEdcndcn, h.dr, h.tmp);
  hipLaunchKernelGGL(hess_assemble_kernel, dim3(blocks_for(ndim, kHT), blocks_for(ndim, kHT), b.nsys),
                     dim3(kHT, kHT), 0, st, b, h.dEdcndr, h.dr, h.tmp, h.hess);
}

}  // namespace d3
