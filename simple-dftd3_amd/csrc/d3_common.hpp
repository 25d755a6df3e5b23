// Shared definitions of the D3 library: element tables, per-system
// descriptors and the device data layout.
#pragma once

#include <cstddef>
#include <cstdint>

namespace d3 {

constexpr int kMaxElem = 103;   // elements with reference data (reference.f90:27)
constexpr int kMaxRef = 7;      // reference systems per element (reference.f90:28)
constexpr int kNumPairs = kMaxElem * (kMaxElem + 1) / 2;
constexpr int kMaxSpecies = kMaxElem; // distinct elements in one batch: any set of elements with reference data
// Fixed device work area of the allocation-free entry point, per GPU.
constexpr size_t kNoAllocBytes = size_t(64) << 20;

// Damping families of the two-body term (the three-body term is the same
// for all): rational (Becke-Johnson; the modified rational parameters use it
// too), zero, modified zero, optimized power, CSO and Z damping.
enum DampingKind : int {
  kRational = 0, kZero = 1, kModifiedZero = 2, kOptimizedPower = 3, kCSO = 4, kZDamping = 5
};
enum Precision : int { kFP64 = 0, kMixed = 1 };

// Damping parameters and cutoffs of one system (param.f90:38-48 d3_param,
// the *_damping_param types of damping/*.f90, cutoff.f90:42-66).  For the
// zero and modified zero damping a1/a2 hold rs6/rs8; bet is the modified
// zero / optimized power exponent, a3/a4 the CSO parameters (rs6/rs8 of
// d3_param).
struct SysParam {
  int kind;
  int grad;
  int ewald;  // two-body term in reciprocal space (the real-space pass adds nothing)
  double s6, s8, s9, a1, a2, alp, bet, a3, a4;
  double cut_cn, cut2, cut3, width2, width3;
  // externally managed work partition (partition.f90:57-148): this system's
  // pair, triple and reciprocal-space loops keep only the terms of part
  // `part` of `nparts` (nparts = 1: every term)
  int part, nparts;
};

// One term pref r^m / (r^alpha + radius^alpha) of a damped pair potential
// (fourier/kernel.f90:42-58); the rational and zero damping are sums of two.
struct FourierTerm {
  double pref, radius;
  int m, alpha;
};
constexpr int kMaxEwaldSpecies = 8;  // species whose structure factors the reciprocal-space kernels keep in registers (more: a generic instantiation)

// Atom range and lattice-image ranges of one system of a batch.  Images are
// stored per cutoff (CN, two-body, three-body) exactly as get_lattice_points
// enumerates them (cutoff.f90:140-186), so the pair and triple sets are the
// reference's.
enum ImageSet { kImgCn = 0, kImg2 = 1, kImg3 = 2 };

struct SysGeom {
  int atom_begin, natoms;
  int img_cn_begin, img_cn_count;
  int img2_begin, img2_count;
  int img3_begin, img3_count;
  long long c6_begin;  // offset of the n x n C6 block of this system
  // tiles of 64 consecutive positions of the system's spatial order, the
  // first at tile_begin of the batch's tile list; cull = 1 (molecules of
  // more than 8 tiles): pair loops skip a tile whose bounding box is beyond
  // the cutoff
  int tile_begin, cull;
  // 3D-periodic systems: the images of set s (kImgCn, kImg2, kImg3) form the
  // index box |n_d| <= box[s][d] in list order (n_x outer, n_z inner); the
  // rows of frac (b_d / 2 pi) give fractional coordinates and span[s][d] =
  // cut_s |b_d| / 2 pi.  box[s][0] < 0: no box, every image is visited.
  int box[3][3];
  double frac[9];
  double span[3][3];
};

// Host-side element tables (generated from data/d3_tables.npz).
struct ElementTables {
  int nref[kMaxElem];
  double refcn[kMaxElem][kMaxRef];
  const double* c6ref;  // [kNumPairs][kMaxRef][kMaxRef], [ic][jref][iref]
  double r4_over_r2[118];
  double vdwrad_aa[kNumPairs];
  double covrad_aa[118];
};

const ElementTables& element_tables();

// Angstrom -> bohr (mctc-lib): Bohr radius hbar / (m_e c alpha), CODATA 2018.
double aatoau();

double covalent_radius(int z);           // 4/3 * Pyykkoe radius, bohr
double r4r2_value(int z);                // sqrt(0.5 r4/r2 sqrt(Z)) (r4r2.f90:91-105)
double vdw_radius(int z1, int z2);       // pairwise radius, bohr (vdwrad.f90:1163-1197)
double reference_c6(int iref, int jref, int zi, int zj);  // get_c6, 0-based refs

}  // namespace d3
