// Device data of a batch and the launch wrappers of the D3 kernels
// (implemented in d3_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "d3_common.hpp"

namespace d3 {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kTile = 64;  // slots per three-body tile

// Device view of one prepared batch (pointers into device memory).
struct DevBatch {
  int natoms, nsys, nsp, mref;
  const double4* pos;   // [N] (x, y, z, 0)
  const int* sp;        // [N] species index
  const int* sys;       // [N] system index
  const SysGeom* geom;  // [nsys]
  const SysParam* par;  // [nsys]
  const double4* img;   // lattice translations of all systems and cutoffs
  const double* rcov;   // [nsp]
  const double* r4r2;   // [nsp]
  const int* nref;      // [nsp]
  const double* refcn;  // [nsp][kMaxRef]
  const double* c6ref;  // [nsp*nsp][kMaxRef*kMaxRef], [(si*nsp+sj)][ri*7+rj]
  const double* rvdw;   // [nsp*nsp]
  const int* ghost;     // [N] or null
  const int* sid;       // [N] species identifier within the atom's system: its element numbered
                        // 1, 2, ... by first appearance (structure_type%id; Z damping)
  // Spatial (Morton) order of the atoms of each system, the storage order of
  // the pair records and of the three-body lists (no result depends on it):
  const int* rank;      // [N] position of the atom in its system's order
  const int* order;     // [N] atom at position atom_begin + r of that order
  const double4* posr;  // [N] positions in that order, posr[x] = pos[order[x]] (refreshed with pos)
  int ntiles;
  const int2* tiles;    // [ntiles] first position and count of each 64-position tile of that order
  const double4* tbox;  // [2 ntiles] lower and upper corner of each tile's positions (refreshed with pos)
};

// Per-evaluation device arrays.
struct DevWork {
  double* cn;       // [N]
  double* gw;       // [N][mref]
  double* dgw;      // [N][mref] or null (energy only)
  void* pair;       // per system n x n block of R4 {sqrt C6, dC6/dCN_i / C6, dC6/dCN_j / C6, r0^(alp/3)},
                    // record (i, j) at [rank_i][rank_j]
  double* e_atom;   // [N]
  double* dedcn;    // [N]
  double* grad;     // [N][3]
  double* sig_atom; // [N][6]
  double4* atm_acc; // [N] three-body slot sums per atom in spatial order (G xyz, D), E+grad only
  double* gwr;      // [2 mref][N] weights and dw/dCN in spatial order (pair records)
  int* spr;         // [N] species | ghost flag in spatial order
};

// Size in bytes of one pair record and one three-body slot.
size_t pair_record_bytes(bool mixed);
size_t slot_record_bytes(bool mixed);

// Launchers (all asynchronous on `stream`).
void launch_pack(const double* xyz, int n, double4* pos, hipStream_t stream);
// Spatial order (rank, order of DevBatch) from the packed positions: Morton
// codes of 4-bohr cells within each system, sorted on the device.  temp is
// scratch of spatial_order_temp_bytes(n) bytes.
size_t spatial_order_temp_bytes(int n);
void launch_spatial_order(const DevBatch& b, void* temp, size_t temp_bytes, int* rank, int* order,
                          hipStream_t stream);
void launch_cn(const DevBatch& b, bool mixed, double* cn, hipStream_t stream);
void launch_weights(const DevBatch& b, const double* cn, double* gw, double* dgw, hipStream_t stream);
// posr[x] = pos[order[x]] for every position x of the batch's spatial order,
// and the bounding box of every 64-position tile of that order (tbox).
void launch_rank_pos(const DevBatch& b, hipStream_t stream);
// gwr [(dgw ? 2 : 1) * mref][natoms] and spr [natoms]: scratch for the
// weights in spatial order.
void launch_pair(const DevBatch& b, bool mixed, const double* gw, const double* dgw, double* gwr, int* spr,
                 void* pair, hipStream_t stream);
// pair2 (energy only, may be null): per system an n x n block (offset
// c6_begin, atom order) of pairwise two-body energies.
void launch_disp2(const DevBatch& b, bool mixed, bool grad, const DevWork& w, double* pair2, hipStream_t stream);
// Three-body lists: count per atom (pair may be null: no C6 test), fill slots.
// The count pass also clears acc (DevWork::atm_acc, may be null).
void launch_atm_count(const DevBatch& b, const void* pair, bool mixed, long long* count, double4* acc,
                      hipStream_t stream);
// offset[0..n] and item_off[0..n]: exclusive scans of the list sizes and of
// their 64-slot tile counts (one block).
void launch_list_scan(const long long* count, int n, long long* offset, long long* item_off, hipStream_t stream);
// Writes nothing when the lists hold more than `capacity` slots.
void launch_atm_fill(const DevBatch& b, const void* pair, bool mixed, const long long* offset, void* slots,
                     long long capacity, hipStream_t stream);
// When the lists hold more than `capacity` slots: sets *overflow = 1,
// empties the work-item range (item_off[natoms] = 0), so no triple is
// evaluated, and writes NaN into the energy, dE/dCN, gradient and virial
// accumulators of w (the last three with grad) of every atom of a system with
// a three-body term.  Runs after the two-body term.
void launch_atm_guard(const DevBatch& b, const long long* offset, long long capacity, long long* item_off,
                      int* overflow, const DevWork& w, bool grad, hipStream_t stream);

// Launch shape of the three-body kernel: wavefronts per block (1, 2 or 4, the
// block shares one 64-slot tile of a list) and either one block per work
// item (blocks_per_cu = 0) or a persistent grid of blocks_per_cu blocks per
// compute unit drawing items from a work counter.  The engine takes waves = 0
// as "choose from the number of work items".
struct AtmShape {
  int waves = 4;
  int blocks_per_cu = 0;
  int num_cus = 0;
};
// smooth: some system has a nonzero three-body switching width.  max_items
// bounds the number of work items (item_off[natoms]); counter is device
// memory for the work counter of a persistent shape.
// partitioned: some system has a work partition of more than one part.
void launch_atm(const DevBatch& b, bool mixed, bool grad, bool smooth, bool partitioned, const AtmShape& shape,
                const long long* item_off, long long max_items, const long long* offset, const void* slots,
                const DevWork& w, unsigned long long* counter, hipStream_t stream);
// Pairwise three-body energies: adds the pair sums of every triple term to
// pair3s (per system an n x n block in storage order, zeroed by the caller);
// pair3_finalize writes the symmetric pair energies in atom order to out.
void launch_atm_pairs(const DevBatch& b, bool mixed, bool smooth, const long long* item_off, long long max_items,
                      const long long* offset, const void* slots, const DevWork& w, double* pair3s,
                      hipStream_t stream);
void launch_pair3_finalize(const DevBatch& b, const double* pair3s, double* out, hipStream_t stream);
// Analytic Hessian (FP64) of a batch; per system, atom order within the
// system, offsets from c6_begin: c6, dc6, d2c6, d2c6ij and dEdcndcn n x n at
// c6_begin; dr and dEdcndr 3n x n and tmp n x 3n at 3 c6_begin; hess 3n x 3n
// at 9 c6_begin.  hess, dedcn, dEdcndr and dEdcndcn must be zeroed; the
// FP64 three-body lists (offset, slots; total_slots entries) must be filled
// and cn computed.  max_natoms: largest system of the batch.
struct HessWork {
  const double* cn;
  double *gw, *gwd, *gwd2;
  double *c6, *dc6, *d2c6, *d2c6ij;
  double *dr, *dedcn, *dEdcndr, *dEdcndcn, *tmp, *hess;
};
void launch_hessian(const DevBatch& b, const HessWork& h, const long long* offset, const void* slots,
                    long long total_slots, int max_natoms, hipStream_t stream);
// Reciprocal-space two-body term of one 3D-periodic system with a low-rank
// model (fourier/ewald.f90:47-252): one vector of each +-k pair within the
// cutoff, in tiles {h, k, l0, count} of up to kWave vectors h b1 + k b2 +
// l b3 (flat vector index tile * kWave + lane); the low-rank expansion in
// batch species order; the atoms of the system grouped by species, cut into
// chunks (structure factors) and groups of up to kEwaldGroup (per-atom pass).
constexpr int kEwaldGroup = 4;
struct EwaldBatch {
  int ntiles, rank, rank_pad, nsp, natoms, atom_begin, nchunks, ngroups;  // rank_pad: rank rounded up to 8
  double inv_vol;
  double rec[9];             // b_1, b_2, b_3 (rows)
  int rep0, rep1, rep2;      // |h| <= rep0, |k| <= rep1, |l| <= rep2
  int tstride, toff0, toff1, toff2;  // phase table row: direction d at toff_d + rep_d + n
  const int4* tiles;         // [ntiles]
  const double* lambda;      // [rank]
  const double* vec;         // [nsp][kMaxRef][rank]
  const double* pval;        // [nsp] pair potential of (a, a) at r = 0
  const FourierTerm* terms;  // [nsp][nsp][2]
  const int* nterm;          // [nsp][nsp]
  const int* sorder;         // [natoms] atoms by species (global indices)
  const int* spos;           // [natoms] position of each atom of the system in sorder
  const int4* chunks;        // [nchunks] {species, begin, end, 0} into sorder
  const int* chunk_seg;      // [nsp + 1] chunks of each species
  const int4* groups;        // [ngroups] {species, begin, end, 0} into sorder, end - begin <= kEwaldGroup
  // work partition: vectors by their index in get_lattice_points' box
  // (direct sum, self-interaction on part 0) or expansion terms l (mesh)
  int part, nparts, mesh;
};
struct EwaldWork {
  // rows and columns by position in species order (sorder)
  double2 *etab;             // [natoms][tstride] exp(i n b_d.x)
  double *c6l;               // [natoms][rank_pad], zero padded
  double *lc6l, *ldc6l;      // [rank][natoms + kEwaldGroup]: lambda c6l, lambda dc6l (null without gradient)
  double2 *part;             // [nchunks][rank][nk] partial structure factors (nk = ntiles * kWave)
  double2 *phi;              // [nsp (nsp + 1) / 2][nk] transform and its k-derivative
  double2 *F;                // [nsp][rank][nk] sum_b phihat_ab S_bl
  double2 *S;                // [nsp][rank][nk] S_al, only for more than kMaxEwaldSpecies species
};
// The reciprocal-space energy (and dE/dCN, gradient, virial), added to the
// accumulators of w after launch_disp2, in this order: the phase tables (if
// ew.etab) and low-rank factors (and the self-interaction), the partial structure
// factors, the per-vector transforms F and virial, the per-atom sums.
void launch_lowrank(const DevBatch& b, const EwaldBatch& e, bool grad, const DevWork& w, const EwaldWork& ew,
                    hipStream_t stream);
void launch_ewald_sf(const EwaldBatch& e, const EwaldWork& ew, hipStream_t stream);
void launch_ewald_f(const EwaldBatch& e, bool grad, const DevWork& w, const EwaldWork& ew, hipStream_t stream);
void launch_ewald_atom(const EwaldBatch& e, bool grad, const DevWork& w, const EwaldWork& ew, hipStream_t stream);
// Particle-mesh evaluation of the reciprocal sum (fourier/spme.f90): the
// mesh [m2][m1][m0] (g1 fastest), its half-complex transform [m2][m1][mh]
// (mh = m0 / 2 + 1), order-6 B-splines; the expansion, species and atom order
// of the EwaldBatch (no tiles).
constexpr int kSplineOrder = 6;  // spme.f90:47
struct SpmeBatch {
  int m0, m1, m2, mh, mmax;
  long long ng, nch;         // mesh points, half-complex points
  double inv[9];             // fractional coordinates f_d = inv[3 d ..] . x (rows b_d / 2 pi)
  double rec[9];             // b_1, b_2, b_3 (rows)
  double dudr[9];            // du_d / dx_c = m_d b_d[c] / 2 pi at [3 d + c]
  double inv_vol;
  const double* bfac;        // [3][mmax] Euler factors (spme.f90:503-529)
};
struct SpmeWork {
  int* base;                 // [natoms][3]
  double *theta, *dtheta;    // [natoms][3][kSplineOrder]
  double* Q;                 // [nsp][ng] spread factors of one term (also the back transform)
  double2 *Qh, *Th;          // [nsp][nch] transforms, kernel products
  double* T;                 // [nsp][ng] back-transformed meshes (may alias Q)
  double *K, *Dm, *Dmir;     // [nsp (nsp + 1) / 2][nch] kernel and virial weights (Dm, Dmir with gradient)
};
// Spline weights of an evaluation and (with table) the kernel table, which
// depends on the cell, the parameters and the mesh only; then per term l:
// spread into Q (zeroed by the caller), [Q -> Qh transform], contract
// Qh -> Th (and the term's virial), [Th -> T back transform], gather.
void launch_spme_setup(const DevBatch& b, const EwaldBatch& e, const SpmeBatch& s, bool grad, bool table,
                       const SpmeWork& sw, hipStream_t stream);
void launch_spme_spread(const DevBatch& b, const EwaldBatch& e, const SpmeBatch& s, const EwaldWork& ew,
                        const SpmeWork& sw, int l, hipStream_t stream);
void launch_spme_conv(const EwaldBatch& e, const SpmeBatch& s, bool grad, const DevWork& w, const SpmeWork& sw, int l,
                      hipStream_t stream);
void launch_spme_gather(const DevBatch& b, const EwaldBatch& e, const SpmeBatch& s, bool grad, const DevWork& w,
                        const EwaldWork& ew, const SpmeWork& sw, int l, hipStream_t stream);
void launch_cn_chain(const DevBatch& b, bool mixed, const DevWork& w, hipStream_t stream);
void launch_reduce(const DevBatch& b, const DevWork& w, double* energy, double* sigma, hipStream_t stream);
void launch_count_pairs(const DevBatch& b, unsigned long long* counts, hipStream_t stream);
void launch_count_triples(const DevBatch& b, bool mixed, const long long* offset, const void* slots,
                          unsigned long long* counts, hipStream_t stream);

// The allocation-free entry point's work area: a static device array of
// kNoAllocBytes per GPU (held from the loading of the library to the end of
// the process).  Returns its device address on the current device (looked
// up once per device with hipGetSymbolAddress); hip_err receives the HIP
// status of the lookup.
void* noalloc_area(int device, hipError_t* hip_err);

}  // namespace d3
