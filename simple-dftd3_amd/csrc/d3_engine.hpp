// Batch engine of the D3 library: plans a batch of systems on the host and
// runs the energy / gradient / virial kernels on one HIP stream.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "d3_common.hpp"
#include "d3_kernels.hpp"

namespace d3 {

// A failure that the C API reports through its error handle.
struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

void hip_check(hipError_t err, const char* what);

// Low-rank expansion of the reference C6 coefficients of a structure's
// species (fourier/decomposition.f90:35-97): C6ref(si ri, sj rj) ~
// sum_l lambda_l vec(si, ri, l) vec(sj, rj, l) over the kept eigenpairs.
struct LowRank {
  int rank = 0;                 // kept eigenpairs
  int ndim = 0;                 // size of the reference C6 matrix
  double error = 0.0;           // largest relative error of the kept expansion
  double kcut = 0.0;            // reciprocal cutoff, 0: from the damping radii
  int mesh = 0;                 // < 0: direct sum over the vectors; 0: particle mesh from the cutoff; > 0: mesh points
  std::vector<int> species;     // atomic numbers in first-appearance order
  std::vector<double> lambda;   // [rank]
  std::vector<double> vec;      // [species][kMaxRef][rank]
  // the last reciprocal-space evaluation: its cutoff and number of vectors
  // (one of each +-k pair and the zero vector)
  mutable double last_kcut = 0.0;
  mutable long long last_vectors = 0;
  double reference_c6(int si, int ri, int sj, int rj) const;
};

// The expansion of these species with a fixed rank (> 0) or the smallest
// rank within the tolerance (decomposition.f90:104-175).
LowRank make_lowrank(const std::vector<int>& species, int rank, double tolerance);

bool supports_ewald(const SysParam& p);
// Terms of the pair potential of elements zi, zj (0 or 2; rational and zero
// damping only).
int fourier_terms(const SysParam& p, int zi, int zj, FourierTerm terms[2]);
bool fourier_supported(const FourierTerm& t);
double fourier_potential_zero(const FourierTerm& t);
double fourier_cutoff(const FourierTerm& t, double tolerance);
// Reciprocal lattice vectors within kcut, one of each +-k pair (and k = 0),
// as tiles {h, k, l0, count} of up to kWave vectors h b1 + k b2 + l b3 with
// consecutive l; the reciprocal vectors b (rows), the index range |n_d| <=
// rep[d] and the cell volume.
struct ReciprocalTiles {
  double rec[9] = {0};
  int rep[3] = {0, 0, 0};
  double volume = 0.0;
  long long vectors = 0;
  std::vector<int4> tiles;
};
ReciprocalTiles reciprocal_tiles(const double lattice[9], double kcut);
// Particle mesh of a cell: points per lattice direction, reciprocal basis,
// volume and Euler factors [3][mmax].
struct SpmeMesh {
  int m[3] = {1, 1, 1};
  int mmax = 1;
  double rec[9] = {0};
  double volume = 0.0;
  std::vector<double> bfac;
};
SpmeMesh spme_mesh(const double lattice[9], double kcut, int mesh_req);
void bspline6_host(double w, double* theta, double* dtheta);

// One system of a batch, as the host describes it.
struct SystemDesc {
  int natoms = 0;
  double lattice[9] = {0};  // rows = lattice vectors (bohr)
  int periodic[3] = {0, 0, 0};
  SysParam param{};
  // Low-rank model (model.f90:185-196, 379-382): every C6 from the
  // expansion; with 3D periodicity the two-body term is summed over the
  // reciprocal lattice (disp.f90:104-111).  A batch with a low-rank system
  // holds only that system.
  const LowRank* lowrank = nullptr;
  // Externally managed work partition (partition.f90:57-148): only the
  // pair, triple and reciprocal-space terms of part `part` of `nparts`
  // (disp.f90:141-175); summing the results of all parts gives the whole.
  int part = 0;
  int nparts = 1;
};

// Outputs of a batch run (device pointers; any may be null except energy).
struct BatchOutput {
  double* energy = nullptr;    // [nsys]
  double* gradient = nullptr;  // [natoms][3]
  double* sigma = nullptr;     // [nsys][9], sigma(a,b) at [b*3+a]
};

// Optional per-atom intermediates for testing (device pointers).
struct BatchDebug {
  double* cn = nullptr;     // [natoms]
  double* dedcn = nullptr;  // [natoms]
};

struct RunStats {
  long long atm_list_entries = 0;  // size of the three-body neighbour lists
};

// Lattice points within a cutoff, in the order of cutoff.f90:140-186.
std::vector<double> lattice_points(const int periodic[3], const double lattice[9], double rthr);

struct Workspace;

// Allocation-free evaluation of a batch: as Engine::run (real-space path, no
// low-rank model, no ghost atoms), with every work buffer taken from the
// library's fixed work area of the current GPU (kNoAllocBytes), metadata
// copied from pageable host memory, and no allocation, free, event or plan
// creation.  Throws Error, before any output is written, when the batch needs
// more than the area.  Calls that overlap on one GPU race: they share its
// area, which keeps the contents of the last call.
void run_noalloc(const std::vector<SystemDesc>& systems, const int* numbers_host, const double* positions_dev,
                 int precision, const BatchOutput& out, hipStream_t stream);

class Engine {
 public:
  Engine();
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  // Evaluate a batch.  numbers are host arrays (atomic numbers, 1..103),
  // positions a device array [natoms][3] in bohr.  Systems are consecutive
  // in the atom arrays.  ghost_host (optional, host, [natoms]) marks atoms
  // whose C6 coefficients vanish (model.f90:166-174, 436-440).  The size of
  // the three-body lists is read back once per call.
  void run(const std::vector<SystemDesc>& systems, const int* numbers_host,
           const double* positions_dev, int precision, const BatchOutput& out,
           hipStream_t stream, const BatchDebug* debug = nullptr, RunStats* stats = nullptr,
           const int* ghost_host = nullptr);

  // Pairwise-resolved energies (disp.f90:403-457) of a batch: pair2 and
  // pair3 (device) hold per system an n x n block in atom order, the systems
  // one after another (offset sum of n^2 of the systems before); pair2 is the
  // two-body (additive) and pair3 the three-body (non-additive) energy of
  // every ordered pair.  Both blocks are symmetric.  The size of the
  // three-body lists is read back once per call.
  void run_pairwise(const std::vector<SystemDesc>& systems, const int* numbers_host, const double* positions_dev,
                    int precision, double* pair2, double* pair3, hipStream_t stream,
                    const int* ghost_host = nullptr);

  // Energy and analytic Hessian (disp.f90:236-308) of a batch, FP64:
  // energy [nsys] (device, may be null) and hessian (device) holding per
  // system a 3n x 3n block, [3i+a][3j+b] = d2E / dR_ia dR_jb, the blocks one
  // after another (offset 9 times the sum of n^2 of the systems before).
  void run_hessian(const std::vector<SystemDesc>& systems, const int* numbers_host, const double* positions_dev,
                   double* energy, double* hessian, hipStream_t stream, const int* ghost_host = nullptr);

  // Work units of the reference's loops per system: counts_host[3*s + 0..2]
  // = CN terms, two-body pair terms, ATM triple terms (host array).
  void count_terms(const std::vector<SystemDesc>& systems, const int* numbers_host,
                   const double* positions_dev, long long* counts_host, hipStream_t stream);

  // Per-kernel timing with HIP events on the launch stream (resets totals).
  void set_kernel_timing(bool enable);
  // Total milliseconds and number of launches of a kernel group since the
  // timing was enabled: pack, order, cn, weights, c6, disp2, lowrank,
  // ewald_sf, ewald_f, ewald_atom, atm_list, atm, cn_chain, reduce, hessian.
  std::pair<double, long long> kernel_time(const std::string& name) const;

 private:
  Workspace* ws_;
};

// A batch planned once and evaluated many times with new coordinates of the
// same systems: the metadata upload, the buffer sizes and the launch shape
// are fixed when the plan is made, and an evaluation is a fixed sequence of
// kernel launches on one stream without host synchronisation or allocation,
// which can be captured in a HIP graph and replayed.
//
// Three-body lists: a batch of molecules gets the exact bound n(n+1)/2 slots
// per molecule; a periodic system gets 1.5 times the list size of the
// coordinates the plan is made with (and at least 64 more slots per atom of a
// system with a three-body term),
// capped by the exact bound.  An evaluation whose lists do not fit adds no
// three-body terms, returns NaN energies, gradients and virials for every
// system with a three-body term, and raises the overflow flag that
// overflowed() reports.
class Plan {
 public:
  Plan(const std::vector<SystemDesc>& systems, const int* numbers_host, const double* positions_dev,
       int precision, bool grad, const int* ghost_host, hipStream_t stream);
  ~Plan();
  Plan(const Plan&) = delete;
  Plan& operator=(const Plan&) = delete;

  int natoms() const;
  int nsys() const;
  long long slot_capacity() const;
  const AtmShape& shape() const;
  void set_shape(int waves, int blocks_per_cu);

  // One evaluation (positions: device [natoms][3], bohr).
  void execute(const double* positions_dev, const BatchOutput& out, hipStream_t stream);

  // Candidate launch shapes of the three-body kernel.
  static std::vector<AtmShape> candidate_shapes();
  // Times every candidate shape (one warm-up and `reps` timed evaluations
  // each, HIP events on `stream`), keeps the fastest and returns its index;
  // times_ms (optional) receives the mean time per evaluation of each.
  int autotune(const double* positions_dev, const BatchOutput& out, hipStream_t stream, int reps,
               std::vector<float>* times_ms);

  // Records one evaluation with these device pointers into a HIP graph;
  // replay() launches it (new coordinates go into the same positions buffer).
  void capture(const double* positions_dev, const BatchOutput& out, hipStream_t stream);
  bool captured() const;
  void replay(hipStream_t stream);

  // Synchronises `stream`; true when an evaluation since the plan was made
  // had more three-body slots than the plan holds.
  bool overflowed(hipStream_t stream);

 private:
  struct State;
  State* st_;
};

}  // namespace d3
