// Host side of the batch engine: plans a batch (species tables, lattice
// images, C6 block offsets), owns the device work buffers and launches the
// kernels of d3_kernels.hip on one stream.
//
// Order of one evaluation (disp.f90:56-185):
//   pack -> cn -> weights -> pair (C6 and derivatives) -> disp2 ->
//   [reciprocal space: lowrank -> ewald_sf -> ewald_f -> ewald_atom, or on a
//    particle mesh: lowrank -> spme_setup -> per term (spread -> FFT -> conv ->
//    FFT -> gather)] ->
//   atm_list (count, scan, fill) -> atm -> cn_chain -> reduce
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "d3_engine.hpp"
#include "d3_kernels.hpp"

namespace d3 {

void hip_check(hipError_t err, const char* what) {
  if (err != hipSuccess) throw Error(std::string("HIP error in ") + what + ": " + hipGetErrorString(err));
}

namespace {

constexpr double kEps = 2.220446049250313e-16;

// A fixed region of device memory handed out front to back (256-byte
// aligned); reset() makes all of it free again.  Used by the allocation-free
// entry point, whose region is the library's static work area.
class Arena {
 public:
  void bind(void* base, size_t bytes) {
    base_ = static_cast<char*>(base);
    size_ = bytes;
    used_ = 0;
  }
  void reset() { used_ = 0; }
  void* take(size_t bytes) {
    const size_t at = (used_ + 255) / 256 * 256;
    if (at > size_ || bytes > size_ - at)
      throw Error("The batch needs more device memory than the fixed work area of the allocation-free entry point (" +
                  std::to_string(size_ >> 20) + " MiB)");
    used_ = at + bytes;
    return base_ + at;
  }

 private:
  char* base_ = nullptr;
  size_t size_ = 0, used_ = 0;
};

// Grow-only device allocation; with an arena, taken from it instead (no
// allocation call).
class DeviceBuffer {
 public:
  ~DeviceBuffer() {
    if (ptr_ && !arena_) (void)hipFree(ptr_);
  }
  void use_arena(Arena* a) {
    if (ptr_ && !arena_) hip_check(hipFree(ptr_), "hipFree");
    arena_ = a;
    ptr_ = nullptr;
    cap_ = 0;
  }
  // arena mode: forget the region (the arena is reset by its owner)
  void release() {
    if (arena_) {
      ptr_ = nullptr;
      cap_ = 0;
    }
  }
  void* get(size_t bytes) {
    if (bytes > cap_) {
      if (arena_) {
        ptr_ = arena_->take(bytes);
        cap_ = bytes;
        return ptr_;
      }
      if (ptr_) hip_check(hipFree(ptr_), "hipFree");
      ptr_ = nullptr;
      cap_ = 0;
      const size_t want = std::max<size_t>(bytes + bytes / 4, 256);
      hip_check(hipMalloc(&ptr_, want), "hipMalloc");
      cap_ = want;
    }
    return ptr_;
  }
  template <typename T>
  T* as(size_t count) {
    return static_cast<T*>(get(std::max<size_t>(count, 1) * sizeof(T)));
  }

 private:
  void* ptr_ = nullptr;
  size_t cap_ = 0;
  Arena* arena_ = nullptr;
};

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
      const size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
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
    hip_check(hipEventCreate(&open_), "hipEventCreate");
    hip_check(hipEventRecord(open_, st), "hipEventRecord");
  }
  void end(const char* name, hipStream_t st) {
    if (!enabled) return;
    hipEvent_t b;
    hip_check(hipEventCreate(&b), "hipEventCreate");
    hip_check(hipEventRecord(b, st), "hipEventRecord");
    pending_.push_back({name, open_, b});
    open_ = nullptr;
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

}  // namespace

struct Prepared {
  DevBatch db{};
  long long pair_total = 0;    // sum over systems of n^2
  bool any_atm = false;        // some system has s9 != 0
  bool smooth = false;         // some three-body system has a nonzero width3
  bool periodic_atm = false;   // some three-body system has lattice images
  bool partitioned = false;    // some system evaluates one part of a work partition
  bool fixed_order = false;    // one system of at most 8 tiles: storage order = atom order, no sort
  long long slot_bound = 0;    // sum over three-body systems of (il + 1) * images
  long long atm_atoms = 0;     // atoms of the systems with a three-body term
  long long item_bound = 0;    // work items of those bounds
  bool ewald = false;          // the (single) system sums its two-body term in reciprocal space
  EwaldBatch eb{};             // its vectors, expansion and species chunks (nk = 0: no two-body term)
  bool spme = false;           // ... on a particle mesh instead of the vectors
  SpmeBatch sb{};
  int spme_nsp = 0;
  std::vector<char> spme_key;  // what the kernel table depends on: mesh, basis, pair terms
};

namespace {

template <typename T>
size_t stage(std::vector<char>& buf, const std::vector<T>& v) {
  const size_t off = (buf.size() + 255) & ~size_t(255);
  buf.resize(off + v.size() * sizeof(T));
  if (!v.empty()) std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
  return off;
}

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
      const double len = std::sqrt(nrm[k][0] * nrm[k][0] + nrm[k][1] * nrm[k][1] + nrm[k][2] * nrm[k][2]);
      for (int c = 0; c < 3; ++c) nrm[k][c] = nrm[k][c] / len;
    }
    for (int k = 0; k < 3; ++k) {
      const double cs = nrm[k][0] * l[k][0] + nrm[k][1] * l[k][1] + nrm[k][2] * l[k][2];
      rep[k] = static_cast<int>(std::ceil(std::fabs(rthr / cs)));
    }
  };
  int rep[3] = {0, 0, 0};
  if (np == 3) {
    translations(lat, rep);
  } else if (np == 1) {
    for (int k = 0; k < 3; ++k) {
      const double len = std::sqrt(lat[k][0] * lat[k][0] + lat[k][1] * lat[k][1] + lat[k][2] * lat[k][2]);
      rep[k] = periodic[k] ? static_cast<int>(std::ceil(rthr / len)) : 0;
    }
  } else {
    double nrm[3][3];
    for (int k = 0; k < 3; ++k) cross(lat[(k + 1) % 3], lat[(k + 2) % 3], nrm[k]);
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c)
        if (periodic[c]) nrm[k][c] = lat[k][c];  // where(spread(periodic, 2, 3)) norms = lat
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

// The index box of one image set of a 3D-periodic system (SysGeom::box):
// the reciprocal rows, the repetitions read back from the images, and a check
// that the images are that box in list order (else no box is used).
void image_box(const SystemDesc& sd, double cut, const double4* imgs, int count, int set, SysGeom& g) {
  g.box[set][0] = g.box[set][1] = g.box[set][2] = -1;
  if (!(sd.periodic[0] && sd.periodic[1] && sd.periodic[2])) return;
  const double* a = sd.lattice;
  const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
  if (!(std::fabs(det) > 0.0)) return;
  double fr[9];
  for (int j = 0; j < 3; ++j) {  // b_j / 2 pi = (a_k x a_l) / det
    const double* u = a + 3 * ((j + 1) % 3);
    const double* v = a + 3 * ((j + 2) % 3);
    fr[3 * j + 0] = (u[1] * v[2] - u[2] * v[1]) / det;
    fr[3 * j + 1] = (u[2] * v[0] - u[0] * v[2]) / det;
    fr[3 * j + 2] = (u[0] * v[1] - u[1] * v[0]) / det;
  }
  int rep[3] = {0, 0, 0};
  std::vector<int> idx(3 * size_t(count));
  for (int t = 0; t < count; ++t)
    for (int d = 0; d < 3; ++d) {
      const double x = fr[3 * d] * imgs[t].x + fr[3 * d + 1] * imgs[t].y + fr[3 * d + 2] * imgs[t].z;
      idx[3 * t + d] = static_cast<int>(std::lround(x));
      rep[d] = std::max(rep[d], std::abs(idx[3 * t + d]));
    }
  if (count != (2 * rep[0] + 1) * (2 * rep[1] + 1) * (2 * rep[2] + 1)) return;
  for (int t = 0; t < count; ++t) {
    const int want = ((idx[3 * t] + rep[0]) * (2 * rep[1] + 1) + (idx[3 * t + 1] + rep[1])) * (2 * rep[2] + 1) +
                     (idx[3 * t + 2] + rep[2]);
    if (want != t) return;
  }
  for (int d = 0; d < 3; ++d) {
    g.box[set][d] = rep[d];
    g.span[set][d] = cut * std::sqrt(fr[3 * d] * fr[3 * d] + fr[3 * d + 1] * fr[3 * d + 1] + fr[3 * d + 2] * fr[3 * d + 2]);
  }
  std::memcpy(g.frac, fr, sizeof fr);
}

// Largest slot buffer (bytes) sized by the exact list bound of molecules
// instead of a read-back of the list size.
constexpr long long kUnsyncedSlotBytes = 64LL << 20;

int device_cu_count() {
  int dev = 0, cus = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
  return cus;
}

// Three-body launch shape of a call without a plan (waves == 0 asks for it).
// Below kLargeAtmItems work items (one 200-atom molecule has about 400) four
// wavefronts per block and one block per item, so that the few items spread
// over the chip.  Above, a persistent grid, whose smaller blocks leave fewer
// wavefronts idle at the end of an item: two wavefronts per block and 8
// blocks per CU, or, for lists with lattice images in mixed precision, one
// and 16.  Measured (whole evaluations, ms, this shape / four wavefronts and
// one block per item / best of the ten shapes): 2,000-atom cell mixed 13.08 /
// 14.13 / 13.08, FP64 19.01 / 20.51 / 19.01; 20,000-atom ellipsoid mixed
// 109.8 / 116.5 / 109.2, FP64 161.0 / 177.3 / 161.0; 1,000- and 2,000-atom
// molecules mixed 0.93 / 0.99 / 0.90 and 3.36 / 3.60 / 3.26.
constexpr long long kLargeAtmItems = 8192;
const AtmShape kPerCallShape{0, 0, 0};

AtmShape per_call_shape(bool mixed, bool periodic, long long items) {
  AtmShape a;
  if (items >= kLargeAtmItems) {
    const bool one = mixed && periodic;
    a.waves = one ? 1 : 2;
    a.blocks_per_cu = one ? 16 : 8;
    a.num_cus = device_cu_count();
  }
  return a;
}

struct Workspace {
  DeviceBuffer meta, pos, rank, order, sort_tmp, cn, gw, dgw, pair, e_atom, dedcn, grad, sig, atm_acc, cnt, off,
      item_off, slots, gw_rank, sp_rank, pos_rank, tbox, counts, counter, flag, pair3s, hbuf, e_out, ew_tab, ew_c6l, ew_lc6l, ew_ldc6l, ew_part,
      ew_phi, ew_f, ew_s, sp_base, sp_theta, sp_dtheta, sp_q, sp_qh, sp_th, sp_k, sp_dm, sp_dmir;
  // batched real <-> half-complex 3D transforms of the particle mesh (one per
  // species), kept while the mesh and the species count stay the same
  hipfftHandle fft_fwd = 0, fft_inv = 0;
  int fft_dims[4] = {0, 0, 0, 0};
  std::vector<char> spme_key;  // the kernel table in sp_k / sp_dm / sp_dmir
  bool spme_key_grad = false;  // ... with its virial weights
  ~Workspace() {
    if (fft_fwd) (void)hipfftDestroy(fft_fwd);
    if (fft_inv) (void)hipfftDestroy(fft_inv);
  }
  static void fft_check(hipfftResult r, const char* what) {
    if (r != HIPFFT_SUCCESS) throw Error(std::string("hipFFT error in ") + what + ": " + std::to_string(int(r)));
  }
  void fft_plans(const SpmeBatch& q, int nsp, hipStream_t stream) {
    const int want[4] = {q.m0, q.m1, q.m2, nsp};
    if (!fft_fwd || std::memcmp(want, fft_dims, sizeof want) != 0) {
      if (fft_fwd) fft_check(hipfftDestroy(fft_fwd), "hipfftDestroy");
      if (fft_inv) fft_check(hipfftDestroy(fft_inv), "hipfftDestroy");
      fft_fwd = fft_inv = 0;
      int n[3] = {q.m2, q.m1, q.m0};  // row-major, g1 fastest
      fft_check(hipfftPlanMany(&fft_fwd, 3, n, nullptr, 1, 0, nullptr, 1, 0, HIPFFT_D2Z, nsp), "hipfftPlanMany");
      fft_check(hipfftPlanMany(&fft_inv, 3, n, nullptr, 1, 0, nullptr, 1, 0, HIPFFT_Z2D, nsp), "hipfftPlanMany");
      std::memcpy(fft_dims, want, sizeof want);
    }
    fft_check(hipfftSetStream(fft_fwd, stream), "hipfftSetStream");
    fft_check(hipfftSetStream(fft_inv, stream), "hipfftSetStream");
  }
  std::vector<char> staging;
  PinnedStaging pinned;
  KernelTimer timer;
  // allocation-free mode: every device buffer comes from the arena, the
  // metadata is copied from pageable host memory (no pinned staging)
  Arena* arena = nullptr;
  std::vector<DeviceBuffer*> buffers() {
    return {&meta, &pos, &rank, &order, &sort_tmp, &cn, &gw, &dgw, &pair, &e_atom, &dedcn, &grad, &sig, &atm_acc, &cnt, &off,
            &item_off, &slots, &gw_rank, &sp_rank, &pos_rank, &tbox, &counts, &counter, &flag, &pair3s, &hbuf,
            &e_out, &ew_tab, &ew_c6l, &ew_lc6l, &ew_ldc6l, &ew_part, &ew_phi, &ew_f, &ew_s, &sp_base, &sp_theta,
            &sp_dtheta, &sp_q, &sp_qh, &sp_th, &sp_k, &sp_dm, &sp_dmir};
  }
  void use_arena(Arena* a) {
    arena = a;
    for (DeviceBuffer* b : buffers()) b->use_arena(a);
  }
  // start of an allocation-free call: the whole arena is free again
  void reset_arena() {
    arena->reset();
    for (DeviceBuffer* b : buffers()) b->release();
  }

  template <typename F>
  void timed(const char* name, hipStream_t st, F&& launch) {
    timer.begin(st);
    launch();
    timer.end(name, st);
  }

  // Host planning of a batch: species of the batch, per-species tables,
  // lattice images per system and cutoff, pair-block offsets; one metadata
  // upload (stream-ordered).
  Prepared upload(const std::vector<SystemDesc>& systems, const int* numbers, hipStream_t stream,
                  const int* ghost_host);
  // Packs the positions; with order_atoms also (re)computes the spatial order
  // of the atoms (the storage order of pair records and three-body lists).
  void pack(const Prepared& pb, const double* positions_dev, hipStream_t stream, bool order_atoms) {
    timed("pack", stream,
          [&] { launch_pack(positions_dev, pb.db.natoms, const_cast<double4*>(pb.db.pos), stream); });
    if (order_atoms && !pb.fixed_order) {
      const size_t bytes = spatial_order_temp_bytes(pb.db.natoms);
      void* t = sort_tmp.get(bytes);
      timed("order", stream, [&] {
        launch_spatial_order(pb.db, t, bytes, const_cast<int*>(pb.db.rank), const_cast<int*>(pb.db.order), stream);
      });
    }
    timed("pack", stream, [&] { launch_rank_pos(pb.db, stream); });
  }
  DevWork work(const Prepared& pb, bool mixed, bool grad, const BatchOutput& out, const BatchDebug* debug);
  // count -> scans of the list sizes and of their work items (three-body lists of every atom)
  void list_offsets(const Prepared& pb, const void* pair, bool mixed, double4* acc, hipStream_t stream);
  // pack .. disp2; then the three-body lists and kernel; then the chain
  // rule and the reduction.  capacity < 0: read the list size back and size
  // the slot buffer to it; otherwise use the fixed capacity and item bound.
  void evaluate(const Prepared& pb, const double* positions_dev, bool order_atoms, bool mixed, bool grad,
                bool smooth, const BatchOutput& out, const BatchDebug* debug, RunStats* stats, long long capacity,
                long long max_items, const AtmShape& shape, hipStream_t stream);
};

Engine::Engine() : ws_(new Workspace) {}
Engine::~Engine() { delete ws_; }

void Engine::set_kernel_timing(bool enable) {
  ws_->timer.clear();
  ws_->timer.enabled = enable;
}

std::pair<double, long long> Engine::kernel_time(const std::string& name) const { return ws_->timer.get(name); }

Prepared Workspace::upload(const std::vector<SystemDesc>& systems, const int* numbers, hipStream_t stream,
                           const int* ghost_host) {
  Prepared pb;
  const int nsys = static_cast<int>(systems.size());
  int natoms = 0;
  for (const auto& s : systems) {
    if (s.natoms <= 0) throw Error("Invalid molecular structure data provided");
    // the n x n pair records of a system (16 or 32 bytes each) outgrow any
    // device memory long before n^2 reaches 2^47 (11.8 million atoms); the
    // bound keeps every index product far from 64-bit overflow
    if (static_cast<long long>(s.natoms) * s.natoms >= (1LL << 47))
      throw Error("Molecular structure too large: " + std::to_string(s.natoms) + " atoms");
    if (natoms > std::numeric_limits<int>::max() - s.natoms)
      throw Error("Too many atoms in one batch");
    natoms += s.natoms;
  }
  const auto& tab = element_tables();
  std::map<int, int> zmap;
  std::vector<int> zlist;
  std::vector<int> sp(natoms), sysidx(natoms), sid(natoms);
  for (int k = 0, a = 0; k < nsys; ++k) {
    std::map<int, int> local;  // element -> species identifier within system k
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
      sid[a] = local.emplace(z, static_cast<int>(local.size()) + 1).first->second;
    }
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
  // A low-rank model replaces every reference C6 by its expansion
  // (model.f90:379-382), so the pair records, the real-space two-body term
  // and the three-body term use the low-rank C6 unchanged.
  const LowRank* lr = nullptr;
  for (const auto& s : systems)
    if (s.lowrank) lr = s.lowrank;
  if (lr) {
    if (nsys != 1) throw Error("A low-rank model evaluates one structure per call");
    std::vector<int> lsp(nsp, -1);
    for (int a = 0; a < nsp; ++a)
      for (size_t q = 0; q < lr->species.size(); ++q)
        if (lr->species[q] == zlist[a]) lsp[a] = static_cast<int>(q);
    for (int a = 0; a < nsp; ++a)
      if (lsp[a] < 0) throw Error("Low-rank model does not match the molecular structure");
    for (int a = 0; a < nsp; ++a)
      for (int c = 0; c < nsp; ++c)
        for (int ri = 0; ri < nref[a]; ++ri)
          for (int rj = 0; rj < nref[c]; ++rj)
            c6ref[(size_t(a) * nsp + c) * kMaxRef * kMaxRef + ri * kMaxRef + rj] =
                lr->reference_c6(lsp[a], ri, lsp[c], rj);
  }
  std::vector<SysGeom> geom(nsys);
  std::vector<SysParam> par(nsys);
  std::vector<double4> img;
  std::vector<int2> tiles;  // 64-position tiles of every system's spatial order
  long long pair_total = 0, slot_bound = 0, item_bound = 0, atm_atoms = 0;
  bool any_atm = false, smooth = false, periodic_atm = false;
  for (int k = 0, a = 0; k < nsys; ++k) {
    const SystemDesc& sd = systems[k];
    SysGeom& g = geom[k];
    g.atom_begin = a;
    g.natoms = sd.natoms;
    a += sd.natoms;
    par[k] = sd.param;
    if (sd.nparts < 1 || sd.part < 0 || sd.part >= sd.nparts) throw Error("Invalid dispersion work partition");
    par[k].part = sd.part;
    par[k].nparts = sd.nparts;
    pb.partitioned = pb.partitioned || sd.nparts > 1;
    if (lr && sd.periodic[0] && sd.periodic[1] && sd.periodic[2]) {  // disp.f90:104-111
      if (!supports_ewald(sd.param)) throw Error("Damping function does not support Ewald summation");
      par[k].ewald = 1;
      pb.ewald = true;
    }
    auto add = [&](double cut, int& begin, int& count) {
      const std::vector<double> tr = lattice_points(sd.periodic, sd.lattice, cut);
      begin = static_cast<int>(img.size());
      count = static_cast<int>(tr.size() / 3);
      for (int t = 0; t < count; ++t) img.push_back(make_double4(tr[3 * t], tr[3 * t + 1], tr[3 * t + 2], 0.0));
    };
    add(sd.param.cut_cn, g.img_cn_begin, g.img_cn_count);
    add(sd.param.cut2, g.img2_begin, g.img2_count);
    add(sd.param.cut3, g.img3_begin, g.img3_count);
    image_box(sd, sd.param.cut_cn, &img[g.img_cn_begin], g.img_cn_count, kImgCn, g);
    image_box(sd, sd.param.cut2, &img[g.img2_begin], g.img2_count, kImg2, g);
    image_box(sd, sd.param.cut3, &img[g.img3_begin], g.img3_count, kImg3, g);
    g.c6_begin = pair_total;
    pair_total += static_cast<long long>(sd.natoms) * sd.natoms;
    g.tile_begin = static_cast<int>(tiles.size());
    // tile culling for molecules of more than 8 tiles (in smaller ones every
    // tile is usually within the cutoffs and the box tests only cost)
    g.cull = !(sd.periodic[0] || sd.periodic[1] || sd.periodic[2]) && sd.natoms > 8 * kTile;
    for (int t = 0; t * kTile < sd.natoms; ++t)
      tiles.push_back(make_int2(g.atom_begin + t * kTile, std::min(kTile, sd.natoms - t * kTile)));
    if (std::fabs(sd.param.s9) >= kEps) {
      any_atm = true;
      smooth |= sd.param.width3 > 0.0;
      if (g.img3_count > 1) periodic_atm = true;
      atm_atoms += sd.natoms;
      // slot j <= i of atom i: at most (il + 1) images each
      for (long long il = 0; il < sd.natoms; ++il) {
        const long long m = (il + 1) * g.img3_count;
        slot_bound += m;
        item_bound += (m + kTile - 1) / kTile;
      }
    }
  }
  // Reciprocal-space set-up (damping.f90:306-385, ewald.f90:113-121).
  ReciprocalTiles rt;
  SpmeMesh sm;
  std::vector<double> lam, vecb, pval;
  std::vector<FourierTerm> terms;
  std::vector<int> nterm, sorder, spos, chunk_seg;
  std::vector<int4> chunks, groups;
  if (pb.ewald) {
    const SystemDesc& sd = systems[0];
    terms.assign(size_t(nsp) * nsp * 2, FourierTerm{0.0, 1.0, 0, 2});
    nterm.assign(size_t(nsp) * nsp, 0);
    double kcut = 0.0;
    for (int a = 0; a < nsp; ++a)
      for (int c = 0; c < nsp; ++c) {
        FourierTerm* t = &terms[2 * (size_t(a) * nsp + c)];
        const int nt = fourier_terms(par[0], zlist[a], zlist[c], t);
        nterm[size_t(a) * nsp + c] = nt;
        for (int q = 0; q < nt; ++q) {
          if (!fourier_supported(t[q])) throw Error("Damped pair potential has no known Fourier transform");
          kcut = std::max(kcut, fourier_cutoff(t[q], 1.0e-8));
        }
      }
    if (kcut > 0.0) {  // else every scaling factor vanishes: no two-body term
      if (lr->kcut > 0.0) kcut = lr->kcut;
      if (lr->mesh >= 0) {
        pb.spme = true;
        sm = spme_mesh(sd.lattice, kcut, lr->mesh);
      } else {
        rt = reciprocal_tiles(sd.lattice, kcut);
      }
      const int R = lr->rank;
      lam = lr->lambda;
      vecb.assign(size_t(nsp) * kMaxRef * R, 0.0);
      pval.assign(nsp, 0.0);
      for (int a = 0; a < nsp; ++a) {
        int q = 0;
        while (lr->species[q] != zlist[a]) ++q;
        for (int r = 0; r < kMaxRef; ++r)
          for (int l = 0; l < R; ++l) vecb[(size_t(a) * kMaxRef + r) * R + l] = lr->vec[(size_t(q) * kMaxRef + r) * R + l];
        for (int t = 0; t < nterm[size_t(a) * nsp + a]; ++t)
          pval[a] += fourier_potential_zero(terms[2 * (size_t(a) * nsp + a) + t]);
      }
      // atoms grouped by species; each species cut into chunks so that the
      // structure-factor pass has some 8192 wavefronts, and into groups of
      // kEwaldGroup for the per-atom pass
      const int ntiles = static_cast<int>(rt.tiles.size());
      const int want = std::max(nsp, (8192 + ntiles - 1) / std::max(ntiles, 1));
      const int chunk = std::max(16, (natoms + want - 1) / want);
      chunk_seg.assign(nsp + 1, 0);
      for (int a = 0; a < nsp; ++a) {
        const int begin = static_cast<int>(sorder.size());
        for (int i = 0; i < natoms; ++i)
          if (sp[i] == a) sorder.push_back(i);
        const int end = static_cast<int>(sorder.size());
        chunk_seg[a] = static_cast<int>(chunks.size());
        for (int q = begin; q < end; q += chunk) chunks.push_back(make_int4(a, q, std::min(end, q + chunk), 0));
        for (int q = begin; q < end; q += kEwaldGroup)
          groups.push_back(make_int4(a, q, std::min(end, q + kEwaldGroup), 0));
      }
      chunk_seg[nsp] = static_cast<int>(chunks.size());
      spos.assign(natoms, 0);
      for (int q = 0; q < natoms; ++q) spos[sorder[q]] = q;
    }
    lr->last_kcut = kcut;
    lr->last_vectors = pb.spme ? static_cast<long long>(sm.m[0]) * sm.m[1] * sm.m[2] : rt.vectors;
  }

  std::vector<char>& sb = staging;
  sb.clear();
  std::vector<int> ghost;
  if (ghost_host) ghost.assign(ghost_host, ghost_host + natoms);
  const size_t o_ghost = stage(sb, ghost), o_sp = stage(sb, sp), o_sys = stage(sb, sysidx), o_sid = stage(sb, sid),
               o_geom = stage(sb, geom),
               o_par = stage(sb, par), o_img = stage(sb, img), o_rcov = stage(sb, rcov), o_r4 = stage(sb, r4r2),
               o_nref = stage(sb, nref), o_refcn = stage(sb, refcn), o_c6r = stage(sb, c6ref),
               o_rvdw = stage(sb, rvdw), o_ptiles = stage(sb, tiles);
  // a single small system (the usual call of the drop-in interface) keeps
  // the atom order as its storage order: in one or a few tiles the spatial
  // order gains nothing and its sort would run with every evaluation
  pb.fixed_order = nsys == 1 && natoms <= 8 * kTile;
  std::vector<int> iota;
  if (pb.fixed_order) {
    iota.resize(natoms);
    for (int k = 0; k < natoms; ++k) iota[k] = k;
  }
  const size_t o_iota = stage(sb, iota);
  const size_t o_tiles = stage(sb, rt.tiles), o_lam = stage(sb, lam), o_vec = stage(sb, vecb),
               o_pval = stage(sb, pval), o_terms = stage(sb, terms), o_nterm = stage(sb, nterm),
               o_sorder = stage(sb, sorder), o_chunks = stage(sb, chunks), o_cseg = stage(sb, chunk_seg),
               o_groups = stage(sb, groups), o_spos = stage(sb, spos), o_bfac = stage(sb, sm.bfac);
  char* dmeta = static_cast<char*>(meta.get(sb.size()));
  if (arena) {
    // from pageable memory: the copy has read the host data when it returns
    hip_check(hipMemcpyAsync(dmeta, sb.data(), sb.size(), hipMemcpyHostToDevice, stream), "upload");
  } else {
    char* host = pinned.get(sb.size());
    std::memcpy(host, sb.data(), sb.size());
    hip_check(hipMemcpyAsync(dmeta, host, sb.size(), hipMemcpyHostToDevice, stream), "upload");
    pinned.record(stream);
  }

  DevBatch& db = pb.db;
  db.natoms = natoms;
  db.nsys = nsys;
  db.nsp = nsp;
  db.mref = mref;
  db.pos = pos.as<double4>(natoms);
  db.sp = reinterpret_cast<const int*>(dmeta + o_sp);
  db.sys = reinterpret_cast<const int*>(dmeta + o_sys);
  db.sid = reinterpret_cast<const int*>(dmeta + o_sid);
  db.geom = reinterpret_cast<const SysGeom*>(dmeta + o_geom);
  db.par = reinterpret_cast<const SysParam*>(dmeta + o_par);
  db.img = reinterpret_cast<const double4*>(dmeta + o_img);
  db.rcov = reinterpret_cast<const double*>(dmeta + o_rcov);
  db.r4r2 = reinterpret_cast<const double*>(dmeta + o_r4);
  db.nref = reinterpret_cast<const int*>(dmeta + o_nref);
  db.refcn = reinterpret_cast<const double*>(dmeta + o_refcn);
  db.c6ref = reinterpret_cast<const double*>(dmeta + o_c6r);
  db.rvdw = reinterpret_cast<const double*>(dmeta + o_rvdw);
  db.ghost = ghost_host ? reinterpret_cast<const int*>(dmeta + o_ghost) : nullptr;
  db.rank = pb.fixed_order ? reinterpret_cast<const int*>(dmeta + o_iota) : rank.as<int>(natoms);
  db.order = pb.fixed_order ? reinterpret_cast<const int*>(dmeta + o_iota) : order.as<int>(natoms);
  db.posr = pos_rank.as<double4>(natoms);
  db.ntiles = static_cast<int>(tiles.size());
  db.tiles = reinterpret_cast<const int2*>(dmeta + o_ptiles);
  db.tbox = tbox.as<double4>(2 * tiles.size());
  pb.pair_total = pair_total;
  pb.any_atm = any_atm;
  pb.smooth = smooth;
  pb.periodic_atm = periodic_atm;
  pb.slot_bound = slot_bound;
  pb.atm_atoms = atm_atoms;
  pb.item_bound = item_bound;
  if (pb.ewald) {
    EwaldBatch& e = pb.eb;
    e.ntiles = static_cast<int>(rt.tiles.size());
    e.rank = lr->rank;
    e.rank_pad = (lr->rank + 7) / 8 * 8;
    e.nsp = nsp;
    e.natoms = natoms;
    e.atom_begin = 0;
    e.nchunks = static_cast<int>(chunks.size());
    e.ngroups = static_cast<int>(groups.size());
    e.inv_vol = rt.volume > 0.0 ? 1.0 / rt.volume : 0.0;
    std::memcpy(e.rec, rt.rec, sizeof e.rec);
    e.rep0 = rt.rep[0];
    e.rep1 = rt.rep[1];
    e.rep2 = rt.rep[2];
    e.toff0 = 0;
    e.toff1 = 2 * rt.rep[0] + 1;
    e.toff2 = e.toff1 + 2 * rt.rep[1] + 1;
    e.tstride = e.toff2 + 2 * rt.rep[2] + 1;
    e.tiles = reinterpret_cast<const int4*>(dmeta + o_tiles);
    e.lambda = reinterpret_cast<const double*>(dmeta + o_lam);
    e.vec = reinterpret_cast<const double*>(dmeta + o_vec);
    e.pval = reinterpret_cast<const double*>(dmeta + o_pval);
    e.terms = reinterpret_cast<const FourierTerm*>(dmeta + o_terms);
    e.nterm = reinterpret_cast<const int*>(dmeta + o_nterm);
    e.sorder = reinterpret_cast<const int*>(dmeta + o_sorder);
    e.chunks = reinterpret_cast<const int4*>(dmeta + o_chunks);
    e.chunk_seg = reinterpret_cast<const int*>(dmeta + o_cseg);
    e.groups = reinterpret_cast<const int4*>(dmeta + o_groups);
    e.spos = reinterpret_cast<const int*>(dmeta + o_spos);
    e.part = systems[0].part;
    e.nparts = systems[0].nparts;
    e.mesh = pb.spme ? 1 : 0;
    if (pb.spme) {
      SpmeBatch& q = pb.sb;
      q.m0 = sm.m[0];
      q.m1 = sm.m[1];
      q.m2 = sm.m[2];
      q.mh = sm.m[0] / 2 + 1;
      q.mmax = sm.mmax;
      q.ng = static_cast<long long>(q.m0) * q.m1 * q.m2;
      q.nch = static_cast<long long>(q.mh) * q.m1 * q.m2;
      std::memcpy(q.rec, sm.rec, sizeof q.rec);
      const int md[3] = {q.m0, q.m1, q.m2};
      for (int k = 0; k < 9; ++k) {
        q.inv[k] = sm.rec[k] / (2.0 * M_PI);
        q.dudr[k] = md[k / 3] * sm.rec[k] / (2.0 * M_PI);
      }
      q.inv_vol = 1.0 / sm.volume;
      q.bfac = reinterpret_cast<const double*>(dmeta + o_bfac);
      pb.spme_nsp = nsp;
      e.ntiles = 0;
      auto put = [&](const void* d, size_t n) {
        const char* c = static_cast<const char*>(d);
        pb.spme_key.insert(pb.spme_key.end(), c, c + n);
      };
      put(sm.m, sizeof sm.m);
      put(sm.rec, sizeof sm.rec);
      put(&nsp, sizeof nsp);
      put(terms.data(), terms.size() * sizeof(FourierTerm));
      put(nterm.data(), nterm.size() * sizeof(int));
    }
  }
  return pb;
}

DevWork Workspace::work(const Prepared& pb, bool mixed, bool grad, const BatchOutput& out,
                        const BatchDebug* debug) {
  const DevBatch& db = pb.db;
  const int natoms = db.natoms;
  DevWork w{};
  w.cn = debug && debug->cn ? debug->cn : cn.as<double>(natoms);
  w.gw = gw.as<double>(size_t(natoms) * db.mref);
  w.dgw = grad ? dgw.as<double>(size_t(natoms) * db.mref) : nullptr;
  w.pair = pair.get(std::max<long long>(pb.pair_total, 1) * pair_record_bytes(mixed));
  w.e_atom = e_atom.as<double>(natoms);
  w.dedcn = debug && debug->dedcn ? debug->dedcn : dedcn.as<double>(natoms);
  w.grad = out.gradient ? out.gradient : this->grad.as<double>(size_t(natoms) * 3);
  w.sig_atom = sig.as<double>(size_t(natoms) * 6);
  w.atm_acc = grad && pb.any_atm ? atm_acc.as<double4>(natoms) : nullptr;
  w.gwr = gw_rank.as<double>(size_t(natoms) * 2 * db.mref);
  w.spr = sp_rank.as<int>(natoms);
  return w;
}

void Workspace::list_offsets(const Prepared& pb, const void* pair_records, bool mixed, double4* acc,
                             hipStream_t stream) {
  const int natoms = pb.db.natoms;
  long long* c = cnt.as<long long>(natoms);
  long long* o = off.as<long long>(natoms + 1);
  long long* io = item_off.as<long long>(natoms + 1);
  timed("atm_list", stream, [&] {
    launch_atm_count(pb.db, pair_records, mixed, c, acc, stream);
    launch_list_scan(c, natoms, o, io, stream);
  });
}

void Workspace::evaluate(const Prepared& pb, const double* positions_dev, bool order_atoms, bool mixed, bool grad,
                         bool smooth, const BatchOutput& out, const BatchDebug* debug, RunStats* stats,
                         long long capacity, long long max_items, const AtmShape& shape, hipStream_t stream) {
  const DevBatch& db = pb.db;
  const int natoms = db.natoms;
  pack(pb, positions_dev, stream, order_atoms);
  const DevWork w = work(pb, mixed, grad, out, debug);
  timed("cn", stream, [&] { launch_cn(db, mixed, w.cn, stream); });
  timed("weights", stream, [&] { launch_weights(db, w.cn, w.gw, w.dgw, stream); });
  timed("c6", stream, [&] { launch_pair(db, mixed, w.gw, w.dgw, w.gwr, w.spr, w.pair, stream); });
  timed("disp2", stream, [&] { launch_disp2(db, mixed, grad, w, nullptr, stream); });
  if (pb.ewald && pb.spme) {
    // fourier/spme.f90:78-328, term by term of the expansion
    const EwaldBatch& e = pb.eb;
    const SpmeBatch& q = pb.sb;
    const size_t R = static_cast<size_t>(e.rank), npair = size_t(e.nsp) * (e.nsp + 1) / 2;
    EwaldWork ew{};
    ew.c6l = ew_c6l.as<double>(size_t(natoms) * e.rank_pad);
    ew.lc6l = ew_lc6l.as<double>((size_t(natoms) + kEwaldGroup) * R);
    ew.ldc6l = grad ? ew_ldc6l.as<double>((size_t(natoms) + kEwaldGroup) * R) : nullptr;
    SpmeWork sw{};
    sw.base = sp_base.as<int>(size_t(natoms) * 3);
    sw.theta = sp_theta.as<double>(size_t(natoms) * 3 * kSplineOrder);
    sw.dtheta = sp_dtheta.as<double>(size_t(natoms) * 3 * kSplineOrder);
    sw.Q = sp_q.as<double>(size_t(e.nsp) * q.ng);
    sw.T = sw.Q;  // the spread mesh is dead once transformed
    sw.Qh = sp_qh.as<double2>(size_t(e.nsp) * q.nch);
    sw.Th = sp_th.as<double2>(size_t(e.nsp) * q.nch);
    sw.K = sp_k.as<double>(npair * q.nch);
    sw.Dm = grad ? sp_dm.as<double>(npair * q.nch) : nullptr;
    sw.Dmir = grad ? sp_dmir.as<double>(npair * q.nch) : nullptr;
    fft_plans(q, e.nsp, stream);
    // the kernel table is kept while the cell, parameters and mesh stay
    const bool table = pb.spme_key != spme_key || (grad && !spme_key_grad);
    timed("lowrank", stream, [&] { launch_lowrank(db, e, grad, w, ew, stream); });
    timed("spme_setup", stream, [&] { launch_spme_setup(db, e, q, grad, table, sw, stream); });
    if (table) {
      spme_key = pb.spme_key;
      spme_key_grad = grad;
    }
    for (int l = 0; l < e.rank; ++l) {
      if (e.nparts > 1 && l % e.nparts != e.part) continue;  // spme.f90:198
      timed("spme_spread", stream, [&] {
        hip_check(hipMemsetAsync(sw.Q, 0, sizeof(double) * size_t(e.nsp) * q.ng, stream), "hipMemsetAsync");
        launch_spme_spread(db, e, q, ew, sw, l, stream);
      });
      timed("spme_fft", stream, [&] { fft_check(hipfftExecD2Z(fft_fwd, sw.Q, reinterpret_cast<hipfftDoubleComplex*>(sw.Qh)), "hipfftExecD2Z"); });
      timed("spme_conv", stream, [&] { launch_spme_conv(e, q, grad, w, sw, l, stream); });
      timed("spme_fft", stream, [&] { fft_check(hipfftExecZ2D(fft_inv, reinterpret_cast<hipfftDoubleComplex*>(sw.Th), sw.T), "hipfftExecZ2D"); });
      timed("spme_gather", stream, [&] { launch_spme_gather(db, e, q, grad, w, ew, sw, l, stream); });
    }
  }
  if (pb.ewald && !pb.spme && pb.eb.ntiles > 0) {
    const EwaldBatch& e = pb.eb;
    EwaldWork ew{};
    const size_t nk = static_cast<size_t>(e.ntiles) * kWave, R = static_cast<size_t>(e.rank);
    ew.etab = ew_tab.as<double2>(size_t(natoms) * e.tstride);
    ew.c6l = ew_c6l.as<double>(size_t(natoms) * e.rank_pad);
    ew.lc6l = ew_lc6l.as<double>((size_t(natoms) + kEwaldGroup) * R);
    ew.ldc6l = grad ? ew_ldc6l.as<double>((size_t(natoms) + kEwaldGroup) * R) : nullptr;
    ew.part = ew_part.as<double2>(size_t(e.nchunks) * R * nk);
    ew.phi = ew_phi.as<double2>(size_t(e.nsp) * (e.nsp + 1) / 2 * nk);
    ew.F = ew_f.as<double2>(size_t(e.nsp) * R * nk);
    ew.S = e.nsp > kMaxEwaldSpecies ? ew_s.as<double2>(size_t(e.nsp) * R * nk) : nullptr;
    timed("lowrank", stream, [&] { launch_lowrank(db, e, grad, w, ew, stream); });
    timed("ewald_sf", stream, [&] { launch_ewald_sf(e, ew, stream); });
    timed("ewald_f", stream, [&] { launch_ewald_f(e, grad, w, ew, stream); });
    timed("ewald_atom", stream, [&] { launch_ewald_atom(e, grad, w, ew, stream); });
  }
  if (pb.any_atm) {
    list_offsets(pb, w.pair, mixed, w.atm_acc, stream);
    long long* o = off.as<long long>(natoms + 1);
    long long* io = item_off.as<long long>(natoms + 1);
    void* s = nullptr;
    const long long bound_bytes = pb.slot_bound * static_cast<long long>(slot_record_bytes(mixed));
    if (capacity < 0 && !stats && !arena && !pb.periodic_atm && bound_bytes <= kUnsyncedSlotBytes) {
      // molecules whose exact list bound (j <= i: il + 1 slots per atom)
      // fits a modest buffer: no read-back of the list size, so the call
      // does not wait for the device here
      capacity = std::max<long long>(pb.slot_bound, 1);
      max_items = pb.item_bound;
      s = slots.get(capacity * slot_record_bytes(mixed));
    } else if (capacity < 0) {
      long long total = 0;
      hip_check(hipMemcpyAsync(&total, o + natoms, sizeof total, hipMemcpyDeviceToHost, stream), "list size");
      hip_check(hipStreamSynchronize(stream), "list size sync");
      if (stats) stats->atm_list_entries = total;
      capacity = total;
      max_items = total / kTile + natoms;  // bound of the number of work items
      s = slots.get(std::max<long long>(total, 1) * slot_record_bytes(mixed));
    } else {
      s = slots.get(std::max<long long>(capacity, 1) * slot_record_bytes(mixed));
      launch_atm_guard(db, o, capacity, io, flag.as<int>(1), w, grad, stream);
    }
    timed("atm_list", stream, [&] { launch_atm_fill(db, w.pair, mixed, o, s, capacity, stream); });
    const AtmShape sh = shape.waves == 0 ? per_call_shape(mixed, pb.periodic_atm, max_items) : shape;
    unsigned long long* wc = sh.blocks_per_cu > 0 ? counter.as<unsigned long long>(1) : nullptr;
    timed("atm", stream,
          [&] { launch_atm(db, mixed, grad, smooth, pb.partitioned, sh, io, max_items, o, s, w, wc, stream); });
  }
  if (grad) timed("cn_chain", stream, [&] { launch_cn_chain(db, mixed, w, stream); });
  timed("reduce", stream, [&] { launch_reduce(db, w, out.energy, grad ? out.sigma : nullptr, stream); });
  hip_check(hipGetLastError(), "kernel launch");
}

void Engine::run(const std::vector<SystemDesc>& systems, const int* numbers, const double* positions_dev,
                 int precision, const BatchOutput& out, hipStream_t stream, const BatchDebug* debug,
                 RunStats* stats, const int* ghost_host) {
  if (systems.empty()) return;
  Workspace& m = *ws_;
  const Prepared pb = m.upload(systems, numbers, stream, ghost_host);
  const bool grad = out.gradient || out.sigma;
  m.evaluate(pb, positions_dev, true, precision == kMixed, grad, pb.smooth, out, debug, stats, -1, 0, kPerCallShape,
             stream);
  if (m.timer.enabled) m.timer.collect();
}

void run_noalloc(const std::vector<SystemDesc>& systems, const int* numbers, const double* positions_dev,
                 int precision, const BatchOutput& out, hipStream_t stream) {
  if (systems.empty()) return;
  for (const auto& sd : systems)
    if (sd.lowrank) throw Error("The allocation-free entry point does not take a low-rank model");
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hipError_t err = hipSuccess;
  void* area = noalloc_area(dev, &err);
  hip_check(err, "hipGetSymbolAddress");
  struct State {
    Arena arena;
    Workspace ws;
  };
  static std::mutex mu;
  static std::map<int, State*> states;  // one per device, kept to the end of the process
  std::lock_guard<std::mutex> lock(mu);
  State*& st = states[dev];
  if (!st) {
    st = new State;
    st->arena.bind(area, kNoAllocBytes);
    st->ws.use_arena(&st->arena);
  }
  Workspace& m = st->ws;
  m.reset_arena();
  const Prepared pb = m.upload(systems, numbers, stream, nullptr);
  const bool grad = out.gradient || out.sigma;
  // the gradient is summed in the work area and copied out at the end, so
  // that a batch rejected for its size leaves every output untouched
  BatchOutput inner = out;
  inner.gradient = nullptr;
  m.evaluate(pb, positions_dev, true, precision == kMixed, grad, pb.smooth, inner, nullptr, nullptr, -1, 0,
             kPerCallShape, stream);
  if (out.gradient)
    hip_check(hipMemcpyAsync(out.gradient, m.grad.as<double>(size_t(pb.db.natoms) * 3),
                             sizeof(double) * 3 * size_t(pb.db.natoms), hipMemcpyDeviceToDevice, stream),
              "gradient copy");
}

void Engine::run_pairwise(const std::vector<SystemDesc>& systems, const int* numbers, const double* positions_dev,
                          int precision, double* pair2, double* pair3, hipStream_t stream, const int* ghost_host) {
  // disp.f90:403-457: CN, weights and C6 as for the energy (no derivatives),
  // then the pairwise two-body and three-body energies
  if (systems.empty()) return;
  for (const auto& sd : systems)  // disp.f90:429-433
    if (sd.lowrank && sd.periodic[0] && sd.periodic[1] && sd.periodic[2])
      throw Error("Pairwise analysis is only available for the real space summation");
  Workspace& m = *ws_;
  const Prepared pb = m.upload(systems, numbers, stream, ghost_host);
  const DevBatch& db = pb.db;
  const int natoms = db.natoms;
  const bool mixed = precision == kMixed;
  m.pack(pb, positions_dev, stream, true);
  const DevWork w = m.work(pb, mixed, false, BatchOutput{}, nullptr);
  m.timed("cn", stream, [&] { launch_cn(db, mixed, w.cn, stream); });
  m.timed("weights", stream, [&] { launch_weights(db, w.cn, w.gw, nullptr, stream); });
  m.timed("c6", stream, [&] { launch_pair(db, mixed, w.gw, nullptr, w.gwr, w.spr, w.pair, stream); });
  m.timed("disp2", stream, [&] { launch_disp2(db, mixed, false, w, pair2, stream); });
  const size_t bytes = sizeof(double) * static_cast<size_t>(pb.pair_total);
  if (!pb.any_atm) {
    hip_check(hipMemsetAsync(pair3, 0, bytes, stream), "hipMemsetAsync");
  } else {
    double* S = m.pair3s.as<double>(static_cast<size_t>(pb.pair_total));
    hip_check(hipMemsetAsync(S, 0, bytes, stream), "hipMemsetAsync");
    m.list_offsets(pb, w.pair, mixed, nullptr, stream);
    long long* o = m.off.as<long long>(natoms + 1);
    long long* io = m.item_off.as<long long>(natoms + 1);
    long long total = 0;
    hip_check(hipMemcpyAsync(&total, o + natoms, sizeof total, hipMemcpyDeviceToHost, stream), "list size");
    hip_check(hipStreamSynchronize(stream), "list size sync");
    void* sl = m.slots.get(std::max<long long>(total, 1) * slot_record_bytes(mixed));
    m.timed("atm_list", stream, [&] { launch_atm_fill(db, w.pair, mixed, o, sl, total, stream); });
    m.timed("atm", stream, [&] {
      launch_atm_pairs(db, mixed, pb.smooth, io, total / kTile + natoms, o, sl, w, S, stream);
      launch_pair3_finalize(db, S, pair3, stream);
    });
  }
  hip_check(hipGetLastError(), "kernel launch");
  if (m.timer.enabled) m.timer.collect();
}

void Engine::run_hessian(const std::vector<SystemDesc>& systems, const int* numbers, const double* positions_dev,
                         double* energy, double* hessian, hipStream_t stream, const int* ghost_host) {
  // disp.f90:236-308: the FP64 energy evaluation leaves the coordination
  // numbers, the pair records and the three-body lists in the workspace; the
  // Hessian kernels reuse them.
  if (systems.empty()) return;
  for (const auto& sd : systems)  // disp.f90:113-119
    if (sd.lowrank) throw Error("Hessian is only available for the real space summation");
  Workspace& m = *ws_;
  const Prepared pb = m.upload(systems, numbers, stream, ghost_host);
  const DevBatch& db = pb.db;
  const int natoms = db.natoms;
  double* e = energy ? energy : m.e_out.as<double>(static_cast<size_t>(db.nsys));
  BatchOutput out;
  out.energy = e;
  RunStats stats;
  m.evaluate(pb, positions_dev, true, false, false, pb.smooth, out, nullptr, &stats, -1, 0, kPerCallShape, stream);
  const size_t nn = static_cast<size_t>(pb.pair_total), nm = static_cast<size_t>(natoms) * db.mref;
  const size_t words = 3 * nm + 4 * nn + 3 * nn + natoms + 3 * nn + nn + 3 * nn;
  double* base = m.hbuf.as<double>(words);
  HessWork h{};
  h.cn = m.cn.as<double>(natoms);
  h.gw = base;
  h.gwd = h.gw + nm;
  h.gwd2 = h.gwd + nm;
  h.c6 = h.gwd2 + nm;
  h.dc6 = h.c6 + nn;
  h.d2c6 = h.dc6 + nn;
  h.d2c6ij = h.d2c6 + nn;
  h.dr = h.d2c6ij + nn;
  h.dedcn = h.dr + 3 * nn;
  h.dEdcndr = h.dedcn + natoms;
  h.dEdcndcn = h.dEdcndr + 3 * nn;
  h.tmp = h.dEdcndcn + nn;
  h.hess = hessian;
  hip_check(hipMemsetAsync(hessian, 0, sizeof(double) * 9 * nn, stream), "hipMemsetAsync");
  hip_check(hipMemsetAsync(h.dedcn, 0, sizeof(double) * (natoms + 4 * nn), stream), "hipMemsetAsync");
  int max_natoms = 0;
  for (const auto& sd : systems) max_natoms = std::max(max_natoms, sd.natoms);
  const long long total = pb.any_atm ? stats.atm_list_entries : 0;
  m.timed("hessian", stream, [&] {
    launch_hessian(db, h, m.off.as<long long>(natoms + 1), total > 0 ? m.slots.get(0) : nullptr, total, max_natoms,
                   stream);
  });
  hip_check(hipGetLastError(), "kernel launch");
  if (m.timer.enabled) m.timer.collect();
}

void Engine::count_terms(const std::vector<SystemDesc>& systems, const int* numbers, const double* positions_dev,
                         long long* counts_host, hipStream_t stream) {
  if (systems.empty()) return;
  Workspace& m = *ws_;
  std::vector<SystemDesc> sys = systems;
  for (auto& s : sys)
    if (std::fabs(s.param.s9) < kEps) s.param.s9 = 1.0;  // triples are counted whatever s9 is
  const Prepared pb = m.upload(sys, numbers, stream, nullptr);
  m.pack(pb, positions_dev, stream, true);
  const int natoms = pb.db.natoms;
  unsigned long long* dc = m.counts.as<unsigned long long>(size_t(pb.db.nsys) * 3);
  hip_check(hipMemsetAsync(dc, 0, sizeof(unsigned long long) * 3 * pb.db.nsys, stream), "hipMemsetAsync");
  launch_count_pairs(pb.db, dc, stream);
  // lists without the C6 test: every pair within the cutoff is a slot
  m.list_offsets(pb, nullptr, false, nullptr, stream);
  long long* o = m.off.as<long long>(natoms + 1);
  long long total = 0;
  hip_check(hipMemcpyAsync(&total, o + natoms, sizeof total, hipMemcpyDeviceToHost, stream), "list size");
  hip_check(hipStreamSynchronize(stream), "list size sync");
  void* s = m.slots.get(std::max<long long>(total, 1) * slot_record_bytes(false));
  launch_atm_fill(pb.db, nullptr, false, o, s, total, stream);
  launch_count_triples(pb.db, false, o, s, dc, stream);
  hip_check(hipGetLastError(), "kernel launch");
  std::vector<unsigned long long> h(3 * pb.db.nsys);
  hip_check(hipMemcpyAsync(h.data(), dc, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, stream),
            "download counts");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  for (size_t k = 0; k < h.size(); ++k) counts_host[k] = static_cast<long long>(h[k]);
}

// ------------------------------------------------------------------ Plan
struct Plan::State {
  Workspace ws;
  Prepared pb;
  bool mixed = false, grad = false;
  long long capacity = 0, max_items = 0;
  AtmShape shape{};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~State() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  void drop_graph() {
    if (exec) hip_check(hipGraphExecDestroy(exec), "hipGraphExecDestroy");
    if (graph) hip_check(hipGraphDestroy(graph), "hipGraphDestroy");
    exec = nullptr;
    graph = nullptr;
  }
};


Plan::Plan(const std::vector<SystemDesc>& systems, const int* numbers_host, const double* positions_dev,
           int precision, bool grad, const int* ghost_host, hipStream_t stream)
    : st_(new State) {
  try {
    if (systems.empty()) throw Error("Invalid molecular structure data provided");
    for (const auto& sd : systems)
      if (sd.lowrank) throw Error("A planned batch does not take a low-rank model");
    State& s = *st_;
    s.mixed = precision == kMixed;
    s.grad = grad;
    s.pb = s.ws.upload(systems, numbers_host, stream, ghost_host);
    s.shape.num_cus = device_cu_count();
    const int natoms = s.pb.db.natoms;
    hip_check(hipMemsetAsync(s.ws.flag.as<int>(1), 0, sizeof(int), stream), "hipMemsetAsync");
    // the spatial order of the planning coordinates is kept for every evaluation
    s.ws.pack(s.pb, positions_dev, stream, true);
    if (s.pb.any_atm) {
      long long cap = s.pb.slot_bound;
      if (s.pb.periodic_atm) {
        // lists of the initial coordinates, with headroom for later ones
        const DevWork w = s.ws.work(s.pb, s.mixed, grad, BatchOutput{}, nullptr);
        launch_cn(s.pb.db, s.mixed, w.cn, stream);
        launch_weights(s.pb.db, w.cn, w.gw, w.dgw, stream);
        launch_pair(s.pb.db, s.mixed, w.gw, w.dgw, w.gwr, w.spr, w.pair, stream);
        s.ws.list_offsets(s.pb, w.pair, s.mixed, w.atm_acc, stream);
        long long total = 0;
        hip_check(hipMemcpyAsync(&total, s.ws.off.as<long long>(natoms + 1) + natoms, sizeof total,
                                 hipMemcpyDeviceToHost, stream),
                  "list size");
        hip_check(hipStreamSynchronize(stream), "list size sync");
        cap = std::min(cap, std::max(total + total / 2, total + 64LL * s.pb.atm_atoms));
      }
      s.capacity = cap;
      s.max_items = std::min(s.pb.item_bound, cap / kTile + natoms);
      const AtmShape a = per_call_shape(s.mixed, s.pb.periodic_atm, s.max_items);  // until set_shape or autotune
      s.shape.waves = a.waves;
      s.shape.blocks_per_cu = a.blocks_per_cu;
      // size every buffer now, so that an evaluation allocates nothing
      (void)s.ws.slots.get(std::max<long long>(cap, 1) * slot_record_bytes(s.mixed));
      (void)s.ws.counter.as<unsigned long long>(1);
    }
    (void)s.ws.work(s.pb, s.mixed, grad, BatchOutput{}, nullptr);
    hip_check(hipStreamSynchronize(stream), "plan sync");
  } catch (...) {
    delete st_;
    throw;
  }
}

Plan::~Plan() { delete st_; }

int Plan::natoms() const { return st_->pb.db.natoms; }
int Plan::nsys() const { return st_->pb.db.nsys; }
long long Plan::slot_capacity() const { return st_->capacity; }
const AtmShape& Plan::shape() const { return st_->shape; }

void Plan::set_shape(int waves, int blocks_per_cu) {
  if (waves != 1 && waves != 2 && waves != 4) throw Error("three-body launch: wavefronts per block must be 1, 2 or 4");
  if (blocks_per_cu < 0 || blocks_per_cu > 64) throw Error("three-body launch: blocks per CU out of range");
  st_->shape.waves = waves;
  st_->shape.blocks_per_cu = blocks_per_cu;
  st_->drop_graph();
}

void Plan::execute(const double* positions_dev, const BatchOutput& out, hipStream_t stream) {
  State& s = *st_;
  if (!out.energy) throw Error("Energy output is missing");
  if (s.grad != (out.gradient != nullptr || out.sigma != nullptr))
    throw Error("The plan was made for " + std::string(s.grad ? "energy and gradient" : "the energy only"));
  s.ws.evaluate(s.pb, positions_dev, false, s.mixed, s.grad, s.pb.smooth, out, nullptr, nullptr, s.capacity,
                s.max_items, s.shape, stream);
}

std::vector<AtmShape> Plan::candidate_shapes() {
  std::vector<AtmShape> c;
  for (int waves : {4, 2})
    for (int bpc : {0, 2, 4, 8}) {
      AtmShape a;
      a.waves = waves;
      a.blocks_per_cu = bpc;
      c.push_back(a);
    }
  // one wavefront per block: for short lists (small molecules), whose work
  // items have one or two tiles to pair with
  for (int bpc : {0, 16}) {
    AtmShape a;
    a.waves = 1;
    a.blocks_per_cu = bpc;
    c.push_back(a);
  }
  return c;
}

int Plan::autotune(const double* positions_dev, const BatchOutput& out, hipStream_t stream, int reps,
                   std::vector<float>* times_ms) {
  State& s = *st_;
  const std::vector<AtmShape> cand = candidate_shapes();
  std::vector<float> t(cand.size(), 0.0f);
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "hipEventCreate");
  hip_check(hipEventCreate(&e1), "hipEventCreate");
  int best = 0;
  try {
    for (size_t k = 0; k < cand.size(); ++k) {
      s.shape.waves = cand[k].waves;
      s.shape.blocks_per_cu = cand[k].blocks_per_cu;
      execute(positions_dev, out, stream);  // warm-up
      hip_check(hipEventRecord(e0, stream), "hipEventRecord");
      for (int r = 0; r < std::max(reps, 1); ++r) execute(positions_dev, out, stream);
      hip_check(hipEventRecord(e1, stream), "hipEventRecord");
      hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
      hip_check(hipEventElapsedTime(&t[k], e0, e1), "hipEventElapsedTime");
      t[k] /= float(std::max(reps, 1));
      if (t[k] < t[best]) best = static_cast<int>(k);
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  s.shape.waves = cand[best].waves;
  s.shape.blocks_per_cu = cand[best].blocks_per_cu;
  s.drop_graph();
  if (times_ms) *times_ms = t;
  return best;
}

void Plan::capture(const double* positions_dev, const BatchOutput& out, hipStream_t stream) {
  State& s = *st_;
  s.drop_graph();
  const bool timing = s.ws.timer.enabled;
  s.ws.timer.enabled = false;
  hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  try {
    execute(positions_dev, out, stream);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(stream, &g);
    if (g) (void)hipGraphDestroy(g);
    s.ws.timer.enabled = timing;
    throw;
  }
  hip_check(hipStreamEndCapture(stream, &s.graph), "hipStreamEndCapture");
  s.ws.timer.enabled = timing;
  hip_check(hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

bool Plan::captured() const { return st_->exec != nullptr; }

void Plan::replay(hipStream_t stream) {
  if (!st_->exec) throw Error("The plan has no captured graph");
  hip_check(hipGraphLaunch(st_->exec, stream), "hipGraphLaunch");
}

bool Plan::overflowed(hipStream_t stream) {
  int h = 0;
  hip_check(hipMemcpyAsync(&h, st_->ws.flag.as<int>(1), sizeof h, hipMemcpyDeviceToHost, stream), "overflow flag");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  return h != 0;
}

}  // namespace d3
