// C ABI of libd3hip.so (declared in include/d3hip.h).
//
// Part 1 mirrors src/dftd3/api.f90 of simple-dftd3 v1.6.0 for the energy /
// gradient / virial path: opaque handles, the error handle, NULL means
// "absent", and the same error messages.  The computation itself runs on the
// GPU through d3::Engine.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "d3_engine.hpp"

#include "../../include/d3hip.h"

namespace {
#include "d3_params_data.inc"
}  // namespace

struct _dftd3_error {
  bool set = false;
  std::string message;
};

struct _dftd3_structure {
  int nat = 0;
  std::vector<int> num;
  std::vector<double> xyz;  // [nat][3]
  double lattice[9] = {0};
  int periodic[3] = {0, 0, 0};
};

struct _dftd3_model {
  bool has_structure = false;
  int nat = 0;
  std::vector<int> num;
  double cutoff[5] = {40.0, 60.0, 40.0, 0.0, 0.0};  // cn, disp2, disp3, width2, width3 (cutoff.f90:26-32)
  std::vector<int> ghost;
  std::unique_ptr<d3::LowRank> lowrank;  // set_model_ewald (model.f90:185-196)
  int part = 0, nparts = 1;               // set_model_work_partition (partition.f90:57-84)
  int precision = d3::kFP64;
  std::unique_ptr<d3::Engine> engine;
  hipStream_t stream = nullptr;
  // device buffers of the host-pointer entry point
  double* d_xyz = nullptr;
  double* d_out = nullptr;  // energy | sigma[9] | gradient[nat][3]
  size_t cap_atoms = 0;
  double* d_pairs = nullptr;  // pairwise energies pair2 | pair3, or the Hessian
  size_t cap_pairs = 0;       // doubles allocated at d_pairs
  ~_dftd3_model() {
    if (d_xyz) (void)hipFree(d_xyz);
    if (d_out) (void)hipFree(d_out);
    if (d_pairs) (void)hipFree(d_pairs);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct _dftd3_param {
  int kind = d3::kRational;
  // param.f90:38-48 defaults; a1/a2 hold rs6/rs8 for the (modified) zero
  // damping, a3/a4 are the CSO parameters (rs6/rs8 of d3_param)
  double s6 = 1.0, s8 = 1.0, s9 = 0.0, a1 = 0.4, a2 = 5.0, alp = 14.0, bet = 0.0, a3 = 1.0, a4 = 1.0;
};

namespace {

void fatal(dftd3_error err, const std::string& msg) {
  if (!err) return;
  err->set = true;
  err->message = msg;
}

template <typename F>
void guarded(dftd3_error err, F&& f) {
  try {
    f();
  } catch (const std::exception& e) {
    fatal(err, e.what());
  }
}

// mctc-lib matinv_3x3 on the Fortran lattice(3,3) (columns = vectors).
void wrap_to_central_cell(int nat, double* xyz, const double* lat, const int* periodic) {
  // utils.f90:29-64
  if (!(periodic[0] || periodic[1] || periodic[2])) return;
  double m[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[r][c] = lat[3 * c + r];
  const double det = m[0][0] * m[1][1] * m[2][2] - m[0][0] * m[1][2] * m[2][1] - m[0][1] * m[1][0] * m[2][2] +
                     m[0][1] * m[1][2] * m[2][0] + m[0][2] * m[1][0] * m[2][1] - m[0][2] * m[1][1] * m[2][0];
  const double di = 1.0 / det;
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
  auto shift_back = [](double in) {
    const double eps = 1.0e-14;
    double out = in;
    if (in < (0.0 - eps)) out = in + std::ceil(-in);
    if (in > (1.0 + eps)) out = in - std::floor(in);
    if (std::fabs(in - 1.0) < eps) out = in - 1.0;
    return out;
  };
  for (int i = 0; i < nat; ++i) {
    double* x = xyz + 3 * i;
    double f[3];
    for (int r = 0; r < 3; ++r) f[r] = shift_back(inv[r][0] * x[0] + inv[r][1] * x[1] + inv[r][2] * x[2]);
    for (int r = 0; r < 3; ++r) x[r] = m[r][0] * f[0] + m[r][1] * f[1] + m[r][2] * f[2];
  }
}

// api.f90:1517-1530 verify_structure.  Small structures test every pair;
// larger ones sort the atoms by x and sweep: only atoms whose x coordinates
// differ by less than 1e-9 bohr can be closer than that, so the test is
// O(N log N) for any structure without many atoms in one thin slab.
bool too_close(int nat, const double* xyz) {
  const double thr = 1.0e-9;
  if (nat < 2) return false;
  if (nat <= 2048) {
    for (int i = 0; i < nat; ++i)
      for (int j = 0; j < i; ++j) {
        const double d0 = xyz[3 * j] - xyz[3 * i], d1 = xyz[3 * j + 1] - xyz[3 * i + 1],
                     d2 = xyz[3 * j + 2] - xyz[3 * i + 2];
        if (std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) < thr) return true;
      }
    return false;
  }
  // sort by x and compare neighbours within thr in x (sweep)
  std::vector<int> idx(nat);
  for (int i = 0; i < nat; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return xyz[3 * a] < xyz[3 * b]; });
  for (int p = 0; p < nat; ++p) {
    const int i = idx[p];
    for (int q = p + 1; q < nat; ++q) {
      const int j = idx[q];
      if (xyz[3 * j] - xyz[3 * i] >= thr) break;
      const double d0 = xyz[3 * j] - xyz[3 * i], d1 = xyz[3 * j + 1] - xyz[3 * i + 1],
                   d2 = xyz[3 * j + 2] - xyz[3 * i + 2];
      if (std::sqrt(d0 * d0 + d1 * d1 + d2 * d2) < thr) return true;
    }
  }
  return false;
}

// param.f90:96-300 get_method_id: drop '-' and compare lower case.
const char* method_id(const char* method) {
  std::string key;
  for (const char* c = method; *c; ++c)
    if (*c != '-') key.push_back(static_cast<char>(std::tolower(static_cast<unsigned char>(*c))));
  while (!key.empty() && key.back() == ' ') key.pop_back();
  for (const auto& a : kMethodAliases)
    if (key == a.name) return a.id;
  return nullptr;
}

const ParamRecord* find_record(const ParamRecord* table, int n, const char* id) {
  if (!id) return nullptr;
  for (int k = 0; k < n; ++k)
    if (std::strcmp(table[k].id, id) == 0) return &table[k];
  return nullptr;
}

void ensure_model_buffers(_dftd3_model* m, int nat) {
  if (!m->stream) d3::hip_check(hipStreamCreate(&m->stream), "hipStreamCreate");
  if (static_cast<size_t>(nat) > m->cap_atoms || !m->d_xyz) {
    if (m->d_xyz) d3::hip_check(hipFree(m->d_xyz), "hipFree");
    if (m->d_out) d3::hip_check(hipFree(m->d_out), "hipFree");
    m->d_xyz = nullptr;
    m->d_out = nullptr;
    d3::hip_check(hipMalloc(&m->d_xyz, sizeof(double) * 3 * std::max(nat, 1)), "hipMalloc");
    d3::hip_check(hipMalloc(&m->d_out, sizeof(double) * (10 + 3 * std::max(nat, 1))), "hipMalloc");
    m->cap_atoms = nat;
  }
}

// Device buffer of n x n results (pair energies, Hessian) of the host entry points.
void ensure_result_buffer(_dftd3_model* m, size_t words) {
  if (words <= m->cap_pairs && m->d_pairs) return;
  if (m->d_pairs) d3::hip_check(hipFree(m->d_pairs), "hipFree");
  m->d_pairs = nullptr;
  m->cap_pairs = 0;
  d3::hip_check(hipMalloc(&m->d_pairs, sizeof(double) * std::max<size_t>(words, 1)), "hipMalloc");
  m->cap_pairs = words;
}

d3::SysParam make_sys_param(const _dftd3_param* p, const double* cutoff) {
  d3::SysParam sp{};
  sp.kind = p->kind;
  sp.s6 = p->s6; sp.s8 = p->s8; sp.s9 = p->s9; sp.a1 = p->a1; sp.a2 = p->a2; sp.alp = p->alp;
  sp.bet = p->bet; sp.a3 = p->a3; sp.a4 = p->a4;
  sp.cut_cn = cutoff[0]; sp.cut2 = cutoff[1]; sp.cut3 = cutoff[2];
  sp.width2 = cutoff[3]; sp.width3 = cutoff[4];
  return sp;
}

}  // namespace

// ------------------------------------------------------------ error handle
int dftd3_get_version(void) { return 10600; }

// feature.F90:37-52: "mpi" is the only optional feature; this library has no
// MPI backend, so every feature is unavailable.
bool dftd3_has_feature(char* feature) {
  (void)feature;
  return false;
}

dftd3_error dftd3_new_error(void) { return new _dftd3_error(); }

int dftd3_check_error(dftd3_error error) { return error && error->set ? 1 : 0; }

void dftd3_get_error(dftd3_error error, char* buffer, const int* buffersize) {
  if (!error || !buffer) return;
  size_t maxlen = buffersize ? static_cast<size_t>(std::max(*buffersize, 1)) : error->message.size() + 1;
  size_t n = std::min(error->message.size(), maxlen - 1);
  std::memcpy(buffer, error->message.data(), n);
  buffer[n] = '\0';
}

void dftd3_delete_error(dftd3_error* error) {
  if (error && *error) {
    delete *error;
    *error = nullptr;
  }
}

// ------------------------------------------------------------- structure
dftd3_structure dftd3_new_structure(dftd3_error error, const int natoms, const int* numbers,
                                    const double* positions, const double* lattice, const bool* periodic) {
  if (!error) return nullptr;
  dftd3_structure mol = nullptr;
  guarded(error, [&] {
    if (natoms <= 0 || !numbers || !positions) throw d3::Error("Invalid molecular structure data provided");
    mol = new _dftd3_structure();
    mol->nat = natoms;
    mol->num.assign(numbers, numbers + natoms);
    mol->xyz.assign(positions, positions + 3 * natoms);
    if (lattice) std::memcpy(mol->lattice, lattice, sizeof mol->lattice);
    if (periodic) {
      for (int k = 0; k < 3; ++k) mol->periodic[k] = periodic[k] ? 1 : 0;
    } else if (lattice) {
      mol->periodic[0] = mol->periodic[1] = mol->periodic[2] = 1;
    }
    wrap_to_central_cell(natoms, mol->xyz.data(), mol->lattice, mol->periodic);
    if (too_close(natoms, mol->xyz.data())) fatal(error, "Too close interatomic distances found");
  });
  return mol;
}

void dftd3_delete_structure(dftd3_structure* mol) {
  if (mol && *mol) {
    delete *mol;
    *mol = nullptr;
  }
}

void dftd3_update_structure(dftd3_error error, dftd3_structure mol, const double* positions,
                            const double* lattice) {
  if (!error) return;
  if (!mol) return fatal(error, "Molecular structure data is missing");
  if (mol->nat <= 0 || mol->num.empty()) return fatal(error, "Invalid molecular structure data provided");
  guarded(error, [&] {
    std::memcpy(mol->xyz.data(), positions, sizeof(double) * 3 * mol->nat);
    if (lattice) std::memcpy(mol->lattice, lattice, sizeof mol->lattice);
    wrap_to_central_cell(mol->nat, mol->xyz.data(), mol->lattice, mol->periodic);
    if (too_close(mol->nat, mol->xyz.data())) fatal(error, "Too close interatomic distances found");
  });
}

// ----------------------------------------------------------------- model
dftd3_model dftd3_new_d3_model(dftd3_error error, dftd3_structure mol) {
  if (!error) return nullptr;
  if (!mol) {
    fatal(error, "Molecular structure data is missing");
    return nullptr;
  }
  dftd3_model disp = nullptr;
  guarded(error, [&] {
    for (int z : mol->num)
      if (z < 1 || z > d3::kMaxElem) throw d3::Error("Unsupported element with atomic number " + std::to_string(z));
    auto m = std::make_unique<_dftd3_model>();
    m->has_structure = true;
    m->nat = mol->nat;
    m->num = mol->num;
    m->ghost.assign(mol->nat, 0);
    m->engine = std::make_unique<d3::Engine>();
    disp = m.release();
  });
  return disp;
}

dftd3_model d3hip_new_workspace(dftd3_error error) {
  if (!error) return nullptr;
  dftd3_model disp = nullptr;
  guarded(error, [&] {
    auto m = std::make_unique<_dftd3_model>();
    m->engine = std::make_unique<d3::Engine>();
    disp = m.release();
  });
  return disp;
}

void dftd3_set_model_realspace_cutoff(dftd3_error error, dftd3_model model, double disp2, double disp3, double cn) {
  dftd3_set_model_realspace_cutoff_smooth(error, model, disp2, disp3, cn, 0.0, 0.0);
}

void dftd3_set_model_realspace_cutoff_smooth(dftd3_error error, dftd3_model model, double disp2, double disp3,
                                             double cn, double width2, double width3) {
  if (!error) return;
  if (!model) return fatal(error, "D3 dispersion model is missing");
  model->cutoff[0] = cn;
  model->cutoff[1] = disp2;
  model->cutoff[2] = disp3;
  model->cutoff[3] = width2;
  model->cutoff[4] = width3;
}

void dftd3_set_model_ewald(dftd3_error error, dftd3_model model, int rank, double tolerance, double kcut,
                           int mesh) {
  // api.f90:425-453: non-positive values select the defaults (rank from the
  // tolerance 1e-4, cutoff from the damping radii); mesh < 0 sums the
  // vectors directly, 0 uses a particle mesh derived from the cutoff, > 0 a
  // mesh of that many points per direction (rounded up to a power of two).
  // The expansion of the model's species is made here, once
  // (model.f90:185-196).
  if (!error) return;
  if (!model) return fatal(error, "D3 dispersion model is missing");
  guarded(error, [&] {
    std::vector<int> species;
    for (int z : model->num)
      if (std::find(species.begin(), species.end(), z) == species.end()) species.push_back(z);
    auto lr = std::make_unique<d3::LowRank>(d3::make_lowrank(species, rank > 0 ? rank : 0,
                                                             tolerance > 0.0 ? tolerance : 1.0e-4));
    lr->kcut = kcut > 0.0 ? kcut : 0.0;
    lr->mesh = mesh;
    model->lowrank = std::move(lr);
  });
}

void dftd3_set_model_work_partition(dftd3_error error, dftd3_model model, int part, int nparts) {
  // api.f90:459-480, partition.f90:57-84: parts are zero based; an invalid
  // partition is an error and leaves the model serial (the partition is
  // reinitialised before the check)
  if (!error) return;
  if (!model) return fatal(error, "D3 dispersion model is missing");
  model->part = 0;
  model->nparts = 1;
  if (nparts <= 0 || part < 0 || part >= nparts) return fatal(error, "Invalid dispersion work partition");
  model->part = part;
  model->nparts = nparts;
}

void dftd3_set_model_mpi_comm(dftd3_error error, dftd3_model model, int comm) {
  // api.f90:484-507 in a build without MPI (mpi_utils.F90:39-92): the
  // partition is reset and the missing backend reported
  (void)comm;
  if (!error) return;
  if (!model) return fatal(error, "D3 dispersion model is missing");
  model->part = 0;
  model->nparts = 1;
  fatal(error, "s-dftd3 was built without MPI support");
}

void d3hip_get_model_lowrank(dftd3_error error, dftd3_model model, int* rank, double* max_error, double* kcut,
                             long long* vectors) {
  if (!error) return;
  if (!model) return fatal(error, "D3 dispersion model is missing");
  const d3::LowRank* lr = model->lowrank.get();
  if (rank) *rank = lr ? lr->rank : 0;
  if (max_error) *max_error = lr ? lr->error : 0.0;
  if (kcut) *kcut = lr ? lr->last_kcut : 0.0;
  if (vectors) *vectors = lr ? lr->last_vectors : 0;
}

void dftd3_set_model_ghost_index(dftd3_error error, dftd3_model model, const int* ghost, int nidx) {
  if (!error) return;
  if (!model) return fatal(error, "D3 dispersion model is missing");
  std::fill(model->ghost.begin(), model->ghost.end(), 0);
  for (int k = 0; k < nidx; ++k)  // model.f90:166-174: out-of-range indices are ignored
    if (ghost[k] > 0 && ghost[k] <= model->nat) model->ghost[ghost[k] - 1] = 1;
}

void d3hip_set_model_precision(dftd3_error error, dftd3_model model, int precision) {
  if (!error) return;
  if (!model) return fatal(error, "D3 dispersion model is missing");
  if (precision != d3::kFP64 && precision != d3::kMixed) return fatal(error, "Unknown precision mode");
  model->precision = precision;
}

void dftd3_delete_model(dftd3_model* disp) {
  if (disp && *disp) {
    delete *disp;
    *disp = nullptr;
  }
}

// ------------------------------------------------------------ parameters
dftd3_param dftd3_new_rational_damping(dftd3_error error, double s6, double s8, double s9, double a1, double a2,
                                       double alp) {
  if (!error) return nullptr;
  auto* p = new _dftd3_param();
  p->kind = d3::kRational;
  p->s6 = s6; p->s8 = s8; p->s9 = s9; p->a1 = a1; p->a2 = a2; p->alp = alp;
  return p;
}

dftd3_param dftd3_new_zero_damping(dftd3_error error, double s6, double s8, double s9, double rs6, double rs8,
                                   double alp) {
  if (!error) return nullptr;
  auto* p = new _dftd3_param();
  p->kind = d3::kZero;
  p->s6 = s6; p->s8 = s8; p->s9 = s9; p->a1 = rs6; p->a2 = rs8; p->alp = alp;
  return p;
}

// api.f90:748-777 (rs6, rs8 in a1, a2)
dftd3_param dftd3_new_mzero_damping(dftd3_error error, double s6, double s8, double s9, double rs6, double rs8,
                                    double alp, double bet) {
  if (!error) return nullptr;
  auto* p = new _dftd3_param();
  p->kind = d3::kModifiedZero;
  p->s6 = s6; p->s8 = s8; p->s9 = s9; p->a1 = rs6; p->a2 = rs8; p->alp = alp; p->bet = bet;
  return p;
}

// api.f90:694-708: the rational damping with other parameters
dftd3_param dftd3_new_mrational_damping(dftd3_error error, double s6, double s8, double s9, double a1, double a2,
                                        double alp) {
  return dftd3_new_rational_damping(error, s6, s8, s9, a1, a2, alp);
}

// api.f90:817-846
dftd3_param dftd3_new_optimizedpower_damping(dftd3_error error, double s6, double s8, double s9, double a1,
                                             double a2, double alp, double bet) {
  if (!error) return nullptr;
  auto* p = new _dftd3_param();
  p->kind = d3::kOptimizedPower;
  p->s6 = s6; p->s8 = s8; p->s9 = s9; p->a1 = a1; p->a2 = a2; p->alp = alp; p->bet = bet;
  return p;
}

// api.f90:886-915 (cso.f90:74-90: a3, a4 are d3_param's rs6, rs8; no s8)
dftd3_param dftd3_new_cso_damping(dftd3_error error, double s6, double s9, double a1, double a2, double a3,
                                  double a4, double alp) {
  if (!error) return nullptr;
  auto* p = new _dftd3_param();
  p->kind = d3::kCSO;
  p->s6 = s6; p->s9 = s9; p->a1 = a1; p->a2 = a2; p->a3 = a3; p->a4 = a4; p->alp = alp;
  return p;
}

// api.f90:955-982
dftd3_param dftd3_new_z_damping(dftd3_error error, double s6, double s8, double s9, double a1, double alp) {
  if (!error) return nullptr;
  auto* p = new _dftd3_param();
  p->kind = d3::kZDamping;
  p->s6 = s6; p->s8 = s8; p->s9 = s9; p->a1 = a1; p->alp = alp;
  return p;
}

namespace {
// Parameter tables of param.f90, one per loader.
enum class Family { kRational, kZero, kMRational, kMZero, kOptimizedPower, kCSO, kZ };

dftd3_param load_damping(dftd3_error error, const char* method, bool atm, Family family) {
  if (!error) return nullptr;
  if (!method) {
    fatal(error, "No entry for '' present");
    return nullptr;
  }
  const char* id = method_id(method);
  const ParamRecord* rec = nullptr;
  switch (family) {
    case Family::kRational: rec = find_record(kRationalParams, kNumRational, id); break;
    case Family::kZero: rec = find_record(kZeroParams, kNumZero, id); break;
    case Family::kMRational: rec = find_record(kMRationalParams, kNumMRational, id); break;
    case Family::kMZero: rec = find_record(kMZeroParams, kNumMZero, id); break;
    case Family::kOptimizedPower: rec = find_record(kOptimizedPowerParams, kNumOptimizedPower, id); break;
    case Family::kCSO: rec = find_record(kCSOParams, kNumCSO, id); break;
    case Family::kZ: rec = find_record(kZParams, kNumZ, id); break;
  }
  if (!rec) {
    fatal(error, std::string("No entry for '") + method + "' present");
    return nullptr;
  }
  auto* p = new _dftd3_param();
  p->s6 = rec->s6;
  p->s8 = rec->s8;
  p->s9 = atm ? 1.0 : rec->s9;  // api.f90:606 (atm -> s9 = 1), param.f90:798-800
  p->alp = rec->alp;
  p->bet = rec->bet;
  p->a1 = rec->a1;
  p->a2 = rec->a2;
  switch (family) {
    case Family::kRational:
    case Family::kMRational: p->kind = d3::kRational; break;
    case Family::kZero: p->kind = d3::kZero; p->a1 = rec->rs6; p->a2 = rec->rs8; break;
    case Family::kMZero: p->kind = d3::kModifiedZero; p->a1 = rec->rs6; p->a2 = rec->rs8; break;
    case Family::kOptimizedPower: p->kind = d3::kOptimizedPower; break;
    case Family::kCSO: p->kind = d3::kCSO; p->a3 = rec->rs6; p->a4 = rec->rs8; break;
    case Family::kZ: p->kind = d3::kZDamping; break;
  }
  return p;
}
}  // namespace

dftd3_param dftd3_load_rational_damping(dftd3_error error, char* method, bool atm) {
  return load_damping(error, method, atm, Family::kRational);
}

dftd3_param dftd3_load_zero_damping(dftd3_error error, char* method, bool atm) {
  return load_damping(error, method, atm, Family::kZero);
}

dftd3_param dftd3_load_mzero_damping(dftd3_error error, char* method, bool atm) {
  return load_damping(error, method, atm, Family::kMZero);
}

dftd3_param dftd3_load_mrational_damping(dftd3_error error, char* method, bool atm) {
  return load_damping(error, method, atm, Family::kMRational);
}

dftd3_param dftd3_load_optimizedpower_damping(dftd3_error error, char* method, bool atm) {
  return load_damping(error, method, atm, Family::kOptimizedPower);
}

dftd3_param dftd3_load_cso_damping(dftd3_error error, char* method, bool atm) {
  return load_damping(error, method, atm, Family::kCSO);
}

dftd3_param dftd3_load_z_damping(dftd3_error error, char* method, bool atm) {
  return load_damping(error, method, atm, Family::kZ);
}

void dftd3_delete_param(dftd3_param* param) {
  if (param && *param) {
    delete *param;
    *param = nullptr;
  }
}

// ------------------------------------------------------------ evaluation
void dftd3_get_dispersion(dftd3_error error, dftd3_structure mol, dftd3_model disp, dftd3_param param,
                          double* energy, double* gradient, double* sigma) {
  // api.f90:1038-1111
  if (!error) return;
  if (!mol) return fatal(error, "Molecular structure data is missing");
  if (!disp) return fatal(error, "Dispersion model is missing");
  if (!param) return fatal(error, "Damping parameters are missing");
  if (!disp->has_structure || disp->nat != mol->nat)
    return fatal(error, "Dispersion model does not match the molecular structure");
  guarded(error, [&] {
    const int nat = mol->nat;
    ensure_model_buffers(disp, nat);
    std::vector<d3::SystemDesc> sys(1);
    sys[0].natoms = nat;
    std::memcpy(sys[0].lattice, mol->lattice, sizeof mol->lattice);
    std::memcpy(sys[0].periodic, mol->periodic, sizeof mol->periodic);
    sys[0].param = make_sys_param(param, disp->cutoff);
    sys[0].lowrank = disp->lowrank.get();
    sys[0].part = disp->part;  // api.f90:1099
    sys[0].nparts = disp->nparts;
    const bool grad = gradient || sigma;
    d3::hip_check(hipMemcpyAsync(disp->d_xyz, mol->xyz.data(), sizeof(double) * 3 * nat, hipMemcpyHostToDevice,
                                 disp->stream),
                  "upload positions");
    d3::BatchOutput out;
    out.energy = disp->d_out;
    out.sigma = grad ? disp->d_out + 1 : nullptr;
    out.gradient = grad ? disp->d_out + 10 : nullptr;
    bool any_ghost = false;
    for (int g : disp->ghost) any_ghost |= g != 0;
    disp->engine->run(sys, mol->num.data(), disp->d_xyz, disp->precision, out, disp->stream, nullptr, nullptr,
                      any_ghost ? disp->ghost.data() : nullptr);
    std::vector<double> host(grad ? 10 + 3 * nat : 1);
    d3::hip_check(hipMemcpyAsync(host.data(), disp->d_out, sizeof(double) * host.size(), hipMemcpyDeviceToHost,
                                 disp->stream),
                  "download results");
    d3::hip_check(hipStreamSynchronize(disp->stream), "hipStreamSynchronize");
    if (energy) *energy = host[0];
    if (sigma) std::memcpy(sigma, host.data() + 1, sizeof(double) * 9);
    if (gradient) std::memcpy(gradient, host.data() + 10, sizeof(double) * 3 * nat);
  });
}

void dftd3_get_pairwise_dispersion(dftd3_error error, dftd3_structure mol, dftd3_model disp, dftd3_param param,
                                   double* pair_energy2, double* pair_energy3) {
  // api.f90:1180-1233
  if (!error) return;
  if (!mol) return fatal(error, "Molecular structure data is missing");
  if (!disp) return fatal(error, "Dispersion model is missing");
  if (!param) return fatal(error, "Damping parameters are missing");
  if (!disp->has_structure || disp->nat != mol->nat)
    return fatal(error, "Dispersion model does not match the molecular structure");
  if (!pair_energy2 || !pair_energy3) return fatal(error, "Pairwise energy arrays are missing");
  guarded(error, [&] {
    const int nat = mol->nat;
    ensure_model_buffers(disp, nat);
    const size_t nn = static_cast<size_t>(nat) * static_cast<size_t>(nat);
    ensure_result_buffer(disp, 2 * nn);
    std::vector<d3::SystemDesc> sys(1);
    sys[0].natoms = nat;
    std::memcpy(sys[0].lattice, mol->lattice, sizeof mol->lattice);
    std::memcpy(sys[0].periodic, mol->periodic, sizeof mol->periodic);
    sys[0].param = make_sys_param(param, disp->cutoff);
    sys[0].lowrank = disp->lowrank.get();
    d3::hip_check(hipMemcpyAsync(disp->d_xyz, mol->xyz.data(), sizeof(double) * 3 * nat, hipMemcpyHostToDevice,
                                 disp->stream),
                  "upload positions");
    bool any_ghost = false;
    for (int g : disp->ghost) any_ghost |= g != 0;
    disp->engine->run_pairwise(sys, mol->num.data(), disp->d_xyz, disp->precision, disp->d_pairs,
                               disp->d_pairs + nn, disp->stream, any_ghost ? disp->ghost.data() : nullptr);
    d3::hip_check(hipMemcpyAsync(pair_energy2, disp->d_pairs, sizeof(double) * nn, hipMemcpyDeviceToHost,
                                 disp->stream),
                  "download results");
    d3::hip_check(hipMemcpyAsync(pair_energy3, disp->d_pairs + nn, sizeof(double) * nn, hipMemcpyDeviceToHost,
                                 disp->stream),
                  "download results");
    d3::hip_check(hipStreamSynchronize(disp->stream), "hipStreamSynchronize");
  });
}

void dftd3_get_dispersion_hessian(dftd3_error error, dftd3_structure mol, dftd3_model disp, dftd3_param param,
                                  double* energy, double* hessian) {
  // api.f90:1115-1176
  if (!error) return;
  if (!mol) return fatal(error, "Molecular structure data is missing");
  if (!disp) return fatal(error, "Dispersion model is missing");
  if (!param) return fatal(error, "Damping parameters are missing");
  if (!disp->has_structure || disp->nat != mol->nat)
    return fatal(error, "Dispersion model does not match the molecular structure");
  if (!hessian) return fatal(error, "Hessian array is missing");
  guarded(error, [&] {
    const int nat = mol->nat;
    ensure_model_buffers(disp, nat);
    const size_t nh = static_cast<size_t>(3 * nat) * static_cast<size_t>(3 * nat);
    ensure_result_buffer(disp, nh);
    std::vector<d3::SystemDesc> sys(1);
    sys[0].natoms = nat;
    std::memcpy(sys[0].lattice, mol->lattice, sizeof mol->lattice);
    std::memcpy(sys[0].periodic, mol->periodic, sizeof mol->periodic);
    sys[0].param = make_sys_param(param, disp->cutoff);
    sys[0].lowrank = disp->lowrank.get();
    sys[0].part = disp->part;  // api.f90:1170
    sys[0].nparts = disp->nparts;
    d3::hip_check(hipMemcpyAsync(disp->d_xyz, mol->xyz.data(), sizeof(double) * 3 * nat, hipMemcpyHostToDevice,
                                 disp->stream),
                  "upload positions");
    bool any_ghost = false;
    for (int g : disp->ghost) any_ghost |= g != 0;
    disp->engine->run_hessian(sys, mol->num.data(), disp->d_xyz, disp->d_out, disp->d_pairs, disp->stream,
                              any_ghost ? disp->ghost.data() : nullptr);
    double e = 0.0;
    d3::hip_check(hipMemcpyAsync(&e, disp->d_out, sizeof(double), hipMemcpyDeviceToHost, disp->stream),
                  "download results");
    d3::hip_check(hipMemcpyAsync(hessian, disp->d_pairs, sizeof(double) * nh, hipMemcpyDeviceToHost, disp->stream),
                  "download results");
    d3::hip_check(hipStreamSynchronize(disp->stream), "hipStreamSynchronize");
    if (energy) *energy = e;
  });
}

namespace {
constexpr int kParamStride = 10;  // {kind, s6, s8, s9, a1|rs6, a2|rs8, alp, bet, a3, a4}

std::vector<d3::SystemDesc> batch_systems(int nsys, const int* natoms, const double* lattices, const bool* periodic,
                                          const double* params, const double* cutoffs) {
  std::vector<d3::SystemDesc> sys(nsys);
  for (int k = 0; k < nsys; ++k) {
    d3::SystemDesc& s = sys[k];
    s.natoms = natoms[k];
    if (s.natoms <= 0) throw d3::Error("Invalid molecular structure data provided");
    if (lattices) std::memcpy(s.lattice, lattices + 9 * k, sizeof s.lattice);
    if (periodic)
      for (int c = 0; c < 3; ++c) s.periodic[c] = periodic[3 * k + c] ? 1 : 0;
    const double* p = params + kParamStride * k;
    _dftd3_param prm;
    prm.kind = static_cast<int>(p[0]);
    if (prm.kind < d3::kRational || prm.kind > d3::kZDamping || p[0] != static_cast<double>(prm.kind))
      throw d3::Error("Unknown damping kind");
    prm.s6 = p[1]; prm.s8 = p[2]; prm.s9 = p[3]; prm.a1 = p[4]; prm.a2 = p[5]; prm.alp = p[6];
    prm.bet = p[7]; prm.a3 = p[8]; prm.a4 = p[9];
    double cut[5] = {40.0, 60.0, 40.0, 0.0, 0.0};
    if (cutoffs) {
      const double* c = cutoffs + 5 * k;
      cut[0] = c[2]; cut[1] = c[0]; cut[2] = c[1]; cut[3] = c[3]; cut[4] = c[4];
    }
    s.param = make_sys_param(&prm, cut);
  }
  return sys;
}
}  // namespace

void d3hip_get_dispersion_batch(dftd3_error error, dftd3_model workspace, int nsys, const int* natoms,
                                const int* numbers, const double* lattices, const bool* periodic,
                                const double* params, const double* cutoffs, int precision,
                                const double* positions, double* energy, double* gradient, double* sigma,
                                void* stream) {
  if (!error) return;
  if (!workspace) return fatal(error, "Dispersion model is missing");
  if (nsys <= 0 || !natoms || !numbers || !params || !positions || !energy)
    return fatal(error, "Invalid batch description");
  if (precision != d3::kFP64 && precision != d3::kMixed) return fatal(error, "Unknown precision mode");
  guarded(error, [&] {
    std::vector<d3::SystemDesc> sys = batch_systems(nsys, natoms, lattices, periodic, params, cutoffs);
    d3::BatchOutput out;
    out.energy = energy;
    out.gradient = gradient;
    out.sigma = sigma;
    workspace->engine->run(sys, numbers, positions, precision, out, static_cast<hipStream_t>(stream));
  });
}

static_assert(D3HIP_NOALLOC_BYTES == d3::kNoAllocBytes, "work area size of the header and the library");

void d3hip_get_dispersion_batch_noalloc(dftd3_error error, int nsys, const int* natoms, const int* numbers,
                                        const double* lattices, const bool* periodic, const double* params,
                                        const double* cutoffs, int precision, const double* positions,
                                        double* energy, double* gradient, double* sigma, void* stream) {
  if (!error) return;
  if (nsys <= 0 || !natoms || !numbers || !params || !positions || !energy)
    return fatal(error, "Invalid batch description");
  if (precision != d3::kFP64 && precision != d3::kMixed) return fatal(error, "Unknown precision mode");
  guarded(error, [&] {
    std::vector<d3::SystemDesc> sys = batch_systems(nsys, natoms, lattices, periodic, params, cutoffs);
    d3::BatchOutput out;
    out.energy = energy;
    out.gradient = gradient;
    out.sigma = sigma;
    d3::run_noalloc(sys, numbers, positions, precision, out, static_cast<hipStream_t>(stream));
  });
}

void d3hip_get_pairwise_dispersion_batch(dftd3_error error, dftd3_model workspace, int nsys, const int* natoms,
                                         const int* numbers, const double* lattices, const bool* periodic,
                                         const double* params, const double* cutoffs, int precision,
                                         const double* positions, double* pair_energy2, double* pair_energy3,
                                         void* stream) {
  if (!error) return;
  if (!workspace) return fatal(error, "Dispersion model is missing");
  if (nsys <= 0 || !natoms || !numbers || !params || !positions || !pair_energy2 || !pair_energy3)
    return fatal(error, "Invalid batch description");
  if (precision != d3::kFP64 && precision != d3::kMixed) return fatal(error, "Unknown precision mode");
  guarded(error, [&] {
    std::vector<d3::SystemDesc> sys = batch_systems(nsys, natoms, lattices, periodic, params, cutoffs);
    workspace->engine->run_pairwise(sys, numbers, positions, precision, pair_energy2, pair_energy3,
                                    static_cast<hipStream_t>(stream));
  });
}

void d3hip_get_dispersion_hessian_batch(dftd3_error error, dftd3_model workspace, int nsys, const int* natoms,
                                        const int* numbers, const double* lattices, const bool* periodic,
                                        const double* params, const double* cutoffs, const double* positions,
                                        double* energy, double* hessian, void* stream) {
  if (!error) return;
  if (!workspace) return fatal(error, "Dispersion model is missing");
  if (nsys <= 0 || !natoms || !numbers || !params || !positions || !hessian)
    return fatal(error, "Invalid batch description");
  guarded(error, [&] {
    std::vector<d3::SystemDesc> sys = batch_systems(nsys, natoms, lattices, periodic, params, cutoffs);
    workspace->engine->run_hessian(sys, numbers, positions, energy, hessian, static_cast<hipStream_t>(stream));
  });
}

void d3hip_count_terms_batch(dftd3_error error, dftd3_model workspace, int nsys, const int* natoms,
                             const int* numbers, const double* lattices, const bool* periodic,
                             const double* params, const double* cutoffs, const double* positions,
                             long long* counts, void* stream) {
  if (!error) return;
  if (!workspace) return fatal(error, "Dispersion model is missing");
  if (nsys <= 0 || !natoms || !numbers || !params || !positions || !counts)
    return fatal(error, "Invalid batch description");
  guarded(error, [&] {
    std::vector<d3::SystemDesc> sys = batch_systems(nsys, natoms, lattices, periodic, params, cutoffs);
    workspace->engine->count_terms(sys, numbers, positions, counts, static_cast<hipStream_t>(stream));
  });
}

void d3hip_set_kernel_timing(dftd3_error error, dftd3_model model, int enable) {
  if (!error) return;
  if (!model) return fatal(error, "Dispersion model is missing");
  model->engine->set_kernel_timing(enable != 0);
}

void d3hip_get_kernel_time(dftd3_error error, dftd3_model model, const char* name, double* total_ms,
                           long long* calls) {
  if (!error) return;
  if (!model) return fatal(error, "Dispersion model is missing");
  if (!name) return fatal(error, "Kernel name is missing");
  auto t = model->engine->kernel_time(name);
  if (total_ms) *total_ms = t.first;
  if (calls) *calls = t.second;
}

// ------------------------------------------------------------------ plans
struct _d3hip_plan {
  std::unique_ptr<d3::Plan> plan;
};

namespace {
d3::BatchOutput outputs(double* energy, double* gradient, double* sigma) {
  d3::BatchOutput out;
  out.energy = energy;
  out.gradient = gradient;
  out.sigma = sigma;
  return out;
}
}  // namespace

d3hip_plan d3hip_new_plan(dftd3_error error, int nsys, const int* natoms, const int* numbers,
                          const double* lattices, const bool* periodic, const double* params, const double* cutoffs,
                          int precision, int grad, const double* positions, void* stream) {
  if (!error) return nullptr;
  if (nsys <= 0 || !natoms || !numbers || !params || !positions) {
    fatal(error, "Invalid batch description");
    return nullptr;
  }
  if (precision != d3::kFP64 && precision != d3::kMixed) {
    fatal(error, "Unknown precision mode");
    return nullptr;
  }
  d3hip_plan h = nullptr;
  guarded(error, [&] {
    std::vector<d3::SystemDesc> sys = batch_systems(nsys, natoms, lattices, periodic, params, cutoffs);
    std::unique_ptr<_d3hip_plan> p(new _d3hip_plan);
    p->plan.reset(new d3::Plan(sys, numbers, positions, precision, grad != 0, nullptr,
                               static_cast<hipStream_t>(stream)));
    h = p.release();
  });
  return h;
}

void d3hip_delete_plan(d3hip_plan* plan) {
  if (!plan || !*plan) return;
  delete *plan;
  *plan = nullptr;
}

void d3hip_plan_execute(dftd3_error error, d3hip_plan plan, const double* positions, double* energy,
                        double* gradient, double* sigma, void* stream) {
  if (!error) return;
  if (!plan) return fatal(error, "Plan is missing");
  if (!positions || !energy) return fatal(error, "Invalid batch description");
  guarded(error, [&] {
    plan->plan->execute(positions, outputs(energy, gradient, sigma), static_cast<hipStream_t>(stream));
  });
}

int d3hip_plan_num_shapes(void) { return static_cast<int>(d3::Plan::candidate_shapes().size()); }

void d3hip_plan_candidate_shape(int index, int* waves, int* blocks_per_cu) {
  const auto c = d3::Plan::candidate_shapes();
  if (index < 0 || index >= static_cast<int>(c.size())) return;
  if (waves) *waves = c[index].waves;
  if (blocks_per_cu) *blocks_per_cu = c[index].blocks_per_cu;
}

void d3hip_plan_get_shape(dftd3_error error, d3hip_plan plan, int* waves, int* blocks_per_cu) {
  if (!error) return;
  if (!plan) return fatal(error, "Plan is missing");
  if (waves) *waves = plan->plan->shape().waves;
  if (blocks_per_cu) *blocks_per_cu = plan->plan->shape().blocks_per_cu;
}

void d3hip_plan_set_shape(dftd3_error error, d3hip_plan plan, int waves, int blocks_per_cu) {
  if (!error) return;
  if (!plan) return fatal(error, "Plan is missing");
  guarded(error, [&] { plan->plan->set_shape(waves, blocks_per_cu); });
}

int d3hip_plan_autotune(dftd3_error error, d3hip_plan plan, const double* positions, double* energy,
                        double* gradient, double* sigma, int reps, float* times_ms, void* stream) {
  if (!error) return -1;
  if (!plan) {
    fatal(error, "Plan is missing");
    return -1;
  }
  if (!positions || !energy) {
    fatal(error, "Invalid batch description");
    return -1;
  }
  int best = -1;
  guarded(error, [&] {
    std::vector<float> t;
    best = plan->plan->autotune(positions, outputs(energy, gradient, sigma), static_cast<hipStream_t>(stream),
                                reps, &t);
    if (times_ms) std::copy(t.begin(), t.end(), times_ms);
  });
  return best;
}

void d3hip_plan_capture(dftd3_error error, d3hip_plan plan, const double* positions, double* energy,
                        double* gradient, double* sigma, void* stream) {
  if (!error) return;
  if (!plan) return fatal(error, "Plan is missing");
  if (!positions || !energy) return fatal(error, "Invalid batch description");
  guarded(error, [&] {
    plan->plan->capture(positions, outputs(energy, gradient, sigma), static_cast<hipStream_t>(stream));
  });
}

void d3hip_plan_replay(dftd3_error error, d3hip_plan plan, void* stream) {
  if (!error) return;
  if (!plan) return fatal(error, "Plan is missing");
  guarded(error, [&] { plan->plan->replay(static_cast<hipStream_t>(stream)); });
}

int d3hip_plan_overflowed(dftd3_error error, d3hip_plan plan, void* stream) {
  if (!error) return -1;
  if (!plan) {
    fatal(error, "Plan is missing");
    return -1;
  }
  int r = -1;
  guarded(error, [&] { r = plan->plan->overflowed(static_cast<hipStream_t>(stream)) ? 1 : 0; });
  return r;
}
