// Host set-up of the reciprocal-space two-body term: the low-rank expansion
// of the reference C6 coefficients of a structure's species, the damped pair
// potential as a sum of terms r^m / (r^alpha + R^alpha) per species pair, and
// the reciprocal lattice vectors within the cutoff.  All of it is per model
// (once per species set) or per geometry (the vectors); the sums over atoms
// and vectors run on the device (d3_kernels.hip).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "d3_engine.hpp"

namespace d3 {

namespace {

constexpr double kEps = std::numeric_limits<double>::epsilon();

// Eigenpairs of a symmetric n x n matrix (column-major) by cyclic Jacobi
// rotations, sweeping the upper triangle row by row, at most 100 sweeps;
// then sorted by decreasing |eigenvalue| with a selection swap (the first of
// equal magnitudes wins).  fourier/jacobi.f90:40-122: the truncation of the
// expansion depends on the order of the eigenpairs, so the sweep order,
// skip threshold (1e-2 eps ||A||) and convergence test (sqrt(2 off) <=
// eps ||A||) are the reference's.
void jacobi(int n, const std::vector<double>& a_in, std::vector<double>& eval, std::vector<double>& evec) {
  auto at = [n](std::vector<double>& m, int r, int c) -> double& { return m[size_t(r) + size_t(c) * n]; };
  evec.assign(size_t(n) * n, 0.0);
  for (int p = 0; p < n; ++p) at(evec, p, p) = 1.0;
  eval.assign(n, 0.0);
  double norm = 0.0;
  for (double v : a_in) norm += v * v;
  norm = std::sqrt(norm);
  if (norm <= 0.0) return;
  const double skip = 0.01 * kEps * norm;
  std::vector<double> w = a_in;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p + 1 < n; ++p)
      for (int q = p + 1; q < n; ++q) off += at(w, p, q) * at(w, p, q);
    if (std::sqrt(2.0 * off) <= kEps * norm) break;
    for (int p = 0; p + 1 < n; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const double apq = at(w, p, q);
        if (std::fabs(apq) <= skip) continue;
        const double theta = 0.5 * (at(w, q, q) - at(w, p, p)) / apq;
        const double t = std::copysign(1.0, theta) / (std::fabs(theta) + std::sqrt(1.0 + theta * theta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int j = 0; j < n; ++j) {  // columns p, q
          const double x = at(w, j, p), y = at(w, j, q);
          at(w, j, p) = c * x - s * y;
          at(w, j, q) = s * x + c * y;
        }
        for (int j = 0; j < n; ++j) {  // rows p, q
          const double x = at(w, p, j), y = at(w, q, j);
          at(w, p, j) = c * x - s * y;
          at(w, q, j) = s * x + c * y;
        }
        at(w, p, q) = 0.0;
        at(w, q, p) = 0.0;
        for (int j = 0; j < n; ++j) {
          const double x = at(evec, j, p), y = at(evec, j, q);
          at(evec, j, p) = c * x - s * y;
          at(evec, j, q) = s * x + c * y;
        }
      }
    }
  }
  for (int p = 0; p < n; ++p) eval[p] = at(w, p, p);
  for (int p = 0; p + 1 < n; ++p) {
    int best = p;
    for (int k = p + 1; k < n; ++k)
      if (std::fabs(eval[k]) > std::fabs(eval[best])) best = k;
    if (best == p) continue;
    std::swap(eval[p], eval[best]);
    for (int j = 0; j < n; ++j) std::swap(at(evec, j, p), at(evec, j, best));
  }
}

}  // namespace

LowRank make_lowrank(const std::vector<int>& species, int rank, double tolerance) {
  // fourier/decomposition.f90:104-175 new_lowrank_c6: the matrix of all
  // reference C6 over (species, reference) in species order, its eigenpairs
  // by decreasing magnitude, kept until the largest relative error of the
  // reconstruction (entries below 1e-6 of the largest count at that floor)
  // is within the tolerance, or for the requested rank.
  const auto& tab = element_tables();
  LowRank lr;
  lr.species = species;
  std::vector<int> spmap, refmap;
  for (size_t s = 0; s < species.size(); ++s)
    for (int r = 0; r < tab.nref[species[s] - 1]; ++r) {
      spmap.push_back(static_cast<int>(s));
      refmap.push_back(r);
    }
  const int n = static_cast<int>(spmap.size());
  std::vector<double> a(size_t(n) * n);
  double amax = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      const double v = reference_c6(refmap[j], refmap[i], species[spmap[j]], species[spmap[i]]);
      a[size_t(j) + size_t(i) * n] = v;
      amax = std::max(amax, std::fabs(v));
    }
  std::vector<double> eval, evec;
  jacobi(n, a, eval, evec);
  const double floor = 1.0e-6 * amax;
  std::vector<double> res = a;
  int keep = n;
  double err = 0.0;
  for (int l = 0; l < n; ++l) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        res[size_t(j) + size_t(i) * n] -= eval[l] * evec[size_t(j) + size_t(l) * n] * evec[size_t(i) + size_t(l) * n];
    err = 0.0;
    for (size_t k = 0; k < res.size(); ++k) err = std::max(err, std::fabs(res[k]) / std::max(std::fabs(a[k]), floor));
    if (rank > 0) {
      if (l + 1 >= std::min(rank, n)) {
        keep = l + 1;
        break;
      }
    } else if (err <= tolerance) {
      keep = l + 1;
      break;
    }
  }
  lr.rank = keep;
  lr.error = err;
  lr.ndim = n;
  lr.lambda.assign(eval.begin(), eval.begin() + keep);
  lr.vec.assign(species.size() * kMaxRef * keep, 0.0);
  for (int l = 0; l < keep; ++l)
    for (int i = 0; i < n; ++i)
      lr.vec[(size_t(spmap[i]) * kMaxRef + refmap[i]) * keep + l] = evec[size_t(i) + size_t(l) * n];
  return lr;
}

double LowRank::reference_c6(int si, int ri, int sj, int rj) const {
  double v = 0.0;
  for (int l = 0; l < rank; ++l)
    v += lambda[l] * vec[(size_t(si) * kMaxRef + ri) * rank + l] * vec[(size_t(sj) * kMaxRef + rj) * rank + l];
  return v;
}

bool supports_ewald(const SysParam& p) {
  // rational.f90:113-123 (always), zero.f90:113-123 (integer exponent);
  // the other families keep the default of damping.f90:241-251
  if (p.kind == kRational) return true;
  if (p.kind == kZero) return std::fabs(p.alp - std::nearbyint(p.alp)) <= std::sqrt(kEps);
  return false;
}

int fourier_terms(const SysParam& p, int zi, int zj, FourierTerm terms[2]) {
  // get_fourier_terms of rational.f90:127-163 and zero.f90:127-164
  if (std::fabs(p.s6) < kEps && std::fabs(p.s8) < kEps) return 0;
  const double rrij = 3.0 * r4r2_value(zi) * r4r2_value(zj);
  if (p.kind == kRational) {
    const double r0 = p.a1 * std::sqrt(rrij) + p.a2;
    terms[0] = {p.s6, r0, 0, 6};
    terms[1] = {p.s8 * rrij, r0, 0, 8};
  } else {
    const int alp6 = static_cast<int>(std::lround(p.alp)), alp8 = alp6 + 2;
    const double r0 = vdw_radius(zj, zi);
    terms[0] = {p.s6, std::pow(6.0, 1.0 / alp6) * p.a1 * r0, alp6 - 6, alp6};
    terms[1] = {p.s8 * rrij, std::pow(6.0, 1.0 / alp8) * p.a2 * r0, alp8 - 8, alp8};
  }
  return 2;
}

bool fourier_supported(const FourierTerm& t) {
  // kernel.f90:63-78: even powers, m + 3 < alpha, positive radius
  return t.alpha > 0 && t.m >= 0 && t.alpha % 2 == 0 && t.m % 2 == 0 && t.m + 3 < t.alpha && t.radius > 0.0;
}

double fourier_potential_zero(const FourierTerm& t) {
  // kernel.f90:133-147: the term at r = 0
  return t.m > 0 ? 0.0 : t.pref / std::pow(t.radius, t.alpha);
}

double fourier_cutoff(const FourierTerm& t, double tolerance) {
  // kernel.f90:151-164: the slowest exponential of the poles falls below
  // the tolerance
  return -std::log(tolerance) / (t.radius * std::sin(M_PI / t.alpha));
}

ReciprocalTiles reciprocal_tiles(const double lattice[9], double kcut) {
  // ewald.f90:118-121: b_j with a_i . b_j = 2 pi delta_ij (rows of lattice
  // are the a_i) and the points h b1 + k b2 + l b3 within kcut, in the box
  // that get_lattice_points spans (cutoff.f90:190-227).  Of each +-k pair
  // the one with (h, k, l) lexicographically positive is kept.  A line (h, k)
  // meets the sphere in one interval of l, found from the quadratic in l and
  // trimmed with the exact test |k| <= kcut on the vector as the reference
  // forms it; the interval is cut into tiles of kWave consecutive l.
  ReciprocalTiles out;
  const double* a = lattice;
  const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
  out.volume = std::fabs(det);
  if (out.volume <= 0.0) throw Error("Lattice vectors are linearly dependent");
  double* b = out.rec;
  for (int j = 0; j < 3; ++j) {  // b_j = 2 pi (a_k x a_l) / det, (j, k, l) cyclic
    const double* u = a + 3 * ((j + 1) % 3);
    const double* v = a + 3 * ((j + 2) % 3);
    b[3 * j + 0] = 2.0 * M_PI * (u[1] * v[2] - u[2] * v[1]) / det;
    b[3 * j + 1] = 2.0 * M_PI * (u[2] * v[0] - u[0] * v[2]) / det;
    b[3 * j + 2] = 2.0 * M_PI * (u[0] * v[1] - u[1] * v[0]) / det;
  }
  // repetitions per direction as get_translations (cutoff.f90:190-227)
  for (int d = 0; d < 3; ++d) {
    const double* u = b + 3 * ((d + 1) % 3);
    const double* v = b + 3 * ((d + 2) % 3);
    double n[3] = {u[1] * v[2] - v[1] * u[2], u[2] * v[0] - v[2] * u[0], u[0] * v[1] - v[0] * u[1]};
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double cs = (n[0] * b[3 * d] + n[1] * b[3 * d + 1] + n[2] * b[3 * d + 2]) / len;
    out.rep[d] = static_cast<int>(std::ceil(std::fabs(kcut / cs)));
  }
  auto norm = [&](int h, int k, int l) {
    double s = 0.0;
    for (int c = 0; c < 3; ++c) {
      const double x = b[c] * h + b[3 + c] * k + b[6 + c] * l;
      s += x * x;
    }
    return std::sqrt(s);
  };
  const double* b3 = b + 6;
  const double bb = b3[0] * b3[0] + b3[1] * b3[1] + b3[2] * b3[2];
  for (int h = 0; h <= out.rep[0]; ++h) {
    for (int k = h == 0 ? 0 : -out.rep[1]; k <= out.rep[1]; ++k) {
      double c[3];
      for (int q = 0; q < 3; ++q) c[q] = b[q] * h + b[3 + q] * k;
      const double cb = c[0] * b3[0] + c[1] * b3[1] + c[2] * b3[2];
      const double cc = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
      const double disc = cb * cb - bb * (cc - kcut * kcut);
      if (disc < 0.0) continue;
      const double sq = std::sqrt(disc);
      int lo = static_cast<int>(std::floor((-cb - sq) / bb)) - 1;
      int hi = static_cast<int>(std::ceil((-cb + sq) / bb)) + 1;
      lo = std::max(lo, h == 0 && k == 0 ? 0 : -out.rep[2]);
      hi = std::min(hi, out.rep[2]);
      while (lo <= hi && norm(h, k, lo) > kcut) ++lo;
      while (hi >= lo && norm(h, k, hi) > kcut) --hi;
      for (int l = lo; l <= hi; l += kWave) {
        const int cnt = std::min(kWave, hi - l + 1);
        out.tiles.push_back(make_int4(h, k, l, cnt));
        out.vectors += cnt;
      }
    }
  }
  return out;
}

SpmeMesh spme_mesh(const double lattice[9], double kcut, int mesh_req) {
  // spme.f90:55-74 get_spme_mesh / damping.f90:368-375: mesh_req > 0 rounded
  // up to a power of two (fft.f90:48-60), else enough points to resolve
  // 2 pi / kcut along each lattice vector; spme.f90:503-529 Euler factors.
  SpmeMesh out;
  auto valid = [](long long n) {
    int m = 1;
    while (m < n) m *= 2;
    return m;
  };
  ReciprocalTiles rt = reciprocal_tiles(lattice, 0.0);  // basis and volume only
  out.volume = rt.volume;
  std::memcpy(out.rec, rt.rec, sizeof out.rec);
  for (int d = 0; d < 3; ++d) {
    const double* a = lattice + 3 * d;
    const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    out.m[d] = mesh_req > 0 ? valid(mesh_req) : valid(static_cast<long long>(std::ceil(kcut * len / M_PI)));
  }
  double spl[6] = {0}, dspl[6];
  bspline6_host(0.0, spl, dspl);
  out.mmax = std::max(out.m[0], std::max(out.m[1], out.m[2]));
  out.bfac.assign(3 * size_t(out.mmax), 0.0);
  for (int d = 0; d < 3; ++d)
    for (int im = 0; im < out.m[d]; ++im) {
      double re = 0.0, im_ = 0.0;
      for (int jm = 0; jm < 5; ++jm) {
        const double arg = 2.0 * M_PI * im * jm / out.m[d];
        re += spl[jm] * std::cos(arg);
        im_ += spl[jm] * std::sin(arg);
      }
      out.bfac[size_t(d) * out.mmax + im] = 1.0 / (re * re + im_ * im_);
    }
  return out;
}

void bspline6_host(double w, double* th, double* dth) {
  // spme.f90:457-499 fill_bspline, order 6
  const int n = 6;
  for (int k = 0; k < n; ++k) th[k] = 0.0;
  th[0] = 1.0 - w;
  th[1] = w;
  for (int korder = 3; korder <= n - 1; ++korder) {
    const double div = 1.0 / (korder - 1);
    th[korder - 1] = div * w * th[korder - 2];
    for (int j = 1; j <= korder - 2; ++j)
      th[korder - j - 1] = div * ((w + j) * th[korder - j - 2] + (korder - j - w) * th[korder - j - 1]);
    th[0] = div * (1.0 - w) * th[0];
  }
  dth[0] = -th[0];
  for (int j = 2; j <= n; ++j) dth[j - 1] = th[j - 2] - th[j - 1];
  const double div = 1.0 / (n - 1);
  th[n - 1] = div * w * th[n - 2];
  for (int j = 1; j <= n - 2; ++j) th[n - j - 1] = div * ((w + j) * th[n - j - 2] + (n - j - w) * th[n - j - 1]);
  th[0] = div * (1.0 - w) * th[0];
}

}  // namespace d3
