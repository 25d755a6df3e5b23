// Element tables of DFT-D3, embedded from data/d3_tables.npz at build time
// (d3_tables_data.inc is written by build.py into the build directory).
#include <cmath>

#include "d3_common.hpp"

namespace {
#include "d3_tables_data.inc"
}  // namespace

namespace d3 {

const ElementTables& element_tables() {
  static const ElementTables tables = [] {
    ElementTables t{};
    for (int z = 0; z < kMaxElem; ++z) {
      t.nref[z] = kTabNref[z];
      for (int r = 0; r < kMaxRef; ++r) t.refcn[z][r] = kTabRefcn[z * kMaxRef + r];
    }
    t.c6ref = kTabC6ref;
    for (int z = 0; z < 118; ++z) {
      t.r4_over_r2[z] = kTabR4r2[z];
      t.covrad_aa[z] = kTabCovrad[z];
    }
    for (int p = 0; p < kNumPairs; ++p) t.vdwrad_aa[p] = kTabVdwrad[p];
    return t;
  }();
  return tables;
}

double aatoau() {
  // hbar / (m_e c alpha) with CODATA 2018 constants, in Angstrom.
  static const double value = [] {
    const double pi = 3.14159265358979323846;
    const double hbar = 6.62607015e-34 / (2.0 * pi);
    const double bohr = hbar / (9.1093837015e-31 * 299792458.0 * 7.2973525693e-3);
    return 1.0 / (bohr * 1.0e10);
  }();
  return value;
}

// mctc-lib covalent_rad_d3 = 4/3 * covalent_rad_2009 (used at model.f90:125).
double covalent_radius(int z) {
  if (z < 1 || z > 118) return 0.0;
  return 4.0 / 3.0 * (aatoau() * element_tables().covrad_aa[z - 1]);
}

// data/r4r2.f90:64-66,91-105
double r4r2_value(int z) {
  if (z < 1 || z > 118) return 0.0;
  return std::sqrt(0.5 * (element_tables().r4_over_r2[z - 1] * std::sqrt(double(z))));
}

// data/vdwrad.f90:1163-1197; elements beyond the table fall back to the
// covalent-radius estimate with cov2vdw = 1.45.
double vdw_radius(int z1, int z2) {
  const auto& t = element_tables();
  const double cov2vdw = 1.45;
  auto tab = [&](int a, int b) {
    if (a < b) std::swap(a, b);
    return aatoau() * t.vdwrad_aa[b + a * (a - 1) / 2 - 1];
  };
  auto cov_aa = [&](int z) { return covalent_radius(z); };
  bool in1 = z1 > 0 && z1 <= kMaxElem, in2 = z2 > 0 && z2 <= kMaxElem;
  if (in1 && in2) return tab(z1, z2);
  if (in1) return 0.5 * tab(z1, z1) + cov2vdw * cov_aa(z2);
  if (in2) return 0.5 * tab(z2, z2) + cov2vdw * cov_aa(z1);
  return cov2vdw * (cov_aa(z1) + cov_aa(z2));
}

// reference.f90:160-171 get_c6 with 0-based reference indices.
double reference_c6(int iref, int jref, int zi, int zj) {
  const auto& t = element_tables();
  if (zi > zj) {
    int ic = zj + zi * (zi - 1) / 2 - 1;
    return t.c6ref[(static_cast<size_t>(ic) * kMaxRef + jref) * kMaxRef + iref];
  }
  int ic = zi + zj * (zj - 1) / 2 - 1;
  return t.c6ref[(static_cast<size_t>(ic) * kMaxRef + iref) * kMaxRef + jref];
}

}  // namespace d3
