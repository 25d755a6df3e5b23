// The three-body Hessian term of one triple (atm.f90:38-429
// get_atm_dispersion_hessian), shared by the device kernel (hess3_kernel,
// one wavefront per (atom i, slot a), lanes over the slots b) and a host
// build of the same code for checks on the CPU.  Plain C++ apart from the
// host/device qualifiers; the accumulation goes through a sink:
//   hf(x, y, k, v)      Hessian block between triple positions x, y in {0, 1}
//                       (atom i, atom of slot a), entry k = 3 c + d
//   h(row, col, v)      Hessian entry [row][col] of the system (3 atom + c)
//   cf(x, y, v)         dEdcndcn between positions x, y in {0, 1}
//   c(atom, atom2, v)   dEdcndcn entry of the system
//   df(x, c, y, v)      dEdcndr [3 atom(x) + c][atom(y)] for x, y in {0, 1}
//   d(row, atom, v)     dEdcndr entry of the system
//   ef(x, v)            dEdcn of position x in {0, 1}
//   e(atom, v)          dEdcn of an atom of the system
// (positions 0 and 1 are the same for every triple of a wavefront, so the
// device sink sums them in registers; atoms are 0-based in the system).
#pragma once

#include <cmath>

#include "d3_common.hpp"

#if defined(__HIPCC__)
#define D3_HD __host__ __device__ __forceinline__
#else
#define D3_HD inline
#endif

namespace d3 {

constexpr double kEpsilon = 2.220446049250313e-16;  // epsilon(1.0_wp)
constexpr double kRs9Atm = 4.0 / 3.0;               // rational.f90:73, zero.f90:73

// vjk = vik - vij (atm.f90:723) and its squared length, as the reference.
D3_HD double diff_norm2(double bx, double by, double bz, double ax, double ay, double az, double& v0, double& v1,
                        double& v2) {
#pragma clang fp contract(off)
  v0 = bx - ax;
  v1 = by - ay;
  v2 = bz - az;
  return v0 * v0 + v1 * v1 + v2 * v2;
}

// partition.f90:127-148 owns_pair with 0-based atom indices il >= jl: the
// pair's index in the lower triangle, il (il + 1) / 2 + jl, modulo nparts.
D3_HD bool owns_pair(const SysParam& p, int il, int jl) {
  if (p.nparts <= 1) return true;
  const long long idx = static_cast<long long>(il) * (il + 1) / 2 + jl;
  return idx % p.nparts == p.part;
}

// cutoff.f90:73-99 smooth_cutoff as a function of u = r^2: sw and its first
// and second derivatives in u.
D3_HD void switch_r2(double r2, double cutoff, double width, double& sw, double& dsw, double& d2sw) {
  const double r1 = sqrt(r2);
  double dswdr = 0.0, d2swdr2 = 0.0;
  sw = 1.0;
  if (width > 0.0 && width < cutoff && r1 > cutoff - width) {
    if (r1 >= cutoff) {
      sw = 0.0;
    } else {
      const double x = (cutoff - r1) / width;
      sw = x * x * x * (10.0 + x * (-15.0 + 6.0 * x));
      dswdr = -30.0 * x * x * (1.0 - x) * (1.0 - x) / width;
      d2swdr2 = 60.0 * x * (1.0 - x) * (1.0 - 2.0 * x) / (width * width);
    }
  }
  dsw = 0.5 * dswdr / r1;
  d2sw = 0.25 * (d2swdr2 - dswdr / r1) / r2;
}

// What the triples of one system share: parameters, the n x n C6 matrices
// of the system ([row][col] = C6 of the atom pair (row, col), dc6 the
// derivative by the CN of the row atom, d2c6 the second, d2c6ij the mixed
// one) and the species-pair van der Waals radii.
struct Hess3System {
  SysParam p;
  int n;
  const double* c6;
  const double* dc6;
  const double* d2c6;
  const double* d2c6ij;
  const double* rvdw;  // [nsp][nsp]
  int nsp;
};

// One triple: atom i (position 0, atom at0, species sp0), slot a (position 1:
// atom at1, vector va = x_a - x_i with |va|^2 = ua) and slot b (position 2).
// selfa / selfb: the slot is an image of atom i (the weights of atm_kernel:
// 1/2 per such slot, 4/3 more when both are).
template <class Sink>
D3_HD void hess3_triple(const Hess3System& s, int at0, int at1, int at2, int sp0, int sp1, int sp2, bool selfa,
                        bool selfb, const double va[3], double ua, const double vb[3], double ub, Sink& sink) {
  const SysParam& p = s.p;
  const int n = s.n;
  const double cut2 = p.cut3 * p.cut3, alp3 = (p.alp + 2.0) / 3.0, aexp = 0.5 * alp3;
  const int sigp[3][3] = {{-1, -1, 0}, {1, 0, -1}, {0, 1, 1}};   // du_pair/dR_atom = 2 sigp[atom][pair] v_pair
  const double cl[3][3] = {{1, -1, 1}, {1, 1, -1}, {-1, 1, 1}};  // coefficient of u_pair in l_l
  double vec[3][3];
  vec[0][0] = va[0]; vec[0][1] = va[1]; vec[0][2] = va[2];
  vec[1][0] = vb[0]; vec[1][1] = vb[1]; vec[1][2] = vb[2];
  double u[3];
  u[2] = diff_norm2(vb[0], vb[1], vb[2], va[0], va[1], va[2], vec[2][0], vec[2][1], vec[2][2]);
  if (u[2] > cut2 || u[2] < kEpsilon) return;
  u[0] = ua;
  u[1] = ub;
  const int at[3] = {at0, at1, at2};
  const double c6t[3] = {s.c6[static_cast<long long>(at1) * n + at0], s.c6[static_cast<long long>(at2) * n + at0],
                         s.c6[static_cast<long long>(at2) * n + at1]};
  if (fabs(c6t[0]) < kEpsilon || fabs(c6t[1]) < kEpsilon || fabs(c6t[2]) < kEpsilon) return;
  // a triple of another part of a work partition (atm.f90:136) gets weight
  // 0: a zero contribution rather than a branch out of the caller's loop
  const double w = owns_pair(p, at0, at1 > at2 ? at1 : at2)
                       ? (selfa ? 0.5 : 1.0) * (selfb ? 0.5 : 1.0) * (selfa && selfb ? 4.0 / 3.0 : 1.0)
                       : 0.0;
  double swp[3], dswp[3], d2swp[3];
  for (int q = 0; q < 3; ++q) switch_r2(u[q], p.cut3, p.width3, swp[q], dswp[q], d2swp[q]);
  const double sval = swp[0] * swp[1] * swp[2];
  const double sp[3] = {dswp[0] * swp[1] * swp[2], swp[0] * dswp[1] * swp[2], swp[0] * swp[1] * dswp[2]};
  double spq[3][3];
  spq[0][0] = d2swp[0] * swp[1] * swp[2];
  spq[1][1] = swp[0] * d2swp[1] * swp[2];
  spq[2][2] = swp[0] * swp[1] * d2swp[2];
  spq[0][1] = spq[1][0] = dswp[0] * dswp[1] * swp[2];
  spq[0][2] = spq[2][0] = dswp[0] * swp[1] * dswp[2];
  spq[1][2] = spq[2][1] = swp[0] * dswp[1] * dswp[2];
  const double ww = u[0] * u[1] * u[2];
  const double wp[3] = {u[1] * u[2], u[0] * u[2], u[0] * u[1]};
  const double wpq[3][3] = {{0.0, u[2], u[1]}, {u[2], 0.0, u[0]}, {u[1], u[0], 0.0}};
  const double lv[3] = {u[0] + u[2] - u[1], u[0] - u[2] + u[1], -u[0] + u[2] + u[1]};
  const double nval = lv[0] * lv[1] * lv[2];
  double np_[3], npq[3][3];
  for (int q = 0; q < 3; ++q) np_[q] = cl[0][q] * lv[1] * lv[2] + cl[1][q] * lv[0] * lv[2] + cl[2][q] * lv[0] * lv[1];
  for (int q = 0; q < 3; ++q)
    for (int r = 0; r < 3; ++r) {
      double tmp = 0.0;
      for (int l = 0; l < 3; ++l)
        for (int m = 0; m < 3; ++m)
          if (l != m) tmp += cl[l][q] * cl[m][r] * lv[3 - l - m];
      npq[q][r] = tmp;
    }
  const double gv = pow(ww, -2.5), gd = -2.5 * pow(ww, -3.5), gdd = 8.75 * pow(ww, -4.5);
  const double hv = pow(ww, -1.5), hd = -1.5 * pow(ww, -2.5), hdd = 3.75 * pow(ww, -3.5);
  const double ang = 0.375 * nval * gv + hv;
  const double r0 = kRs9Atm * s.rvdw[sp1 * s.nsp + sp0] * kRs9Atm * s.rvdw[sp2 * s.nsp + sp0] * kRs9Atm *
                    s.rvdw[sp2 * s.nsp + sp1];
  const double tval = 6.0 * pow(r0, alp3) * pow(ww, -aexp);
  const double fd = 1.0 / (1.0 + tval);
  double angp[3], tp[3], fdp[3], ap[3], kp[3];
  for (int q = 0; q < 3; ++q) {
    angp[q] = 0.375 * (np_[q] * gv + nval * gd * wp[q]) + hd * wp[q];
    tp[q] = -aexp * tval * wp[q] / ww;
    fdp[q] = -tp[q] * fd * fd;
  }
  const double av = ang * fd;
  for (int q = 0; q < 3; ++q) ap[q] = angp[q] * fd + ang * fdp[q];
  const double kv = sval * av;
  for (int q = 0; q < 3; ++q) kp[q] = sp[q] * av + sval * ap[q];
  double kpq[3][3];
  for (int q = 0; q < 3; ++q)
    for (int r = 0; r < 3; ++r) {
      const double angpq = 0.375 * (npq[q][r] * gv + np_[q] * gd * wp[r] + np_[r] * gd * wp[q] +
                                    nval * (gdd * wp[q] * wp[r] + gd * wpq[q][r])) +
                           hdd * wp[q] * wp[r] + hd * wpq[q][r];
      const double tpq = aexp * (aexp + 1.0) * tval * wp[q] * wp[r] / (ww * ww) - aexp * tval * wpq[q][r] / ww;
      const double fdpq = -tpq * fd * fd + 2.0 * tp[q] * tp[r] * fd * fd * fd;
      const double apq = angpq * fd + angp[q] * fdp[r] + angp[r] * fdp[q] + ang * fdpq;
      kpq[q][r] = spq[q][r] * av + sp[q] * ap[r] + sp[r] * ap[q] + sval * apq;
    }
  const double qv = sqrt(fabs(c6t[0] * c6t[1] * c6t[2]));
  const double pref = p.s9 * w, pq = pref * qv;
  // Cartesian second derivatives at fixed CN
  double gk[3][3][3];
  for (int a = 0; a < 3; ++a)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double tmp = 0.0;
        for (int q = 0; q < 3; ++q) tmp += kpq[q][r] * 2.0 * sigp[a][q] * vec[q][c];
        gk[a][r][c] = pq * tmp;
      }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) {
      double dg = 0.0;
      for (int q = 0; q < 3; ++q) dg += kp[q] * sigp[a][q] * sigp[bb][q];
      dg *= 2.0 * pq;
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          double tmp = c == d ? dg : 0.0;
          for (int r = 0; r < 3; ++r) tmp += gk[a][r][c] * 2.0 * sigp[bb][r] * vec[r][d];
          if (a < 2 && bb < 2) sink.hf(a, bb, 3 * c + d, tmp);
          else sink.h(3 * at[a] + c, 3 * at[bb] + d, tmp);
        }
    }
  // derivatives through C6: pairs (i, a), (i, b), (a, b)
  double ec[3], euc[3][3], ecc[3][3];
  for (int q = 0; q < 3; ++q) {
    const double qm = 0.5 * qv / c6t[q];
    ec[q] = pref * qm * kv;
    for (int r = 0; r < 3; ++r) {
      const double qmn = q == r ? -0.25 * qv / (c6t[q] * c6t[q]) : 0.25 * qv / (c6t[q] * c6t[r]);
      ecc[q][r] = pref * qmn * kv;
      euc[r][q] = pref * qm * kp[r];
    }
  }
  // the (position, pair) entries of dC6/dCN: both atoms of each pair, or
  // one entry of twice the derivative when the pair's two atoms coincide
  // (the second then has coefficient 0)
  constexpr int kEntPos[6] = {0, 1, 0, 2, 1, 2}, kEntPair[6] = {0, 0, 1, 1, 2, 2};
  double ent_coef[6];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int a0 = at[kEntPos[2 * q]], a1 = at[kEntPos[2 * q + 1]];
    if (a0 != a1) {
      ent_coef[2 * q] = s.dc6[static_cast<long long>(a0) * n + a1];
      ent_coef[2 * q + 1] = s.dc6[static_cast<long long>(a1) * n + a0];
    } else {
      ent_coef[2 * q] = 2.0 * s.dc6[static_cast<long long>(a0) * n + a0];
      ent_coef[2 * q + 1] = 0.0;
    }
  }
#pragma unroll
  for (int ie = 0; ie < 6; ++ie) {
    const int kpos = kEntPos[ie], kq = kEntPair[ie], ka = at[kpos];
    const double ce = ent_coef[ie];
    if (ce == 0.0) continue;
    if (kpos < 2) sink.ef(kpos, ec[kq] * ce);
    else sink.e(ka, ec[kq] * ce);
#pragma unroll
    for (int jf = 0; jf < 6; ++jf) {
      const int jpos = kEntPos[jf];
      const double v = ecc[kq][kEntPair[jf]] * ce * ent_coef[jf];
      if (kpos < 2 && jpos < 2) sink.cf(kpos, jpos, v);
      else if (ent_coef[jf] != 0.0) sink.c(ka, at[jpos], v);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double egr = 0.0;
        for (int q = 0; q < 3; ++q) egr += euc[q][kq] * 2.0 * sigp[a][q] * vec[q][c];
        if (a < 2 && kpos < 2) sink.df(a, c, kpos, egr * ce);
        else sink.d(3 * at[a] + c, ka, egr * ce);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int p0 = kEntPos[2 * q], p1 = kEntPos[2 * q + 1];  // positions of the pair's atoms
    const int a0 = at[p0], a1 = at[p1];
    const long long o01 = static_cast<long long>(a0) * n + a1;
    if (a0 != a1) {
      const long long o10 = static_cast<long long>(a1) * n + a0;
      const double v00 = ec[q] * s.d2c6[o01], v11 = ec[q] * s.d2c6[o10], vx = ec[q] * s.d2c6ij[o01];
      if (p1 < 2) {  // pair (i, a): positions 0 and 1
        sink.cf(0, 0, v00);
        sink.cf(1, 1, v11);
        sink.cf(0, 1, vx);
        sink.cf(1, 0, vx);
      } else {
        sink.c(a0, a0, v00);
        sink.c(a1, a1, v11);
        sink.c(a0, a1, vx);
        sink.c(a1, a0, vx);
      }
    } else {
      const long long o00 = static_cast<long long>(a0) * n + a0;
      const double v = ec[q] * (2.0 * s.d2c6[o00] + 2.0 * s.d2c6ij[o00]);
      if (p1 < 2) sink.cf(0, 0, v);
      else sink.c(a0, a0, v);
    }
  }
}

}  // namespace d3
