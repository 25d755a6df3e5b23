// Host build of the per-triple three-body Hessian code (d3_hess3.hpp), the
// code hess3_kernel runs per lane, driven the way the kernel drives it: the
// three-body list L(i) of every atom i (slots (j <= i, T) with 0 < |v|^2 <=
// cut3^2, atm.f90:136-150), one (i, slot a) at a time with the entries of
// positions 0 and 1 summed in locals and flushed after the slots b > a.
//
// Built with sanitizers (address + undefined, or memory) by the tests; reads
// one system from a binary file and writes the four accumulators:
//   in : int32 n, nsp, ntr, part, nparts
//        float64 s9, alp, cut3, width3
//        int32 sp[n]; float64 rvdw[nsp * nsp], xyz[3 n], trans[3 ntr],
//        c6[n n], dc6[n n], d2c6[n n], d2c6ij[n n] (row-major, derivatives by
//        the CN of the row atom)
//   out: float64 hess[3n][3n], dedcn[n], dEdcndr[3n][n], dEdcndcn[n][n]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "d3_hess3.hpp"

namespace {

struct HostSink {
  double hf_[2][2][9], df_[2][3][2], cf_[2][2], ef_[2];
  double* hess;
  double* dedcn;
  double* dEdcndr;
  double* dEdcndcn;
  int n, ndim;
  void clear() {
    for (int x = 0; x < 2; ++x) {
      for (int y = 0; y < 2; ++y)
        for (int c = 0; c < 9; ++c) hf_[x][y][c] = 0.0;
      for (int c = 0; c < 3; ++c) df_[x][c][0] = df_[x][c][1] = 0.0;
      cf_[x][0] = cf_[x][1] = 0.0;
      ef_[x] = 0.0;
    }
  }
  void hf(int x, int y, int k, double v) { hf_[x][y][k] += v; }
  void h(int row, int col, double v) { hess[static_cast<long long>(row) * ndim + col] += v; }
  void cf(int x, int y, double v) { cf_[x][y] += v; }
  void c(int r, int col, double v) { dEdcndcn[static_cast<long long>(r) * n + col] += v; }
  void df(int x, int c, int y, double v) { df_[x][c][y] += v; }
  void d(int row, int atom, double v) { dEdcndr[static_cast<long long>(row) * n + atom] += v; }
  void ef(int x, double v) { ef_[x] += v; }
  void e(int atom, double v) { dedcn[atom] += v; }
  void flush(int at0, int at1) {
    const int fat[2] = {at0, at1};
    for (int x = 0; x < 2; ++x) {
      for (int y = 0; y < 2; ++y) {
        for (int c = 0; c < 9; ++c) h(3 * fat[x] + c / 3, 3 * fat[y] + c % 3, hf_[x][y][c]);
        c(fat[x], fat[y], cf_[x][y]);
        for (int c = 0; c < 3; ++c) d(3 * fat[x] + c, fat[y], df_[x][c][y]);
      }
      e(fat[x], ef_[x]);
    }
  }
};

struct Slot {
  int atom;
  double v[3];
  double r2;
};

template <class T>
bool read_n(FILE* f, T* p, size_t n) {
  return fread(p, sizeof(T), n, f) == n;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s input output\n", argv[0]);
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[5];
  double par[4];
  if (!read_n(f, hdr, 5) || !read_n(f, par, 4)) return 2;
  const int n = hdr[0], nsp = hdr[1], ntr = hdr[2];
  const size_t nn = static_cast<size_t>(n) * n;
  std::vector<int> sp(n);
  std::vector<double> rvdw(static_cast<size_t>(nsp) * nsp), xyz(3 * n), trans(3 * ntr), c6(nn), dc6(nn), d2c6(nn),
      d2c6ij(nn);
  if (!read_n(f, sp.data(), n) || !read_n(f, rvdw.data(), rvdw.size()) || !read_n(f, xyz.data(), xyz.size()) ||
      !read_n(f, trans.data(), trans.size()) || !read_n(f, c6.data(), nn) || !read_n(f, dc6.data(), nn) ||
      !read_n(f, d2c6.data(), nn) || !read_n(f, d2c6ij.data(), nn))
    return 2;
  fclose(f);

  d3::SysParam p{};
  p.s9 = par[0];
  p.alp = par[1];
  p.cut3 = par[2];
  p.width3 = par[3];
  p.part = hdr[3];
  p.nparts = hdr[4];
  const d3::Hess3System sys{p, n, c6.data(), dc6.data(), d2c6.data(), d2c6ij.data(), rvdw.data(), nsp};

  const int ndim = 3 * n;
  std::vector<double> hess(static_cast<size_t>(ndim) * ndim, 0.0), dedcn(n, 0.0), dEdcndr(3 * nn, 0.0),
      dEdcndcn(nn, 0.0);
  HostSink sink;
  sink.hess = hess.data();
  sink.dedcn = dedcn.data();
  sink.dEdcndr = dEdcndr.data();
  sink.dEdcndcn = dEdcndcn.data();
  sink.n = n;
  sink.ndim = ndim;

  const double cut2 = p.cut3 * p.cut3;
  long long triples = 0;
  for (int i = 0; i < n; ++i) {
    std::vector<Slot> list;
    for (int j = 0; j <= i; ++j)
      for (int t = 0; t < ntr; ++t) {
        Slot s;
        s.atom = j;
        for (int c = 0; c < 3; ++c) s.v[c] = xyz[3 * j + c] + trans[3 * t + c] - xyz[3 * i + c];
        s.r2 = s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2];
        if (s.r2 > cut2 || s.r2 < d3::kEpsilon) continue;
        list.push_back(s);
      }
    for (size_t a = 0; a < list.size(); ++a) {
      sink.clear();
      for (size_t b = a + 1; b < list.size(); ++b) {
        d3::hess3_triple(sys, i, list[a].atom, list[b].atom, sp[i], sp[list[a].atom], sp[list[b].atom],
                         list[a].atom == i, list[b].atom == i, list[a].v, list[a].r2, list[b].v, list[b].r2, sink);
        ++triples;
      }
      sink.flush(i, list[a].atom);
    }
  }

  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  fwrite(hess.data(), sizeof(double), hess.size(), o);
  fwrite(dedcn.data(), sizeof(double), dedcn.size(), o);
  fwrite(dEdcndr.data(), sizeof(double), dEdcndr.size(), o);
  fwrite(dEdcndcn.data(), sizeof(double), dEdcndcn.size(), o);
  fclose(o);
  printf("triples %lld\n", triples);
  return 0;
}
