// Host side of the EOM-CCSD steps: the packed vector space the Davidson driver
// works in and its map to the dense r1 / r2 arrays of the device right-applies,
// the preconditioner diagonal, and the reordering of the reference's integral
// classes into the engine's.  No device code.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

namespace sisi_host {

// IP: r1[i] (No), r2[i,j,a] (No x No x Nv, antisymmetric in i, j), packed as
// r1 then (pairs i < j, j slowest) x a.  EA: r1[a] (Nv), r2[j,a,b] (No x Nv x
// Nv, antisymmetric in a, b), packed as r1 then (pairs a < b, b slowest) x j.
// Dense arrays are first index fastest, as the C ABI takes them.
struct EomSpace {
  bool ea;
  int64_t o, v;
  int64_t n1, nk;  // length of r1; range of the index that is not paired
  std::vector<std::pair<int64_t, int64_t>> pairs;
  size_t n;

  EomSpace(bool ea_, int64_t o_, int64_t v_) : ea(ea_), o(o_), v(v_), n1(ea_ ? v_ : o_), nk(ea_ ? o_ : v_) {
    const int64_t np = ea ? v : o;
    for (int64_t y = 0; y < np; ++y)
      for (int64_t x = 0; x < y; ++x) pairs.emplace_back(x, y);
    n = size_t(n1) + pairs.size() * size_t(nk);
  }
  size_t r1_size() const { return size_t(n1); }
  size_t r2_size() const { return size_t(o) * size_t(v) * size_t(ea ? v : o); }
  size_t packed(size_t p, int64_t k) const { return size_t(n1) + p * size_t(nk) + size_t(k); }
  // dense offset of r2 with paired indices (x, y) and unpaired k
  size_t dense(int64_t x, int64_t y, int64_t k) const {
    return ea ? size_t(k + o * (x + v * y)) : size_t(x + o * (y + o * k));
  }
  void unpack(const std::vector<double> &xv, std::vector<double> &r1, std::vector<double> &r2) const {
    r1.assign(xv.begin(), xv.begin() + n1);
    r2.assign(r2_size(), 0.0);
    for (size_t p = 0; p < pairs.size(); ++p)
      for (int64_t k = 0; k < nk; ++k) {
        const double val = xv[packed(p, k)];
        r2[dense(pairs[p].first, pairs[p].second, k)] = val;
        r2[dense(pairs[p].second, pairs[p].first, k)] = -val;
      }
  }
  std::vector<double> pack(const std::vector<double> &s1, const std::vector<double> &s2) const {
    std::vector<double> y(n);
    for (int64_t i = 0; i < n1; ++i) y[size_t(i)] = s1[size_t(i)];
    for (size_t p = 0; p < pairs.size(); ++p)
      for (int64_t k = 0; k < nk; ++k) y[packed(p, k)] = s2[dense(pairs[p].first, pairs[p].second, k)];
    return y;
  }
  // IPCcsdPreconditioner / EACcsdPreconditioner diagonals
  // (CcsdPreconditioner.cxx:546-560, 374-397): IP -f_ii and
  // -f_ii - f_jj + f_aa, EA f_aa and -f_jj + f_aa + f_bb
  std::vector<double> diagonal(const double *ei, const double *ea_) const {
    std::vector<double> d(n);
    for (int64_t i = 0; i < n1; ++i) d[size_t(i)] = ea ? ea_[i] : -ei[i];
    for (size_t p = 0; p < pairs.size(); ++p)
      for (int64_t k = 0; k < nk; ++k) {
        const int64_t x = pairs[p].first, y = pairs[p].second;
        d[packed(p, k)] = ea ? -ei[k] + ea_[x] + ea_[y] : -ei[x] - ei[y] + ea_[k];
      }
    return d;
  }
};

// The reference's classes as the engine's, for real antisymmetrized
// integrals: V_abij = V_ijab (PPHH from HHPP), V_abci = -V_icab (PPPH from
// HPPP), V_aibj = V_iajb (PHPH from HPHP).
inline std::vector<double> pphh_from_hhpp(const double *hhpp, int64_t o, int64_t v) {
  std::vector<double> out(static_cast<size_t>(v * v * o * o));
  for (int64_t j = 0; j < o; ++j)
    for (int64_t i = 0; i < o; ++i)
      for (int64_t b = 0; b < v; ++b)
        for (int64_t a = 0; a < v; ++a)
          out[size_t(a + v * (b + v * (i + o * j)))] = hhpp[i + o * (j + o * (a + v * b))];
  return out;
}

inline std::vector<double> ppph_from_hppp(const double *hppp, int64_t o, int64_t v) {
  std::vector<double> out(static_cast<size_t>(v * v * v * o));
  for (int64_t i = 0; i < o; ++i)
    for (int64_t c = 0; c < v; ++c)
      for (int64_t b = 0; b < v; ++b)
        for (int64_t a = 0; a < v; ++a)
          out[size_t(a + v * (b + v * (c + v * i)))] = -hppp[i + o * (c + v * (a + v * b))];
  return out;
}

inline std::vector<double> phph_from_hphp(const double *hphp, int64_t o, int64_t v) {
  std::vector<double> out(static_cast<size_t>(v * o * v * o));
  for (int64_t j = 0; j < o; ++j)
    for (int64_t b = 0; b < v; ++b)
      for (int64_t i = 0; i < o; ++i)
        for (int64_t a = 0; a < v; ++a)
          out[size_t(a + v * (i + o * (b + v * j)))] = hphp[i + o * (a + v * (j + o * b))];
  return out;
}

}  // namespace sisi_host
