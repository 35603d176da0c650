// Uniform electron gas vertex in real orbitals (host, no device needed).
//
// Follows UegVertexGenerator.cxx:49-229 for the orbital grid, the volume
// v = rs^3 4pi/3 2No, the Madelung constant (Ewald sum, cxx:10-30), the
// Hartree-Fock eigenenergies eps_p = |k_p|^2/2 - sum_o w(k_p - k_o) and the
// plane-wave vertex sqrt(w(G)), w(G) = 4pi/(v|G|^2) (Madelung constant at
// G = 0), for the transfer G = g_r - g_p of each orbital pair.
//
// The real CCSD path needs Gamma[G,p,q] == Gamma[G,q,p], which plane waves
// do not satisfy, so the vertex is written in real orbitals: each plane-wave
// pair (k, -k) becomes (cos, sin) = ((|k> + |-k>)/sqrt2, (|k> - |-k>)/(i sqrt2)),
// and each transfer pair (G, -G) becomes the auxiliary pair
// (sqrt2 Re M_G, sqrt2 Im M_G), M_G being the plane-wave vertex of G in the
// real orbitals (M_-G = conj M_G).  The rotation mixes only degenerate
// orbitals inside the hole or the particle block, so correlation energies are
// those of the plane-wave vertex.
//
// Deterministic orders: orbitals by (|g|^2, g0, g1, g2) (the reference sorts
// by |g|^2 alone, leaving shells in unspecified order); transfers in the
// reference's momMap order (g0 slowest), all |t|^2 <= max|g_r - g_p|^2.  The
// reference bounds the components by those of one longest transfer, which
// can leave some transfers out of its map; here every one is present.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/sisi_engine.h"

namespace {

using ivec = std::array<int, 3>;

int64_t sl(const ivec &a) { return int64_t(a[0]) * a[0] + int64_t(a[1]) * a[1] + int64_t(a[2]) * a[2]; }

// UegVertexGenerator::evalMadelung (cxx:10-30)
double madelung_of(double v) {
  const double pi = 3.14159265358979323846;
  const double kappa = std::pow(v, -1.0 / 3.0);
  const double term2 = pi / (kappa * kappa * v);
  const double term4 = 2.0 * kappa / std::sqrt(pi);
  const double box = 1.0 / kappa;
  double recip = 0.0, real = 0.0;
  for (int l1 = -6; l1 <= 6; ++l1)
    for (int l2 = -6; l2 <= 6; ++l2)
      for (int l3 = -6; l3 <= 6; ++l3) {
        const int n2 = l1 * l1 + l2 * l2 + l3 * l3;
        if (n2 == 0) continue;
        const double modr = box * std::sqrt(double(n2));
        const double k2 = kappa * kappa * n2;
        recip -= 1.0 / (pi * k2) * std::exp(-pi * pi * k2 / kappa / kappa) / v;
        real -= std::erfc(kappa * modr) / modr;
      }
  return real + term2 + term4 + recip;
}

struct Ueg {
  int No, Nv, Np;
  std::vector<ivec> grid;
  std::vector<double> eps;
  std::vector<ivec> transfers;
  std::map<ivec, int> transfer_index;
  std::vector<int> row;      // first auxiliary row of a transfer, -1 for the inverse half
  int NG = 0;
  double v = 0, b = 0, mad = 0;
  std::string error;
};

bool setup(Ueg &u, int No, int Nv, double rs, int hartree_fock, double madelung) {
  const double pi = 3.14159265358979323846;
  u.No = No;
  u.Nv = Nv;
  u.Np = No + Nv;
  const int maxg = int(std::pow(5.0 * u.Np, 1.0 / 3.0));   // cxx:69
  std::vector<ivec> g;
  for (int a = -maxg; a <= maxg; ++a)
    for (int b = -maxg; b <= maxg; ++b)
      for (int c = -maxg; c <= maxg; ++c) g.push_back({a, b, c});
  std::sort(g.begin(), g.end(), [](const ivec &x, const ivec &y) {
    const int64_t a = sl(x), b = sl(y);
    return a != b ? a < b : x < y;
  });
  if (int64_t(g.size()) <= u.Np) {
    u.error = "orbital grid smaller than No + Nv";
    return false;
  }
  if (sl(g[No]) == sl(g[No - 1]) && !hartree_fock) {
    u.error = "Zero gap system! Either change No or use Hartree-Fock!";
    return false;
  }
  g.resize(u.Np);
  u.grid = g;
  // every orbital's inverse must be an orbital of the same block
  std::map<ivec, int> where;
  for (int p = 0; p < u.Np; ++p) where[g[p]] = p;
  for (int p = 0; p < u.Np; ++p) {
    auto it = where.find({-g[p][0], -g[p][1], -g[p][2]});
    if (it == where.end() || (p < No) != (it->second < No)) {
      u.error = "UEG real orbitals need closed shells at No and No + Nv";
      return false;
    }
  }
  u.v = rs * rs * rs / 3.0 * 4.0 * pi * No * 2;   // cxx:93
  u.b = 2.0 * pi / std::pow(u.v, 1.0 / 3.0);       // cxx:94-95
  u.mad = madelung < 0 ? madelung_of(u.v) : madelung;
  const double b = u.b;
  auto w = [&](double q2) { return q2 < 1e-8 ? u.mad : 4.0 * pi / u.v / q2; };
  u.eps.resize(u.Np);
  for (int p = 0; p < u.Np; ++p) {                 // cxx:104-109
    const double k2 = b * b * double(sl(g[p]));
    double e = 0.5 * k2;
    if (hartree_fock)
      for (int o = 0; o < No; ++o) {
        const ivec d = {g[p][0] - g[o][0], g[p][1] - g[o][1], g[p][2] - g[o][2]};
        e -= w(b * b * double(sl(d)));
      }
    u.eps[p] = e;
  }
  int64_t maxr = 0;                                // cxx:133-160
  for (int p = 0; p < u.Np; ++p)
    for (int q = 0; q < u.Np; ++q)
      maxr = std::max(maxr, sl({g[q][0] - g[p][0], g[q][1] - g[p][1], g[q][2] - g[p][2]}));
  int bound = int(std::sqrt(double(maxr)));
  while (int64_t(bound + 1) * (bound + 1) <= maxr) ++bound;
  while (int64_t(bound) * bound > maxr) --bound;
  for (int a = -bound; a <= bound; ++a)
    for (int c = -bound; c <= bound; ++c)
      for (int d = -bound; d <= bound; ++d) {
        const ivec t = {a, c, d};
        if (sl(t) > maxr) continue;
        u.transfer_index[t] = int(u.transfers.size());
        u.transfers.push_back(t);
      }
  u.row.assign(u.transfers.size(), -1);
  int rows = 0;
  for (size_t n = 0; n < u.transfers.size(); ++n) {
    const ivec &t = u.transfers[n];
    const int m = u.transfer_index[{-t[0], -t[1], -t[2]}];
    if (m == int(n)) {
      u.row[n] = rows;
      rows += 1;
    } else if (int(n) < m) {
      u.row[n] = rows;
      rows += 2;
    }
  }
  u.NG = rows;
  return true;
}

}  // namespace

extern "C" int sisi_generate_ueg_impl(int No, int Nv, double rs, int hartree_fock, double madelung,
                                      int *NG, double *gamma, double *eps_i, double *eps_a,
                                      const char **error) {
  static thread_local std::string msg;
  *error = "";
  if (No <= 0 || Nv <= 0 || !(rs > 0.0) || !NG) {
    *error = "generate_ueg: No, Nv, rs must be positive and NG non-null";
    return SISI_ERR_INVALID;
  }
  Ueg u;
  if (!setup(u, No, Nv, rs, hartree_fock, madelung)) {
    msg = "generate_ueg: " + u.error;
    *error = msg.c_str();
    return SISI_ERR_INVALID;
  }
  if (gamma && *NG != u.NG) {
    msg = "generate_ueg: gamma holds NG = " + std::to_string(*NG) + " rows, the system has " +
          std::to_string(u.NG);
    *error = msg.c_str();
    return SISI_ERR_INVALID;
  }
  *NG = u.NG;
  if (eps_i) std::copy(u.eps.begin(), u.eps.begin() + No, eps_i);
  if (eps_a) std::copy(u.eps.begin() + No, u.eps.end(), eps_a);
  if (!gamma) return SISI_OK;

  const int64_t Np = u.Np, NGr = u.NG;
  std::fill(gamma, gamma + NGr * Np * Np, 0.0);
  // real orbital p' = sum over at most two plane waves of coef * |p>
  struct Term {
    int p;
    std::complex<double> c;
  };
  std::vector<std::array<Term, 2>> orb(Np);
  std::vector<int> nterm(Np, 0);
  std::map<ivec, int> where;
  for (int p = 0; p < Np; ++p) where[u.grid[p]] = p;
  const double s2 = std::sqrt(0.5);
  for (int i = 0; i < Np; ++i) {
    const int j = where[{-u.grid[i][0], -u.grid[i][1], -u.grid[i][2]}];
    if (j == i) {
      orb[i][0] = {i, 1.0};
      nterm[i] = 1;
    } else if (i < j) {
      orb[i] = {Term{i, s2}, Term{j, s2}};
      orb[j] = {Term{i, {0.0, -s2}}, Term{j, {0.0, s2}}};
      nterm[i] = nterm[j] = 2;
    }
  }
  const double pi = 3.14159265358979323846;
  // M_G[p', r'] = sum conj(U[p',p]) U[r',r] sqrt(w(g_r - g_p)) over the pairs
  // with transfer G; only representatives of (G, -G) and G = 0 are formed.
  for (int64_t rp = 0; rp < Np; ++rp)
    for (int64_t pp = 0; pp < Np; ++pp)
      for (int a = 0; a < nterm[pp]; ++a)
        for (int c = 0; c < nterm[rp]; ++c) {
          const int p = orb[pp][a].p, r = orb[rp][c].p;
          const ivec d = {u.grid[r][0] - u.grid[p][0], u.grid[r][1] - u.grid[p][1],
                          u.grid[r][2] - u.grid[p][2]};
          const int n = u.transfer_index.at(d);
          const int row = u.row[n];
          if (row < 0) continue;
          const int64_t s = sl(d);
          const double val = s ? 4.0 * pi / u.v / (u.b * u.b * double(s)) : u.mad;   // cxx:221-222
          const std::complex<double> m = std::conj(orb[pp][a].c) * orb[rp][c].c * std::sqrt(val);
          const int64_t at = pp + Np * rp;
          if (s == 0) {
            gamma[row + NGr * at] += m.real();
          } else {
            gamma[row + NGr * at] += std::sqrt(2.0) * m.real();
            gamma[row + 1 + NGr * at] += std::sqrt(2.0) * m.imag();
          }
        }
  return SISI_OK;
}
