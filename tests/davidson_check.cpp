// CPU harness for sisi4s_amd/host/davidson.hpp and eom_space.hpp, driven by
// tests/test_davidson.py.  Modes "unpack", "pack", "diag" and "classes" read
// "ea o v" (classes: "o v") and then the vectors described at each branch.
// Reads from stdin: mode ("eig" or "davidson"), n, then for "davidson" the
// root count, tolerance_vectors, tolerance_energy, max_basis, max_iterations,
// min_iterations and refresh flag, then the n x n matrix (row-major) and, for
// "davidson", its preconditioner diagonal.  Writes the eigenvalues one per
// line ("re im"), then for "davidson" a line "iterations N converged C rms R".
#include <cstdio>
#include <iostream>
#include <string>

#include "../sisi4s_amd/host/davidson.hpp"
#include "../sisi4s_amd/host/eom_space.hpp"

static void print_line(const std::vector<double> &x) {
  for (size_t i = 0; i < x.size(); ++i) std::printf(i ? " %.17g" : "%.17g", x[i]);
  std::printf("\n");
}

static std::vector<double> read_vector(size_t n) {
  std::vector<double> x(n);
  for (auto &y : x) std::cin >> y;
  return x;
}

// EomSpace and the class reorderings: returns -1 if mode is not one of them
static int eom_modes(const std::string &mode) {
  if (mode == "classes") {  // o v, then HHPP, HPPP, HPHP (first index fastest)
    int64_t o = 0, v = 0;
    std::cin >> o >> v;
    const auto hhpp = read_vector(size_t(o * o * v * v)), hppp = read_vector(size_t(o * v * v * v)),
               hphp = read_vector(size_t(o * v * o * v));
    if (!std::cin) return 2;
    print_line(sisi_host::pphh_from_hhpp(hhpp.data(), o, v));
    print_line(sisi_host::ppph_from_hppp(hppp.data(), o, v));
    print_line(sisi_host::phph_from_hphp(hphp.data(), o, v));
    return 0;
  }
  if (mode != "unpack" && mode != "pack" && mode != "diag") return -1;
  int ea = 0;
  int64_t o = 0, v = 0;
  std::cin >> ea >> o >> v;
  const sisi_host::EomSpace space(ea != 0, o, v);
  if (mode == "diag") {  // eps_i[o], eps_a[v] -> packed diagonal
    const auto ei = read_vector(size_t(o)), eav = read_vector(size_t(v));
    if (!std::cin) return 2;
    print_line(space.diagonal(ei.data(), eav.data()));
    return 0;
  }
  int m = 0;
  std::cin >> m;
  for (int k = 0; k < m; ++k) {
    if (mode == "unpack") {  // packed vector -> line: r1 then r2
      const auto x = read_vector(space.n);
      if (!std::cin) return 2;
      std::vector<double> r1, r2;
      space.unpack(x, r1, r2);
      r1.insert(r1.end(), r2.begin(), r2.end());
      print_line(r1);
    } else {  // r1 then r2 -> packed line
      const auto s1 = read_vector(space.r1_size()), s2 = read_vector(space.r2_size());
      if (!std::cin) return 2;
      print_line(space.pack(s1, s2));
    }
  }
  return 0;
}

int main() {
  std::string mode;
  std::cin >> mode;
  const int rc = eom_modes(mode);
  if (rc >= 0) return rc;
  int n = 0;
  std::cin >> n;
  sisi_host::DavidsonOptions opt;
  int refresh = 0;
  if (mode == "davidson")
    std::cin >> opt.nroots >> opt.tolerance_vectors >> opt.tolerance_energy >> opt.max_basis >>
        opt.max_iterations >> opt.min_iterations >> refresh;
  opt.refresh_on_max_basis = refresh != 0;
  std::vector<double> A(static_cast<size_t>(n) * n);
  for (auto &x : A) std::cin >> x;
  if (!std::cin) return 2;
  if (mode == "eig") {
    for (const auto &z : sisi_host::small_eigenvalues(A, n)) std::printf("%.17g %.17g\n", z.real(), z.imag());
    return 0;
  }
  std::vector<double> diag(static_cast<size_t>(n));
  for (auto &x : diag) std::cin >> x;
  if (!std::cin) return 2;
  int applies = 0;
  auto apply = [&](const std::vector<double> &x) {
    ++applies;
    std::vector<double> y(size_t(n), 0.0);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) y[size_t(i)] += A[size_t(i) * n + j] * x[size_t(j)];
    return y;
  };
  const auto res = sisi_host::davidson(apply, diag, sisi_host::lowest_diagonal_basis(diag, opt.nroots), opt);
  for (const auto &z : res.values) std::printf("%.17g %.17g\n", z.real(), z.imag());
  // residual check of the returned vectors per root: |A x - theta x| / |x|
  // for a real theta; for one of a complex pair, the distance of A x from the
  // span of the pair's two vectors (their invariant subspace), / |x|
  for (size_t k = 0; k < res.vectors.size(); ++k) {
    const auto &x = res.vectors[k];
    const auto y = apply(x);
    auto dot = [&](const std::vector<double> &a, const std::vector<double> &b) {
      double s = 0.0;
      for (int i = 0; i < n; ++i) s += a[size_t(i)] * b[size_t(i)];
      return s;
    };
    double r = 0.0;
    const double xx = dot(x, x);
    size_t partner = k;
    if (res.values[k].imag() != 0.0)
      for (size_t q = 0; q < res.vectors.size(); ++q)
        if (q != k && std::abs(res.values[q] - std::conj(res.values[k])) <= 1e-9 * (1.0 + std::abs(res.values[k])))
          partner = q;
    if (partner == k) {
      for (int i = 0; i < n; ++i) {
        const double d = y[size_t(i)] - res.values[k].real() * x[size_t(i)];
        r += d * d;
      }
    } else {  // least squares of y on [x, z]
      const auto &z = res.vectors[partner];
      const double a11 = xx, a12 = dot(x, z), a22 = dot(z, z), b1 = dot(x, y), b2 = dot(z, y);
      const double det = a11 * a22 - a12 * a12;
      const double c1 = (b1 * a22 - b2 * a12) / det, c2 = (a11 * b2 - a12 * b1) / det;
      for (int i = 0; i < n; ++i) {
        const double d = y[size_t(i)] - c1 * x[size_t(i)] - c2 * z[size_t(i)];
        r += d * d;
      }
    }
    std::printf("residual %.17g\n", std::sqrt(r / xx));
  }
  std::printf("iterations %d converged %d rms %.17g applies %d\n", res.iterations, int(res.converged), res.rms,
              applies - int(res.vectors.size()));
  return 0;
}
