// CPU harness for sisi4s_amd/host/davidson.hpp, driven by tests/test_davidson.py.
// Reads from stdin: mode ("eig" or "davidson"), n, then for "davidson" the
// root count, tolerance_vectors, tolerance_energy, max_basis, max_iterations,
// min_iterations and refresh flag, then the n x n matrix (row-major) and, for
// "davidson", its preconditioner diagonal.  Writes the eigenvalues one per
// line ("re im"), then for "davidson" a line "iterations N converged C rms R".
#include <cstdio>
#include <iostream>
#include <string>

#include "../sisi4s_amd/host/davidson.hpp"

int main() {
  std::string mode;
  int n = 0;
  std::cin >> mode >> n;
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
  // residual check of the returned vectors, |A x - theta x| / |x| per root
  for (size_t k = 0; k < res.vectors.size(); ++k) {
    const auto y = apply(res.vectors[k]);
    double r = 0.0, xx = 0.0;
    for (int i = 0; i < n; ++i) {
      const double d = y[size_t(i)] - res.values[k].real() * res.vectors[k][size_t(i)];
      r += d * d;
      xx += res.vectors[k][size_t(i)] * res.vectors[k][size_t(i)];
    }
    std::printf("residual %.17g\n", std::sqrt(r / xx));
  }
  std::printf("iterations %d converged %d rms %.17g applies %d\n", res.iterations, int(res.converged), res.rms,
              applies - int(res.vectors.size()));
  return 0;
}
