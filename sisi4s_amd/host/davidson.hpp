// Host side of the EOM eigenproblems: the eigenvalues and right eigenvectors
// of the small non-symmetric projected matrices, and a Davidson driver for
// the lowest eigenpairs of a non-symmetric operator given as a right-apply
// (the role of EigenSystemDavidsonMono, math/EigenSystemDavidson.hpp, with a
// diagonal preconditioner as IPCcsdPreconditioner's, CcsdPreconditioner.cxx).
// Header only, no device code: the right-apply is a callback.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <functional>
#include <stdexcept>
#include <vector>

namespace sisi_host {

// Eigenvalues of the real n x n matrix A (row-major): reduction to upper
// Hessenberg form by Householder reflections, then the implicit double-shift
// QR algorithm of Francis with deflation on negligible subdiagonals; a 2 x 2
// block that splits off gives its two eigenvalues, a complex pair included.  Unordered.  The role of LAPACK's
// dgeev in LapackGeneralEigenSystem (util/LapackGeneralEigenSystem.hpp:35-110)
// for the small projected matrices of the Davidson driver.
inline std::vector<std::complex<double>> small_eigenvalues(std::vector<double> A, int n) {
  auto at = [&](int i, int j) -> double & { return A[size_t(i) * n + j]; };
  // Householder reduction to Hessenberg form
  for (int k = 0; k + 2 < n; ++k) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += at(i, k) * at(i, k);
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    if (at(k + 1, k) > 0) alpha = -alpha;
    std::vector<double> v(size_t(n), 0.0);
    v[size_t(k + 1)] = at(k + 1, k) - alpha;
    for (int i = k + 2; i < n; ++i) v[size_t(i)] = at(i, k);
    double vn = 0.0;
    for (int i = k + 1; i < n; ++i) vn += v[size_t(i)] * v[size_t(i)];
    if (vn == 0.0) continue;
    for (int j = 0; j < n; ++j) {  // A = (I - 2vv'/v'v) A
      double s = 0.0;
      for (int i = k + 1; i < n; ++i) s += v[size_t(i)] * at(i, j);
      s = 2.0 * s / vn;
      for (int i = k + 1; i < n; ++i) at(i, j) -= s * v[size_t(i)];
    }
    for (int i = 0; i < n; ++i) {  // A = A (I - 2vv'/v'v)
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += at(i, j) * v[size_t(j)];
      s = 2.0 * s / vn;
      for (int j = k + 1; j < n; ++j) at(i, j) -= s * v[size_t(j)];
    }
  }
  std::vector<std::complex<double>> ev;
  auto pair_of = [&](int k) {  // eigenvalues of the 2 x 2 block at rows/columns k, k+1
    const double a = at(k, k), b = at(k, k + 1), c = at(k + 1, k), d = at(k + 1, k + 1);
    const double half = 0.5 * (a - d), disc = half * half + b * c, mid = 0.5 * (a + d);
    if (disc >= 0.0) {
      const double r = std::sqrt(disc);
      ev.emplace_back(mid + r, 0.0);
      ev.emplace_back(mid - r, 0.0);
    } else {
      const double r = std::sqrt(-disc);
      ev.emplace_back(mid, r);
      ev.emplace_back(mid, -r);
    }
  };
  // Householder reflector of a 3-vector applied to rows r..r+2 (columns
  // c0..n-1) and columns r..r+2 (rows 0..rmax)
  auto reflect3 = [&](int r, double x, double y, double z, int c0, int rmax) {
    const double nrm = std::sqrt(x * x + y * y + z * z);
    if (nrm == 0.0) return;
    const double alpha = x > 0 ? -nrm : nrm;
    const double v0 = x - alpha, v1 = y, v2 = z;
    const double beta = 2.0 / (v0 * v0 + v1 * v1 + v2 * v2);
    for (int j = c0; j < n; ++j) {
      const double s = beta * (v0 * at(r, j) + v1 * at(r + 1, j) + v2 * at(r + 2, j));
      at(r, j) -= s * v0;
      at(r + 1, j) -= s * v1;
      at(r + 2, j) -= s * v2;
    }
    for (int i = 0; i <= rmax; ++i) {
      const double s = beta * (v0 * at(i, r) + v1 * at(i, r + 1) + v2 * at(i, r + 2));
      at(i, r) -= s * v0;
      at(i, r + 1) -= s * v1;
      at(i, r + 2) -= s * v2;
    }
  };
  int hi = n - 1, iter = 0;
  while (hi >= 0) {
    int lo = hi;  // active unreduced block [lo, hi]
    while (lo > 0) {
      double s = std::fabs(at(lo, lo)) + std::fabs(at(lo - 1, lo - 1));
      if (s == 0.0) s = 1.0;
      if (std::fabs(at(lo, lo - 1)) <= 2.2e-16 * s) {
        at(lo, lo - 1) = 0.0;
        break;
      }
      --lo;
    }
    if (lo == hi) {
      ev.emplace_back(at(hi, hi), 0.0);
      --hi;
      iter = 0;
      continue;
    }
    if (lo == hi - 1) {
      pair_of(hi - 1);
      hi -= 2;
      iter = 0;
      continue;
    }
    if (iter > 300) throw std::runtime_error("small_eigenvalues: QR iteration did not converge");
    // Francis double shift (Golub and Van Loan, Algorithm 7.5.1): the shifts
    // are the eigenvalues of the trailing 2 x 2 block, passed as their sum s
    // and product t; exceptional shifts after 10 and 20 sweeps
    double s = at(hi - 1, hi - 1) + at(hi, hi);
    double t = at(hi - 1, hi - 1) * at(hi, hi) - at(hi - 1, hi) * at(hi, hi - 1);
    if (iter == 10 || iter == 20) {
      const double w = std::fabs(at(hi, hi - 1)) + std::fabs(at(hi - 1, hi - 2));
      s = 1.5 * w;
      t = w * w;
    }
    ++iter;
    double x = at(lo, lo) * at(lo, lo) + at(lo, lo + 1) * at(lo + 1, lo) - s * at(lo, lo) + t;
    double y = at(lo + 1, lo) * (at(lo, lo) + at(lo + 1, lo + 1) - s);
    double z = at(lo + 1, lo) * at(lo + 2, lo + 1);
    for (int k = lo; k <= hi - 2; ++k) {
      reflect3(k, x, y, z, k > lo ? k - 1 : lo, std::min(k + 3, hi));
      if (k < hi - 2) {  // the bulge moved one row down
        x = at(k + 1, k);
        y = at(k + 2, k);
        z = at(k + 3, k);
      }
    }
    {  // closing Givens rotation on rows/columns hi-1, hi
      const double a = at(hi - 1, hi - 2), b = at(hi, hi - 2);
      const double r = std::hypot(a, b);
      if (r != 0.0) {
        const double cc = a / r, ss = b / r;
        for (int j = hi - 2; j < n; ++j) {
          const double p = at(hi - 1, j), q = at(hi, j);
          at(hi - 1, j) = cc * p + ss * q;
          at(hi, j) = -ss * p + cc * q;
        }
        for (int i = 0; i <= hi; ++i) {
          const double p = at(i, hi - 1), q = at(i, hi);
          at(i, hi - 1) = cc * p + ss * q;
          at(i, hi) = -ss * p + cc * q;
        }
      }
    }
  }
  return ev;
}

// Right eigenvector of A (row-major n x n) for the eigenvalue lambda (real,
// or one of a complex-conjugate pair) by inverse iteration: LU with partial
// pivoting of A - (lambda + delta) I in complex arithmetic, three sweeps from
// the all-ones vector, normalised to unit 2-norm.  For a complex lambda the
// real and imaginary parts of the result span the pair's real invariant
// subspace: A Re y = Re(lambda) Re y - Im(lambda) Im y.
inline std::vector<std::complex<double>> small_eigenvector_complex(const std::vector<double> &A, int n,
                                                                   std::complex<double> lambda) {
  using cd = std::complex<double>;
  const double delta = 1e-10 * (1.0 + std::abs(lambda));
  std::vector<cd> M(A.begin(), A.end());
  for (int i = 0; i < n; ++i) M[size_t(i) * n + i] -= lambda + delta;
  std::vector<int> piv(static_cast<size_t>(n));
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(M[size_t(i) * n + k]) > std::abs(M[size_t(p) * n + k])) p = i;
    piv[size_t(k)] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(M[size_t(k) * n + j], M[size_t(p) * n + j]);
    cd pk = M[size_t(k) * n + k];
    if (pk == cd(0.0)) pk = M[size_t(k) * n + k] = 1e-300;
    for (int i = k + 1; i < n; ++i) {
      const cd f = M[size_t(i) * n + k] / pk;
      M[size_t(i) * n + k] = f;
      for (int j = k + 1; j < n; ++j) M[size_t(i) * n + j] -= f * M[size_t(k) * n + j];
    }
  }
  std::vector<cd> x(size_t(n), cd(1.0));
  for (int it = 0; it < 3; ++it) {
    for (int k = 0; k < n; ++k) std::swap(x[size_t(k)], x[size_t(piv[size_t(k)])]);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) x[size_t(i)] -= M[size_t(i) * n + j] * x[size_t(j)];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) x[size_t(i)] -= M[size_t(i) * n + j] * x[size_t(j)];
      x[size_t(i)] /= M[size_t(i) * n + i];
    }
    double nrm = 0.0;
    for (const cd &v : x) nrm += std::norm(v);
    nrm = std::sqrt(nrm);
    for (cd &v : x) v /= nrm;
  }
  return x;
}

// Real right eigenvector for a real eigenvalue lambda (the real part of the
// complex inverse iteration's result, which is real up to rounding).
inline std::vector<double> small_eigenvector(const std::vector<double> &A, int n, double lambda) {
  const auto z = small_eigenvector_complex(A, n, std::complex<double>(lambda, 0.0));
  std::vector<double> x(static_cast<size_t>(n));
  double nrm = 0.0;
  for (int i = 0; i < n; ++i) {
    x[size_t(i)] = z[size_t(i)].real();
    nrm += x[size_t(i)] * x[size_t(i)];
  }
  nrm = std::sqrt(nrm);
  for (double &v : x) v /= nrm;
  return x;
}

struct DavidsonResult {
  std::vector<std::complex<double>> values;  // ascending real part
  std::vector<std::vector<double>> vectors;  // right eigenvectors (real part of the Ritz vector)
  int iterations = 0;
  double rms = 0.0;                          // sum_k |r_k|^2 / |x_k|^2 of the last iteration
  double energy_difference = 0.0;            // sum_k |theta_k - theta_k(previous)|
  bool converged = false;
};

struct DavidsonOptions {
  int nroots = 1;
  double tolerance_vectors = 1e-6;   // amplitudesConvergence
  double tolerance_energy = 1e-6;    // energyConvergence
  int max_basis = 1000;              // maxBasisSize
  int max_iterations = 32;           // maxIterations
  int min_iterations = 1;            // minIterations
  bool refresh_on_max_basis = false; // refreshOnMaxBasisSize
};

// Lowest eigenpairs (by real part) of the non-symmetric operator `apply`
// (dimension n) with the diagonal preconditioner `diag`, following
// EigenSystemDavidsonMono::run (math/EigenSystemDavidson.hpp:114-428):
//  * the initial basis is the unit vectors on the smallest diagonal elements
//    (IPCcsdPreconditioner::getInitialBasis, CcsdPreconditioner.cxx:575-620);
//  * each iteration projects onto the basis, takes the lowest Ritz values,
//    forms the residuals r_k = H x_k - theta_k x_k and appends the corrections
//    r_k / (theta_k - diag), zero where |theta_k - diag| < 1e-4
//    (CcsdPreconditioner.cxx:678-700), orthonormalised against the basis and
//    dropped when orthogonalisation leaves less than 1e-8 of their norm (the
//    reference drops them below an absolute norm of 1e-6, which stalls the
//    residuals near that size; here only linearly dependent ones go), the
//    iteration ending when no correction is left;
//  * it stops once rms = sum_k |r_k|^2/|x_k|^2 < nroots * tolerance_vectors
//    or sum_k |dtheta_k| / nroots < tolerance_energy, or past max_iterations,
//    or when the basis outgrows max_basis unless refresh_on_max_basis, which
//    restarts from the current Ritz vectors instead.
// The reference applies H to every basis vector in every iteration and once
// more per root for the residuals; here the products H b_j are kept with the
// basis and formed only when an iteration uses them, so H is applied once per
// basis vector in all and the residuals come from the kept products (on the
// IP-EOM-CCSD matrix of a 156-dimensional closed-shell case: 4 roots in 7
// iterations take 26 right-applies against the reference scheme's 140).
inline DavidsonResult davidson(const std::function<std::vector<double>(const std::vector<double> &)> &apply,
                               const std::vector<double> &diag, const std::vector<std::vector<double>> &initial,
                               const DavidsonOptions &opt) {
  const int n = int(diag.size());
  const int nroots = opt.nroots;
  if (nroots < 1 || nroots > n) throw std::invalid_argument("davidson: eigenstates must lie in [1, dimension]");
  std::vector<std::vector<double>> V, W;
  auto dot = [&](const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[size_t(i)] * b[size_t(i)];
    return s;
  };
  auto orthogonalise = [&](std::vector<double> &t) {  // against V, twice for stability
    for (int pass = 0; pass < 2; ++pass)
      for (const auto &v : V) {
        const double c = dot(v, t);
        for (int i = 0; i < n; ++i) t[size_t(i)] -= c * v[size_t(i)];
      }
    return std::sqrt(dot(t, t));
  };
  auto append = [&](std::vector<double> t, double min_relative_norm) {
    const double before = std::sqrt(dot(t, t));
    const double nrm = orthogonalise(t);
    if (!(nrm > 0.0) || !(nrm >= min_relative_norm * before)) return false;
    for (double &x : t) x /= nrm;
    V.push_back(std::move(t));  // H b is formed when the next iteration needs it
    return true;
  };
  for (const auto &b : initial) {
    if (int(b.size()) != n) throw std::invalid_argument("davidson: initial vector of the wrong dimension");
    append(b, 1e-12);
  }
  if (int(V.size()) < nroots) throw std::invalid_argument("davidson: initial basis spans fewer vectors than eigenstates");
  DavidsonResult res;
  std::vector<std::complex<double>> prev(size_t(nroots), 0.0);
  for (int it = 1;; ++it) {
    if (opt.refresh_on_max_basis && int(V.size()) >= opt.max_basis && !res.vectors.empty()) {
      // restart from the Ritz vectors; if they span fewer than nroots
      // directions, refill from the old basis in its order
      std::vector<std::vector<double>> keep = res.vectors, old = V;
      V.clear();
      W.clear();
      for (auto &x : keep) append(x, 1e-12);
      for (size_t j = 0; j < old.size() && int(V.size()) < nroots; ++j) append(old[j], 1e-8);
      if (int(V.size()) < nroots)
        throw std::runtime_error("davidson: restarted basis spans fewer vectors than eigenstates");
    }
    while (W.size() < V.size()) {
      W.push_back(apply(V[W.size()]));
      if (int(W.back().size()) != n) throw std::runtime_error("davidson: right-apply changed the dimension");
    }
    const int m = int(V.size());
    std::vector<double> Hs(size_t(m) * m);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) Hs[size_t(i) * m + j] = dot(V[size_t(i)], W[size_t(j)]);
    auto ev = small_eigenvalues(Hs, m);
    std::stable_sort(ev.begin(), ev.end(), [](const std::complex<double> &a, const std::complex<double> &b) {
      return a.real() < b.real();
    });
    if (int(ev.size()) < nroots) throw std::runtime_error("davidson: basis smaller than the number of eigenstates");
    res.values.assign(ev.begin(), ev.begin() + nroots);
    res.vectors.assign(size_t(nroots), std::vector<double>(size_t(n), 0.0));
    res.rms = 0.0;
    res.energy_difference = 0.0;
    for (int k = 0; k < nroots; ++k) {
      res.energy_difference += std::abs(res.values[size_t(k)] - prev[size_t(k)]);
      prev[size_t(k)] = res.values[size_t(k)];
    }
    std::vector<std::vector<double>> corrections;
    // Ritz vectors and residuals.  A real theta gives one real Ritz vector.
    // A complex-conjugate pair (theta, conj theta) gives the complex Ritz
    // vector y; Re y and Im y span the pair's invariant subspace and serve as
    // the two roots' vectors (one shift for both would give the same vector
    // twice), with the complex residual r = H y - theta y split the same way
    // and the corrections Re and Im of r / (theta - diag).  A pair cut by the
    // root count keeps Re y as the last root and still adds both corrections.
    for (int k = 0; k < nroots;) {
      const std::complex<double> theta = ev[size_t(k)];
      const bool cplx = theta.imag() != 0.0;
      const auto y = small_eigenvector_complex(Hs, m, theta);
      std::vector<std::complex<double>> x(size_t(n), 0.0), r(size_t(n), 0.0);
      for (int j = 0; j < m; ++j)
        for (int i = 0; i < n; ++i) {
          x[size_t(i)] += y[size_t(j)] * V[size_t(j)][size_t(i)];
          r[size_t(i)] += y[size_t(j)] * W[size_t(j)][size_t(i)];
        }
      for (int i = 0; i < n; ++i) r[size_t(i)] -= theta * x[size_t(i)];
      const int parts = cplx ? 2 : 1;
      for (int part = 0; part < parts; ++part) {
        std::vector<double> xr(static_cast<size_t>(n)), rr(static_cast<size_t>(n)),
            t(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
          xr[size_t(i)] = part ? x[size_t(i)].imag() : x[size_t(i)].real();
          rr[size_t(i)] = part ? r[size_t(i)].imag() : r[size_t(i)].real();
          const std::complex<double> den = theta - diag[size_t(i)];
          const std::complex<double> c = std::abs(den) < 1e-4 ? 0.0 : r[size_t(i)] / den;
          t[size_t(i)] = part ? c.imag() : c.real();
        }
        corrections.push_back(std::move(t));
        if (k + part < nroots) {
          const double xx = dot(xr, xr);
          res.rms += xx > 0.0 ? dot(rr, rr) / xx : 0.0;
          res.vectors[size_t(k + part)] = std::move(xr);
        }
      }
      k += parts;
    }
    res.iterations = it;
    int appended = 0;
    for (auto &t : corrections) appended += append(std::move(t), 1e-8) ? 1 : 0;
    const bool small_rms = res.rms < nroots * opt.tolerance_vectors;
    const bool small_de = res.energy_difference / nroots < opt.tolerance_energy;
    res.converged = small_rms || small_de;
    if (it + 1 <= opt.min_iterations) continue;
    const bool basis_ok = opt.refresh_on_max_basis || int(V.size()) <= opt.max_basis;
    if (res.converged || !basis_ok || it + 1 > opt.max_iterations || appended == 0) break;
  }
  return res;
}

// Unit vectors on the `count` smallest diagonal elements (ties by index).
inline std::vector<std::vector<double>> lowest_diagonal_basis(const std::vector<double> &diag, int count) {
  std::vector<int> order(diag.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = int(i);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return diag[size_t(a)] < diag[size_t(b)]; });
  std::vector<std::vector<double>> basis;
  for (int k = 0; k < count && k < int(order.size()); ++k) {
    basis.emplace_back(diag.size(), 0.0);
    basis.back()[size_t(order[size_t(k)])] = 1.0;
  }
  return basis;
}

}  // namespace sisi_host
