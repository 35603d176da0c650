// Planar complex tensors: the real part, then the imaginary part of the same
// logical tensor, each dense and first index fastest.  A complex product is
// four real contractions of the planes,
//   Re C = Re A Re B - sA sB Im A Im B,   Im C = sB Re A Im B + sA Im A Re B,
// with sA = -1 when A enters conjugated (likewise sB).
#pragma once

#include "contract.hpp"
#include "tensor_ops.hpp"

namespace sisi {

struct ZV {
  View re, im;
};

inline ZV zdense(double *p, std::initializer_list<int64_t> lens) {
  int64_t n = 1;
  for (int64_t l : lens) n *= l;
  return {dense(p, lens), dense(p + n, lens)};
}
inline ZV Z1(double *p, int64_t a) { return zdense(p, {a}); }
inline ZV Z2(double *p, int64_t a, int64_t b) { return zdense(p, {a, b}); }
inline ZV Z3(double *p, int64_t a, int64_t b, int64_t c) { return zdense(p, {a, b, c}); }
inline ZV Z4(double *p, int64_t a, int64_t b, int64_t c, int64_t d) { return zdense(p, {a, b, c, d}); }
inline ZV zsub(const ZV &v, std::initializer_list<int64_t> lo, std::initializer_list<int64_t> hi) {
  return {subview(v.re, lo, hi), subview(v.im, lo, hi)};
}

// C = alpha * op(A) op(B) + beta * C, op = conj when the flag is set.
inline void zctr(Engine &e, double alpha, const char *spec, const ZV &A, bool ca, const ZV &B, bool cb,
          double beta, const ZV &C, int prec) {
  const double sa = ca ? -1.0 : 1.0, sb = cb ? -1.0 : 1.0;
  ctr(e, alpha, spec, A.re, B.re, beta, C.re, prec);
  ctr(e, -alpha * sa * sb, spec, A.im, B.im, 1.0, C.re, prec);
  ctr(e, alpha * sb, spec, A.re, B.im, beta, C.im, prec);
  ctr(e, alpha * sa, spec, A.im, B.re, 1.0, C.im, prec);
}

// C[ic] = alpha * op(A)[ia] + beta * C[ic] (ia a permutation of ic).
inline void zsum(Engine &e, double alpha, const ZV &A, const char *ia, bool ca, double beta, const ZV &C,
          const char *ic) {
  sum(e, alpha, A.re, ia, beta, C.re, ic);
  sum(e, ca ? -alpha : alpha, A.im, ia, beta, C.im, ic);
}

// dst = alpha * op(src) + beta * dst (same index order, any strides).
inline void zcopy(Engine &e, const ZV &dst, const ZV &src, bool conj = false, double alpha = 1.0,
           double beta = 0.0) {
  copy(e, dst.re, src.re, alpha, beta);
  copy(e, dst.im, src.im, conj ? -alpha : alpha, beta);
}

}  // namespace sisi
