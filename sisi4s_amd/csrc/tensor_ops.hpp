// Bandwidth (HBM-bound) kernels: strided copies / permutations, elementwise
// updates and deterministic FP64 reductions.
#pragma once

#include "engine.hpp"

namespace sisi {

// dst = alpha * src + beta * dst over dst's logical shape; src must have the
// same order and lengths (any strides).  beta == 0 never reads dst.
void copy(Engine &e, const View &dst, const View &src, double alpha = 1.0,
          double beta = 0.0);

// dst = 0 (any strides).
void zero(Engine &e, const View &dst);
void zero(Engine &e, double *p, int64_t n);

// p[i] = deterministic pseudo-random values in [-1, 1) (timing operands).
void fill_pattern(Engine &e, double *p, int64_t n, uint32_t seed);

// y = a*x + b*y on contiguous arrays.
void axpby(Engine &e, int64_t n, double a, const double *x, double b, double *y);

// X[a,b,i,j] = T2[a,b,i,j] + f * T1[a,i] * T1b[b,j]  (X may alias T2;
// T1b == nullptr means T1b = T1).
void add_t1t1(Engine &e, double *X, const double *T2, const double *T1,
              double f, int Nv, int No, const double *T1b = nullptr);

// T2[a,b,i,j] = T2[b,a,j,i] = (T2[a,b,i,j] + T2[b,a,j,i]) / 2 in place (real
// doubles, exactly symmetric afterwards).
void symmetrize_doubles(Engine &e, double *T2, int Nv, int No);

// C[R,S,ij] *= V[R0 + R, S0 + S]: C planar complex (Rx x Ry x nij per plane),
// V planar complex with leading dimension ldV and plane size nV; real: only
// the real planes (both operands real).
void scale_pairs(Engine &e, double *C, int64_t Rx, int64_t Ry, int64_t nij, const double *V,
                 int64_t ldV, int64_t nV, int64_t R0, int64_t S0, bool real);

// Deterministic FP64 sum_k x[k]*y[k] (returned to the host).
double dot(Engine &e, int64_t n, const double *x, const double *y);

// Amplitude estimate of ClusterSinglesDoublesAlgorithm.cxx:302-331 on one
// component (order 2 for T1, 4 for T2; both planes in complex mode, the
// denominators being real), fused with the amplitude change:
//   Test = -(R - shift*T) / (D + shift),  dT = Test - T,
// D from the eigenenergies.  R is overwritten by Test.
void estimate_amplitudes(Engine &e, int order, double *R, const double *T,
                         double *dT, double shift);

// Taped reductions (engines with a tape, batched solver): the deterministic
// sum lands in device memory at out (part: kTapeRedBlocks doubles of
// partials; 2 kTapeRedBlocks and two outputs, direct and exchange halves of
// energy_terms without the factors 2 and -1, for energy_into).
void dot_into(Engine &e, int64_t n, const double *x, const double *y, double *part,
              double *out);
void energy_into(Engine &e, const double *T1, const double *T2, double *part, double *out);

// Energy terms of getEnergy (cxx:131-189): direct = 2 sum tau V,
// exchange = -sum tau Vbaij, tau = T2 + T1 T1.
void energy_terms(Engine &e, const double *T1, const double *T2,
                  double &direct, double &exchange);

}  // namespace sisi
