// Index-string contractions on device tensors.
//
// The reference writes every term of the path as a CTF expression such as
// C["abij"] += A["acik"] * B["cbkj"].  contract() evaluates one such binary
// product with a single MFMA GEMM: labels shared by A, B and C are a batch
// index, labels of A and C only form the GEMM row index m, labels of B and C
// only the column index n, labels of A and B only the contraction index k.
// Operands whose label groups are not contiguous in memory are first copied
// into a dense layout; the output may be written in any order of up to three
// runs per group through the GEMM's split epilogue.
#pragma once

#include <string>

#include "engine.hpp"

namespace sisi {

// C[ic] = alpha * A[ia] * B[ib] + beta * C[ic].  precision < 0 uses the
// engine's current precision.
void contract(Engine &e, double alpha, const View &A, const std::string &ia,
              const View &B, const std::string &ib, double beta, const View &C,
              const std::string &ic, int precision = -1);

// einsum-style spelling: ctr(e, alpha, "cdkl,adkl->ac", A, B, beta, C).
void ctr(Engine &e, double alpha, const char *spec, const View &A, const View &B,
         double beta, const View &C, int precision = -1);

// C[ic] = alpha * A[ia] + beta * C[ic] where ia is a permutation of ic
// (CTF's C["abij"] += A["baji"]).
void sum(Engine &e, double alpha, const View &A, const std::string &ia, double beta,
         const View &C, const std::string &ic);

}  // namespace sisi
