// Plain bf16 GEMMs through hipBLASLt (FP32 accumulation and output).
#pragma once

#include <cstdint>

#include "engine.hpp"

namespace sisi {

// C (M x N fp32, column-major, ld M) = alpha * A^T B + beta * C, where A holds
// M rows of K bf16 (row stride lda) and B N rows of K bf16 (row stride ldb).
// Returns false when hipBLASLt offers no algorithm for the shape.
bool blaslt_bf16_gemm(Engine &e, int M, int N, int K, const uint16_t *A, int64_t lda,
                      const uint16_t *B, int64_t ldb, float *C, float alpha, float beta);

}  // namespace sisi
