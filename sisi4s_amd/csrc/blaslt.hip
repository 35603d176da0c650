// Plain bf16 GEMMs (FP32 accumulation and output) through hipBLASLt, for the
// products of split-BF16x3 contractions on pre-split operand planes:
//   C(M x N, fp32, ld M) = alpha * A^T B + beta * C,
// A stored as M rows of K bf16 (k contiguous, row stride lda), B as N rows of
// K bf16 (row stride ldb).  The engine keeps one hipBLASLt handle, a
// workspace and the heuristic's algorithm per shape.
#include <hipblaslt/hipblaslt.h>

#include <cstring>
#include <map>
#include <tuple>

#include "blaslt.hpp"

namespace sisi {

namespace {

struct LtState {
  hipblasLtHandle_t handle = nullptr;
  DBuf workspace;
  size_t ws_bytes = 0;
  std::map<std::tuple<int, int, int, int64_t, int64_t>, hipblasLtMatmulAlgo_t> algos;
  std::map<std::tuple<int, int, int, int64_t, int64_t>, bool> unavailable;
  ~LtState() {
    if (handle) hipblasLtDestroy(handle);
  }
};

#define LT_CHECK(x)                                                                      \
  do {                                                                                   \
    hipblasStatus_t s_ = (x);                                                            \
    SISI_REQUIRE(s_ == HIPBLAS_STATUS_SUCCESS, SISI_ERR_HIP,                             \
                 std::string("hipBLASLt: ") + #x + " failed (" + std::to_string(int(s_)) + ")"); \
  } while (0)

LtState &state(Engine &e) {
  if (!e.blaslt) {
    auto *s = new LtState();
    e.blaslt = std::shared_ptr<void>(s, [](void *p) { delete static_cast<LtState *>(p); });
    LT_CHECK(hipblasLtCreate(&s->handle));
    s->ws_bytes = size_t(64) << 20;
    s->workspace = DBuf(&e.pool, s->ws_bytes / sizeof(double));
  }
  return *static_cast<LtState *>(e.blaslt.get());
}

// RAII holders for the per-call descriptors
struct Desc {
  hipblasLtMatmulDesc_t d = nullptr;
  hipblasLtMatrixLayout_t a = nullptr, b = nullptr, c = nullptr;
  hipblasLtMatmulPreference_t pref = nullptr;
  ~Desc() {
    if (pref) hipblasLtMatmulPreferenceDestroy(pref);
    if (c) hipblasLtMatrixLayoutDestroy(c);
    if (b) hipblasLtMatrixLayoutDestroy(b);
    if (a) hipblasLtMatrixLayoutDestroy(a);
    if (d) hipblasLtMatmulDescDestroy(d);
  }
};

}  // namespace

bool blaslt_bf16_gemm(Engine &e, int M, int N, int K, const uint16_t *A, int64_t lda,
                      const uint16_t *B, int64_t ldb, float *C, float alpha, float beta) {
  LtState &s = state(e);
  const auto key = std::make_tuple(M, N, K, lda, ldb);
  if (s.unavailable.count(key)) return false;
  Desc d;
  LT_CHECK(hipblasLtMatmulDescCreate(&d.d, HIPBLAS_COMPUTE_32F, HIP_R_32F));
  const hipblasOperation_t ta = HIPBLAS_OP_T, tb = HIPBLAS_OP_N;
  LT_CHECK(hipblasLtMatmulDescSetAttribute(d.d, HIPBLASLT_MATMUL_DESC_TRANSA, &ta, sizeof(ta)));
  LT_CHECK(hipblasLtMatmulDescSetAttribute(d.d, HIPBLASLT_MATMUL_DESC_TRANSB, &tb, sizeof(tb)));
  // column-major views: A is K x M (ld lda), B is K x N (ld ldb), C is M x N (ld M)
  LT_CHECK(hipblasLtMatrixLayoutCreate(&d.a, HIP_R_16BF, uint64_t(K), uint64_t(M), lda));
  LT_CHECK(hipblasLtMatrixLayoutCreate(&d.b, HIP_R_16BF, uint64_t(K), uint64_t(N), ldb));
  LT_CHECK(hipblasLtMatrixLayoutCreate(&d.c, HIP_R_32F, uint64_t(M), uint64_t(N), int64_t(M)));
  auto it = s.algos.find(key);
  if (it == s.algos.end()) {
    LT_CHECK(hipblasLtMatmulPreferenceCreate(&d.pref));
    LT_CHECK(hipblasLtMatmulPreferenceSetAttribute(
        d.pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &s.ws_bytes, sizeof(s.ws_bytes)));
    hipblasLtMatmulHeuristicResult_t res[4];
    int n = 0;
    const hipblasStatus_t st = hipblasLtMatmulAlgoGetHeuristic(s.handle, d.d, d.a, d.b, d.c, d.c,
                                                               d.pref, 4, res, &n);
    if (st != HIPBLAS_STATUS_SUCCESS || n == 0) {
      s.unavailable[key] = true;
      return false;
    }
    it = s.algos.emplace(key, res[0].algo).first;
  }
  LT_CHECK(hipblasLtMatmul(s.handle, d.d, &alpha, A, d.a, B, d.b, &beta, C, d.c, C, d.c,
                           &it->second, s.workspace.get(), s.ws_bytes, e.stream));
  return true;
}

}  // namespace sisi
