// MFMA GEMM family used by every dense contraction of the engine.
//
//   C(m,n) = alpha * sum_k A(m,k) B(k,n) + beta * C(m,n)      (per batch)
//
//   A(m,k) = A[m*sam + k*sak]   with sam == 1 or sak == 1
//   B(k,n) = B[k*sbk + n*sbn]   with sbk == 1 or sbn == 1
//   C(m,n) = C[off_m(m) + off_n(n)], where the row index m is split as
//   m = m0 + dm0*(m1 + dm1*m2) and off_m = m0*cm0 + m1*cm1 + m2*cm2 (same
//   for n).  The split output lets a GEMM write straight into a permuted
//   tensor layout (e.g. the ladder's V_xycd from a (x,c) x (y,d) product).
//
// Split-K partial sums go to a dense FP64 workspace and are reduced in a
// fixed order (deterministic results).
#pragma once

#include "common.hpp"

namespace sisi {

// The precisions of the dense contractions (SISI_FP64/_FP32/_BF16X3/_BF16).
inline bool is_precision(int p) {
  return p == SISI_FP64 || p == SISI_FP32 || p == SISI_BF16X3 || p == SISI_BF16;
}

struct GemmArgs {
  const double *A = nullptr;
  const double *B = nullptr;
  double *C = nullptr;
  double *work = nullptr;
  int64_t sam = 1, sak = 0, sbk = 1, sbn = 0;
  int64_t bsa = 0, bsb = 0, bsc = 0;
  int64_t cm0 = 1, cm1 = 0, cm2 = 0, cn0 = 0, cn1 = 0, cn2 = 0;
  int M = 0, N = 0, K = 0;
  int dm0 = 1, dm1 = 1, dn0 = 1, dn1 = 1;
  int batch = 1, ksplit = 1, kchunk = 0;
  int tiles_m = 0, tiles_n = 0;
  int split_out = 0;  // 1: output index splits need div/mod
  int xcd_map = 1;    // 1: give each of the 8 XCDs a contiguous tile range
  int group = 8;      // m-tiles per group in the tile walk
  double alpha = 1.0, beta = 0.0;
  // split-BF16x3 / bf16 kernels only: out_planes = 1 writes the result as
  // bf16 hi/lo planes instead of FP64 (C is then the hi plane as bf16, the
  // output offsets count bf16 elements, lo = hi + plane_lo; beta must be 0):
  // hi = bf16(x), lo = bf16(x - hi) of x = float(alpha * acc), the split an
  // operand split of the FP64 result would give; out_planes = 2 writes the
  // hi plane only (a result read by plain-bf16 products)
  int out_planes = 0;
  int64_t plane_lo = 0;
  // out_cube = 1: each output index splits into two coordinates (x0 = m %
  // dm0, x1 = m / dm0, likewise for n), and a coordinate x contributes
  // (x & 7) * lo + (x >> 3) * hi: lo/hi = cm0/cm1 for x0 and cm2/cm3 for x1
  // (cn* for n).  This writes tensors stored as contiguous 8 x 8 x 8 cubes.
  int out_cube = 0;
  int64_t cm3 = 0, cn3 = 0;
};

// One product of a segmented GEMM: A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn], k < K, scaled by `scale`.  ak: A is k-fastest
// (sak == 1); bnf: B is n-fastest (sbn == 1).
struct GemmSeg {
  const double *A = nullptr;
  const double *B = nullptr;
  int64_t sam = 1, sak = 0, sbk = 1, sbn = 0;
  int K = 0;
  int ak = 0, bnf = 0;
  double scale = 1.0;
};

// C = alpha * sum_s scale_s A_s B_s + beta * C in one launch, the products
// accumulated in the same registers (no batch, no split-K).  The output
// fields of `g` (C, M, N, cm*/cn*, dm*/dn*, alpha, beta) apply to the sum.
constexpr int kMaxSegs = 4;
struct SegGemmArgs {
  GemmArgs g;
  int nseg = 0;
  GemmSeg seg[kMaxSegs];
};

class Engine;

// Segmented GEMM on the engine's stream, precision SISI_FP64/_FP32/_BF16X3.
void gemm_segments(Engine &eng, int precision, SegGemmArgs a);

// An operand split once into bf16 hi/lo planes for split-BF16x3: `batch`
// matrices of Rp rows by Kp bf16 (k contiguous, zero padded), matrix b's row
// r at hi + (b*Rp + r)*Kp (lo likewise).
struct B3Planes {
  DBuf store;
  const unsigned short *hi = nullptr, *lo = nullptr;
  int64_t Rp = 0, Kp = 0, batch = 0;
  const unsigned short *hi_of(int64_t b) const { return hi + b * Rp * Kp; }
  const unsigned short *lo_of(int64_t b) const { return lo + b * Rp * Kp; }
};

// Split `batch` (R x K) views (element (r,k) of matrix b at
// src[b*bsrc + r*sr + k*sk]), times `scale`, into planes of Rp >= R rows.
void split_planes(Engine &e, const double *src, int64_t sr, int64_t sk, int R, int K, int Rp,
                  int batch, int64_t bsrc, double scale, B3Planes &out);

// Segmented split-BF16x3 GEMM on pre-split planes: C = alpha * sum_s A_s B_s
// + beta * C, A_s rows = output rows (at least the tile-padded M), B_s rows =
// output columns; each segment has its own planes and Kp.
struct B3Seg {
  const unsigned short *Ah = nullptr, *Al = nullptr, *Bh = nullptr, *Bl = nullptr;
  int Kp = 0;
};
struct B3SegArgs {
  GemmArgs g;
  int nseg = 0;
  B3Seg seg[kMaxSegs];
};
constexpr int kB3SegTile = 128;  // rows of A and B planes must cover the tile-padded M, N
void gemm_b3_segments(Engine &e, B3SegArgs a);

// Launch on the engine's stream.  precision: SISI_FP64/_FP32/_BF16X3/_BF16.
// Products too small for the bf16 split run in FP32 unless `exact` is set.
void gemm(Engine &eng, int precision, GemmArgs g, bool exact = false);

// FP64 / FP32 product on `stream` with the split-K partial sums in the
// caller's workspace (work_doubles doubles), no allocation.  ksplit = 0:
// the split gemm() would choose, reduced to what the workspace holds;
// ksplit > 0: that split, or an error when its partial sums do not fit.
void gemm_workspace(hipStream_t stream, int cus, int precision, GemmArgs g, double *work,
                    int64_t work_doubles, int ksplit);

// Split-BF16x3 GEMM on operands split into bf16 planes (batch == 1);
// passes = 1 keeps the hi*hi product only (plain bf16).
void gemm_b3(Engine &eng, GemmArgs g, int passes = 3);

// Backend of the batch-1 split-BF16x3 / bf16 products for the whole process:
// 0 = hand-written kernel (default), 1 = hipBLASLt on the split planes.
void set_b3_backend(int library);
int get_b3_backend();

// Kernel variants of the batch-1 split-BF16x3 (passes = 3) and bf16 (passes
// = 1) products: count, name, and a process-wide forced choice (-1: the
// tuning cache, else the size rule).
int b3_variant_count(int passes);
const char *b3_variant_name(int passes, int variant);
void set_b3_variant(int passes, int variant);

// Split-K factor from s (at most smax) chosen so that groups * split
// workgroups fill their last round on a device of `cus` compute units.
int round_filling_split(int64_t groups, int64_t s, int64_t smax, int cus);

// An A operand (M x K, element (m,k) at A[m*sam + k*sak]) split once into the
// bf16 hi and lo planes of the split-BF16x3 / bf16 kernels (rows padded to
// 256, so every kernel variant reads them) and kept
// across residual evaluations (the refinement anchor's ladder slices);
// `store` owns both planes.
// The planes may instead live in memory owned elsewhere (ext: the hi plane,
// lo_off: bf16 elements from hi to lo; 0 means Mp * Kp), e.g. one pair's
// rows inside a stack of several operands' planes.
struct B3FrozenA {
  DBuf store;
  int M = 0, K = 0, Mp = 0, Kp = 0;
  const void *ext = nullptr;
  int64_t lo_off = 0;
  const void *hi() const { return ext ? ext : static_cast<const void *>(store.get()); }
  bool valid() const { return hi() != nullptr; }
  int64_t lo_offset() const { return lo_off ? lo_off : int64_t(Mp) * Kp; }
};
void b3_freeze_a(Engine &e, const double *A, int64_t sam, int64_t sak, int M, int K,
                 B3FrozenA &out);
// Split A (M x K, element (m,k) at A[m*sam + k*sak]) into caller-owned planes:
// rows [0, Mp) of hi and lo with row stride Kp (rows past M and columns past K
// are written as zeros).
void b3_split_into(Engine &e, const double *A, int64_t sam, int64_t sak, int M, int K, int Mp,
                   int Kp, void *hi, void *lo);
// C = alpha * A B + beta * C with A taken from frozen planes (g.A unused; g.M
// and g.K must match them); passes = 1 (bf16) or 3 (split-BF16x3).
void gemm_b3_frozen_a(Engine &e, GemmArgs g, const B3FrozenA &A, int passes);

// Split-K reduction of g.work into C (deterministic order).
void gemm_reduce(Engine &eng, const GemmArgs &g);

// Products below this many flops run on FP32 MFMA when split-BF16x3 is asked
// for (the split's extra pass over the operands would dominate).
constexpr double kB3MinFlops = 1e10;

// Algorithmic flops of one call (2*M*N*K*batch).
inline double gemm_flops(const GemmArgs &g) {
  return 2.0 * double(g.M) * double(g.N) * double(g.K) * double(g.batch);
}

}  // namespace sisi
