// MFMA GEMM kernels for gfx950 (CDNA4).
//
// Block tile 128 x 128, 256 threads = 4 wave64s in a 2 x 2 arrangement, each
// wave owning a 64 x 64 sub-tile as 4 x 4 MFMA 16x16 tiles.  Operands are
// staged through LDS in k-chunks; the next chunk is fetched into registers
// while the current one is multiplied.
//
// MFMA operand roles: the instruction's "A" operand is fed with B(k,n) and
// its "B" operand with A(m,k), so the 16x16 result tile is C^T in MFMA
// terms: lane l holds C(m = l%16, n = 4*(l/16) + r), r = 0..3, for the f32
// and bf16 forms and C(m = l%16, n = l/16 + 4*r) for the f64 form.  Lanes then
// run along m, which is the unit-stride index of the column-major output,
// and each register's 16-lane group stores 128 contiguous bytes.
//
//   FP64   : v_mfma_f64_16x16x4_f64, FP64 accumulation.
//   FP32   : v_mfma_f32_16x16x4_f32 on operands rounded to FP32, FP32
//            accumulation; the epilogue and split-K reduction are FP64.
//   BF16X3 : v_mfma_f32_16x16x32_bf16 three times per tile,
//            a_hi*b_hi + a_hi*b_lo + a_lo*b_hi with x = hi + lo split of the
//            FP32-rounded operand, FP32 accumulation.
#include <cstdlib>

#include "engine.hpp"
#include "gemm.hpp"
#include "gemm_device.hpp"
#include "tape.hpp"

namespace sisi {

// ---------------------------------------------------------------------------
// FP64 / FP32 kernel (LDS element type LT; k-chunk 16; MFMA k-step 4)
// ---------------------------------------------------------------------------

template <typename LT>
__device__ __forceinline__ void mfma_k4(d4 &acc, double a, double b);

// Body of the FP64 / FP32 kernel for output tile `pid` (of tiles_m x
// tiles_n), k-split `ks` and batch entry `bz` (the eager kernel passes its
// block indices, the grouped kernel of the batched solver the indices of one
// record's blocks).
template <int PREC, bool AK, bool BNF>
__device__ __forceinline__ void gemm_k4_body(const GemmArgs &g, int pid, int ks, int bz) {
  using LT = typename std::conditional<PREC == SISI_FP64, double, float>::type;
  using AT = typename std::conditional<PREC == SISI_FP64, d4, f4>::type;
  constexpr int BK = 16;
  constexpr int PAD = (PREC == SISI_FP64) ? 2 : 4;
  // result layout of the 16x16 tile: row (here n) = LSTR*(lane/16) + RSTR*reg
  // (v_mfma_f64_16x16x4f64: lane/16 + 4*reg; the f32 form: 4*(lane/16) + reg)
  constexpr int LSTR = (PREC == SISI_FP64) ? 1 : 4;
  constexpr int RSTR = (PREC == SISI_FP64) ? 4 : 1;
  __shared__ LT As[BK][BM + PAD];
  __shared__ LT Bs[BK][BN + PAD];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar branches
  int tm, tn;
  tile_coords(g, pid, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int kbeg = ks * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);

  const double *A = g.A + int64_t(bz) * g.bsa;
  const double *B = g.B + int64_t(bz) * g.bsb;

  double ra[8], rb[8];

  // Out-of-range elements load from a clamped (in-range) address, so the
  // loads need no branch; they are zeroed when the chunk is written to LDS
  // (multiplying by 0.0, which the compiler cannot fold away), so the loads
  // stay in flight across the MFMAs of the previous chunk.
  const int mlast = g.M - 1, nlast = g.N - 1;
  int lk0 = 0;  // k0 of the chunk held in ra/rb
  auto load = [&](int k0) {
    const int klast = kend - 1;
    lk0 = k0;
    if (!AK) {  // A m-fastest: 32 threads x 4 consecutive m, k rows t/32 and +8
      const int mm = (tid & 31) * 4, kk = tid >> 5;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int k = min(k0 + kk + 8 * h, klast);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          ra[h * 4 + j] = A[int64_t(min(m0 + mm + j, mlast)) * g.sam + int64_t(k) * g.sak];
      }
    } else {  // A k-fastest: 16 threads along k, m = t/16 + 16 j
      const int kk = tid & 15, mm = tid >> 4;
      const int k = min(k0 + kk, klast);
#pragma unroll
      for (int j = 0; j < 8; ++j)
        ra[j] = A[int64_t(min(m0 + mm + 16 * j, mlast)) * g.sam + int64_t(k) * g.sak];
    }
    if (BNF) {  // B n-fastest
      const int nn = (tid & 31) * 4, kk = tid >> 5;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int k = min(k0 + kk + 8 * h, klast);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          rb[h * 4 + j] = B[int64_t(k) * g.sbk + int64_t(min(n0 + nn + j, nlast)) * g.sbn];
      }
    } else {  // B k-fastest
      const int kk = tid & 15, nn = tid >> 4;
      const int k = min(k0 + kk, klast);
#pragma unroll
      for (int j = 0; j < 8; ++j)
        rb[j] = B[int64_t(k) * g.sbk + int64_t(min(n0 + nn + 16 * j, nlast)) * g.sbn];
    }
  };

  auto store = [&]() {
    auto z = [](bool ok) { return ok ? 1.0 : 0.0; };
    if (!AK) {
      const int mm = (tid & 31) * 4, kk = tid >> 5;
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          As[kk + 8 * h][mm + j] =
              LT(ra[h * 4 + j] * z(m0 + mm + j <= mlast && lk0 + kk + 8 * h < kend));
    } else {
      const int kk = tid & 15, mm = tid >> 4;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        As[kk][mm + 16 * j] = LT(ra[j] * z(m0 + mm + 16 * j <= mlast && lk0 + kk < kend));
    }
    if (BNF) {
      const int nn = (tid & 31) * 4, kk = tid >> 5;
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          Bs[kk + 8 * h][nn + j] =
              LT(rb[h * 4 + j] * z(n0 + nn + j <= nlast && lk0 + kk + 8 * h < kend));
    } else {
      const int kk = tid & 15, nn = tid >> 4;
#pragma unroll
      for (int j = 0; j < 8; ++j)
        Bs[kk][nn + 16 * j] = LT(rb[j] * z(n0 + nn + 16 * j <= nlast && lk0 + kk < kend));
    }
  };

  AT acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = AT{0, 0, 0, 0};

  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int l16 = lane & 15, lk = lane >> 4;
  // a wave whose 64x64 sub-tile lies wholly outside the output (ragged edge
  // tiles) skips its MFMAs; the condition is uniform per wave
  const bool active = (m0 + wm < g.M) && (n0 + wn < g.N);
  auto mma = [&]() {
#pragma unroll
    for (int kq = 0; kq < BK / 4; ++kq) {
      const int k = kq * 4 + lk;
      LT av[4], bv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) av[i] = As[k][wm + i * 16 + l16];
#pragma unroll
      for (int j = 0; j < 4; ++j) bv[j] = Bs[k][wn + j * 16 + l16];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if constexpr (PREC == SISI_FP64)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j], av[i], acc[i][j], 0, 0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[j], av[i], acc[i][j], 0, 0, 0);
        }
    }
  };

  int k0 = kbeg;
  if (k0 < kend) load(k0);
  while (k0 < kend) {
    store();
    __syncthreads();
    const int k1 = k0 + BK;
    if (k1 < kend) load(k1);
    if (active) mma();
    __syncthreads();
    k0 = k1;
  }

  // epilogue
  if (g.ksplit == 1) {
    double *C = g.C + int64_t(bz) * g.bsc;
    // row offsets once per row, each column offset once per column (the
    // output splits cost divisions)
    int64_t om[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + wm + i * 16 + l16;
      om[i] = m < g.M ? out_off_m(g, m) : -1;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn + j * 16 + LSTR * lk + RSTR * r;
        if (n >= g.N) continue;
        const int64_t on = out_off_n(g, n);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (om[i] < 0) continue;
          const int64_t o = om[i] + on;
          double v = g.alpha * double(acc[i][j][r]);
          if (g.beta != 0.0) v += g.beta * C[o];
          C[o] = v;
        }
      }
  } else {
    double *W = g.work + (int64_t(bz) * g.ksplit + ks) * int64_t(g.M) * g.N;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + wm + i * 16 + l16;
      if (m >= g.M) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = n0 + wn + j * 16 + LSTR * lk + RSTR * r;
          if (n < g.N) W[m + int64_t(n) * g.M] = double(acc[i][j][r]);
        }
    }
  }
}

template <int PREC, bool AK, bool BNF>
__global__ __launch_bounds__(256, 2) void gemm_k4_kernel(GemmArgs g) {
  gemm_k4_body<PREC, AK, BNF>(g, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Grouped launch of n products (batched solver): record p owns the blocks
// [start[p], start[p + 1]): tiles_m * tiles_n per k-split and batch entry.
template <int PREC, bool AK, bool BNF>
__global__ __launch_bounds__(256, 2) void gemm_k4_grouped(const GemmArgs *const *args,
                                                          const unsigned *__restrict__ start,
                                                          int n) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  // the record is read in place (a copy kept the whole struct live in
  // registers and made the kernel spill)
  const GemmArgs &g = *args[lo];
  const int tiles = g.tiles_m * g.tiles_n;
  const int local = int(blockIdx.x - start[lo]);
  const int r = local / tiles;
  gemm_k4_body<PREC, AK, BNF>(g, local % tiles, r % g.ksplit, r / g.ksplit);
}

// ---------------------------------------------------------------------------
// BF16x3 kernel (k-chunk 32 = one MFMA k-step; LDS rows along k)
// ---------------------------------------------------------------------------

template <bool AK, bool BNF>
__global__ __launch_bounds__(256, 2) void gemm_bf16x3_kernel(GemmArgs g) {
  constexpr int BK = 32;
  constexpr int RS = BK + 8;  // row stride in bf16 (keeps 16-byte alignment)
  __shared__ __attribute__((aligned(16))) __bf16 Ah[BM * RS];
  __shared__ __attribute__((aligned(16))) __bf16 Al[BM * RS];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[BN * RS];
  __shared__ __attribute__((aligned(16))) __bf16 Bl[BN * RS];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar branches
  int tm, tn;
  tile_coords(g, blockIdx.x, tm, tn);
  const int ks = blockIdx.y;
  const int bz = blockIdx.z;
  const int m0 = tm * BM, n0 = tn * BN;
  const int kbeg = ks * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  const double *A = g.A + int64_t(bz) * g.bsa;
  const double *B = g.B + int64_t(bz) * g.bsb;

  // 128 x 32 elements per operand = 16 per thread
  double ra[16], rb[16];
  // Out-of-range elements load from a clamped (in-range) address (no
  // branch) and are zeroed at the LDS store, one chunk later.
  const int mlast = g.M - 1, nlast = g.N - 1;
  int lk0 = 0;
  auto load = [&](int k0) {
    const int klast = kend - 1;
    lk0 = k0;
    if (!AK) {  // m fastest: 32 threads x 4 m, 8 k rows, 4 passes
      const int mm = (tid & 31) * 4, kk = tid >> 5;
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int k = k0 + kk + 8 * h;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int m = m0 + mm + j;
          ra[h * 4 + j] = A[int64_t(min(m, mlast)) * g.sam + int64_t(min(k, klast)) * g.sak];
        }
      }
    } else {  // k fastest: 32 threads along k, m = t/32 + 8 j
      const int kk = tid & 31, mm = tid >> 5;
      const int k = k0 + kk;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int m = m0 + mm + 8 * j;
        ra[j] = A[int64_t(min(m, mlast)) * g.sam + int64_t(min(k, klast)) * g.sak];
      }
    }
    if (BNF) {
      const int nn = (tid & 31) * 4, kk = tid >> 5;
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int k = k0 + kk + 8 * h;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int n = n0 + nn + j;
          rb[h * 4 + j] = B[int64_t(min(k, klast)) * g.sbk + int64_t(min(n, nlast)) * g.sbn];
        }
      }
    } else {
      const int kk = tid & 31, nn = tid >> 5;
      const int k = k0 + kk;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int n = n0 + nn + 8 * j;
        rb[j] = B[int64_t(min(k, klast)) * g.sbk + int64_t(min(n, nlast)) * g.sbn];
      }
    }
  };
  auto store = [&]() {
    __bf16 h, l;
    if (!AK) {
      const int mm = (tid & 31) * 4, kk = tid >> 5;
#pragma unroll
      for (int hh = 0; hh < 4; ++hh)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          split_bf16(ra[hh * 4 + j] * ((m0 + mm + j <= mlast && lk0 + kk + 8 * hh < kend) ? 1.0 : 0.0), h, l);
          Ah[(mm + j) * RS + kk + 8 * hh] = h;
          Al[(mm + j) * RS + kk + 8 * hh] = l;
        }
    } else {
      const int kk = tid & 31, mm = tid >> 5;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        split_bf16(ra[j] * ((m0 + mm + 8 * j <= mlast && lk0 + kk < kend) ? 1.0 : 0.0), h, l);
        Ah[(mm + 8 * j) * RS + kk] = h;
        Al[(mm + 8 * j) * RS + kk] = l;
      }
    }
    if (BNF) {
      const int nn = (tid & 31) * 4, kk = tid >> 5;
#pragma unroll
      for (int hh = 0; hh < 4; ++hh)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          split_bf16(rb[hh * 4 + j] * ((n0 + nn + j <= nlast && lk0 + kk + 8 * hh < kend) ? 1.0 : 0.0), h, l);
          Bh[(nn + j) * RS + kk + 8 * hh] = h;
          Bl[(nn + j) * RS + kk + 8 * hh] = l;
        }
    } else {
      const int kk = tid & 31, nn = tid >> 5;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        split_bf16(rb[j] * ((n0 + nn + 8 * j <= nlast && lk0 + kk < kend) ? 1.0 : 0.0), h, l);
        Bh[(nn + 8 * j) * RS + kk] = h;
        Bl[(nn + 8 * j) * RS + kk] = l;
      }
    }
  };

  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0, 0, 0, 0};
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int l16 = lane & 15, lk = lane >> 4;

  int k0 = kbeg;
  if (k0 < kend) load(k0);
  while (k0 < kend) {
    store();
    __syncthreads();
    const int k1 = k0 + BK;
    if (k1 < kend) load(k1);
    bf8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = wm + i * 16 + l16;
      ah[i] = *reinterpret_cast<const bf8 *>(&Ah[row * RS + 8 * lk]);
      al[i] = *reinterpret_cast<const bf8 *>(&Al[row * RS + 8 * lk]);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = wn + j * 16 + l16;
      bh[j] = *reinterpret_cast<const bf8 *>(&Bh[col * RS + 8 * lk]);
      bl[j] = *reinterpret_cast<const bf8 *>(&Bl[col * RS + 8 * lk]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah[i], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al[i], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah[i], acc[i][j], 0, 0, 0);
    __syncthreads();
    k0 = k1;
  }

  if (g.ksplit == 1) {
    double *C = g.C + int64_t(bz) * g.bsc;
    // row offsets once per row, each column offset once per column (the
    // output splits cost divisions)
    int64_t om[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + wm + i * 16 + l16;
      om[i] = m < g.M ? out_off_m(g, m) : -1;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn + j * 16 + 4 * lk + r;
        if (n >= g.N) continue;
        const int64_t on = out_off_n(g, n);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          if (om[i] < 0) continue;
          const int64_t o = om[i] + on;
          double v = g.alpha * double(acc[i][j][r]);
          if (g.beta != 0.0) v += g.beta * C[o];
          C[o] = v;
        }
      }
  } else {
    double *W = g.work + (int64_t(bz) * g.ksplit + ks) * int64_t(g.M) * g.N;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int m = m0 + wm + i * 16 + l16;
      if (m >= g.M) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = n0 + wn + j * 16 + 4 * lk + r;
          if (n < g.N) W[m + int64_t(n) * g.M] = double(acc[i][j][r]);
        }
    }
  }
}

// Split-K reduction: C(m,n) = alpha * sum_s W[s](m,n) + beta * C(m,n).  A
// block takes 64 consecutive outputs; its four waves sum the partials
// s = q, q + 4, ... (q = wave) and the four sums are added in a fixed order,
// so the result does not depend on timing.
// Split-K reduction of batch entry bz, workgroup bx of gx: 64 outputs per
// step, the four waves summing every fourth partial, the four sums added in
// a fixed order (deterministic).
__device__ __forceinline__ void gemm_reduce_body(const GemmArgs &g, unsigned bx, unsigned gx,
                                                 int bz) {
  const int64_t MN = int64_t(g.M) * g.N;
  double *C = g.C + int64_t(bz) * g.bsc;
  const double *W = g.work + int64_t(bz) * g.ksplit * MN;
  __shared__ double part[4][64];
  const int t = threadIdx.x & 63, q = threadIdx.x >> 6;
  for (int64_t e0 = int64_t(bx) * 64; e0 < MN; e0 += int64_t(gx) * 64) {
    const int64_t e = e0 + t;
    double s = 0.0;
    if (e < MN) {
#pragma unroll 4
      for (int k = q; k < g.ksplit; k += 4) s += W[k * MN + e];
    }
    part[q][t] = s;
    __syncthreads();
    if (q == 0 && e < MN) {
      const int m = int(e % g.M);
      const int n = int(e / g.M);
      const double sum = (part[0][t] + part[1][t]) + (part[2][t] + part[3][t]);
      const int64_t o = out_off_m(g, m) + out_off_n(g, n);
      double v = g.alpha * sum;
      if (g.beta != 0.0) v += g.beta * C[o];
      C[o] = v;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void gemm_reduce_kernel(GemmArgs g) {
  gemm_reduce_body(g, blockIdx.x, gridDim.x, blockIdx.y);
}

// Grouped split-K reductions: record p owns gx * batch blocks (gx = the
// record's eager grid x).
__global__ __launch_bounds__(256) void gemm_reduce_grouped(const GemmArgs *const *args,
                                                           const unsigned *__restrict__ start,
                                                           const unsigned *__restrict__ gxs,
                                                           int n) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const GemmArgs g = *args[lo];
  const unsigned local = blockIdx.x - start[lo], gx = gxs[lo];
  gemm_reduce_body(g, local % gx, gx, int(local / gx));
}

// ---------------------------------------------------------------------------
// Launcher
// ---------------------------------------------------------------------------

template <int PREC>
static void launch_k4(GemmArgs &g, dim3 grid, hipStream_t s, bool AK, bool BNF) {
  if (!AK && !BNF) hipLaunchKernelGGL((gemm_k4_kernel<PREC, false, false>), grid, dim3(256), 0, s, g);
  else if (!AK && BNF) hipLaunchKernelGGL((gemm_k4_kernel<PREC, false, true>), grid, dim3(256), 0, s, g);
  else if (AK && !BNF) hipLaunchKernelGGL((gemm_k4_kernel<PREC, true, false>), grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL((gemm_k4_kernel<PREC, true, true>), grid, dim3(256), 0, s, g);
}

static void launch_b3(GemmArgs &g, dim3 grid, hipStream_t s, bool AK, bool BNF) {
  if (!AK && !BNF) hipLaunchKernelGGL((gemm_bf16x3_kernel<false, false>), grid, dim3(256), 0, s, g);
  else if (!AK && BNF) hipLaunchKernelGGL((gemm_bf16x3_kernel<false, true>), grid, dim3(256), 0, s, g);
  else if (AK && !BNF) hipLaunchKernelGGL((gemm_bf16x3_kernel<true, false>), grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL((gemm_bf16x3_kernel<true, true>), grid, dim3(256), 0, s, g);
}

static void launch_reduce(const GemmArgs &g, hipStream_t s);

void gemm_reduce(Engine &eng, const GemmArgs &g) {
  Timed t(*eng.timers, Kind::Bandwidth, 8.0 * (g.ksplit + 2) * double(g.M) * g.N * g.batch);
  launch_reduce(g, eng.stream);
}

int round_filling_split(int64_t groups, int64_t s, int64_t smax, int cus) {
  // workgroups of equal length run in rounds of 2 per CU: from s up to 4 s,
  // the smallest split whose last round is at least 95 % full (else the
  // fullest one found)
  const int64_t slots = 2 * int64_t(cus);
  auto fill = [&](int64_t ks) {
    const int64_t blocks = groups * ks;
    return double(blocks) / double(((blocks + slots - 1) / slots) * slots);
  };
  if (s < 1) s = 1;
  if (groups * s < slots) return int(s);
  int64_t best = s;
  for (int64_t ks = s; ks <= std::min(4 * s, smax) && fill(best) < 0.95; ++ks)
    if (fill(ks) > fill(best) + 0.02) best = ks;
  return int(best);
}

static int64_t env_int(const char *name, int64_t fallback) {
  const char *v = getenv(name);
  return v != nullptr && *v != 0 ? int64_t(atoll(v)) : fallback;
}

static void set_ksplit(GemmArgs &g, int ksplit, int kq);

// Tile grid, load mappings and split-K plan of one FP64 / FP32 / batched
// split-BF16x3 product (fields of g filled in); returns the load mappings.
static void plan_k4(GemmArgs &g, int precision, int cus, bool &AKf, bool &BNFf) {
  if (g.K < 0) g.K = 0;  // K == 0 gives C = beta * C through the epilogue
  SISI_REQUIRE(g.sam == 1 || g.sak == 1 || g.K <= 1 || g.M <= 1, SISI_ERR_INVALID,
               "gemm: A needs a unit stride");
  SISI_REQUIRE(g.sbk == 1 || g.sbn == 1 || g.K <= 1 || g.N <= 1, SISI_ERR_INVALID,
               "gemm: B needs a unit stride");
  // The address formula is the same for every load mapping; the mapping only
  // decides which index consecutive threads walk (coalescing).
  AKf = (g.sak == 1) && (g.sam != 1);
  BNFf = (g.sbn == 1) && (g.sbk != 1);
  g.tiles_m = (g.M + BM - 1) / BM;
  g.tiles_n = (g.N + BN - 1) / BN;
  const int64_t tiles = int64_t(g.tiles_m) * g.tiles_n;
  SISI_REQUIRE(tiles < (int64_t(1) << 31), SISI_ERR_INVALID, "gemm: too many tiles");
  g.split_out = !(g.dm0 >= g.M && g.dn0 >= g.N);
  {
    static const int env_map = getenv("SISI_GEMM_XCD") ? atoi(getenv("SISI_GEMM_XCD")) : 1;
    static const int env_group = getenv("SISI_GEMM_GROUP") ? atoi(getenv("SISI_GEMM_GROUP")) : 8;
    g.xcd_map = env_map;
    g.group = env_group > 0 ? env_group : 8;
  }
  const int kq = (precision == SISI_BF16X3) ? 32 : 16;
  // split-K when the tile grid alone cannot fill the 256 CUs: about `want`
  // workgroups, each with at least `minsteps` k-chunks, at most `maxsplit`
  // partial sums
  static const int64_t want = env_int("SISI_GEMM_SPLIT_WANT", 1024);
  static const int64_t minsteps = std::max<int64_t>(1, env_int("SISI_GEMM_SPLIT_MINSTEPS", 8));
  static const int64_t maxsplit = std::max<int64_t>(1, env_int("SISI_GEMM_SPLIT_MAX", 64));
  int ksplit = 1;
  if (g.K > 4 * kq && tiles * g.batch < want) {
    const int64_t tb = tiles * g.batch;
    int64_t s = (want + tb - 1) / tb;
    const int64_t smax = std::min<int64_t>(std::max<int64_t>(1, g.K / (minsteps * kq)), maxsplit);
    ksplit = round_filling_split(tb, std::min(s, smax), smax, cus);
  }
  set_ksplit(g, ksplit, kq);
}

// Chunking of K for a split of (at most) ksplit partial sums.
static void set_ksplit(GemmArgs &g, int ksplit, int kq) {
  int kchunk = (g.K + ksplit - 1) / ksplit;
  kchunk = ((kchunk + kq - 1) / kq) * kq;
  if (kchunk == 0) kchunk = kq;
  ksplit = (g.K + kchunk - 1) / kchunk;
  if (ksplit < 1) ksplit = 1;
  g.ksplit = ksplit;
  g.kchunk = kchunk;
}

// Launch of a planned product; the split-K partial sums go to g.work.
static void launch_planned(GemmArgs &g, int precision, bool AKf, bool BNFf, hipStream_t s) {
  SISI_REQUIRE(g.batch <= 65535 && g.ksplit <= 65535, SISI_ERR_INVALID, "gemm: grid too large");
  dim3 grid(unsigned(int64_t(g.tiles_m) * g.tiles_n), unsigned(g.ksplit), unsigned(g.batch));
  if (precision == SISI_FP64) launch_k4<SISI_FP64>(g, grid, s, AKf, BNFf);
  else if (precision == SISI_FP32) launch_k4<SISI_FP32>(g, grid, s, AKf, BNFf);
  else launch_b3(g, grid, s, AKf, BNFf);
  SISI_HIP(hipGetLastError());
}

static void launch_reduce(const GemmArgs &g, hipStream_t s) {
  const int64_t MN = int64_t(g.M) * g.N;
  int blocks = int(std::min<int64_t>((MN + 63) / 64, 8192));
  hipLaunchKernelGGL(gemm_reduce_kernel, dim3(blocks, g.batch), dim3(256), 0, s, g);
  SISI_HIP(hipGetLastError());
}

// Sub-kind of a taped product: precision (0 FP64, 1 FP32) * 4 + AK * 2 + BNF.
static int tape_gemm_sub(int precision, bool AKf, bool BNFf) {
  return (precision == SISI_FP64 ? 0 : 4) + (AKf ? 2 : 0) + (BNFf ? 1 : 0);
}

void launch_gemm_grouped(int sub, const GemmArgs *const *args, const unsigned *start, int n,
                         unsigned total, hipStream_t s) {
  switch (sub) {
#define SISI_GROUPED_CASE(S, P, A, B)                                                            \
  case S:                                                                                      \
    hipLaunchKernelGGL((gemm_k4_grouped<P, A, B>), dim3(total), dim3(256), 0, s, args, start, n); \
    break;
    SISI_GROUPED_CASE(0, SISI_FP64, false, false)
    SISI_GROUPED_CASE(1, SISI_FP64, false, true)
    SISI_GROUPED_CASE(2, SISI_FP64, true, false)
    SISI_GROUPED_CASE(3, SISI_FP64, true, true)
    SISI_GROUPED_CASE(4, SISI_FP32, false, false)
    SISI_GROUPED_CASE(5, SISI_FP32, false, true)
    SISI_GROUPED_CASE(6, SISI_FP32, true, false)
    SISI_GROUPED_CASE(7, SISI_FP32, true, true)
#undef SISI_GROUPED_CASE
    default: throw Error(SISI_ERR_INVALID, "grouped gemm: unknown sub-kind");
  }
  SISI_HIP(hipGetLastError());
}

void launch_gemm_reduce_grouped(const GemmArgs *const *args, const unsigned *start,
                                const unsigned *gxs, int n, unsigned total, hipStream_t s) {
  hipLaunchKernelGGL(gemm_reduce_grouped, dim3(total), dim3(256), 0, s, args, start, gxs, n);
  SISI_HIP(hipGetLastError());
}

// A product on a taped engine, recorded for the grouped launch, with the
// reduced precisions on FP32 MFMA.  The systems of a grouped launch fill the
// device together, so only a product with very few output tiles and a long
// K is split along K: into parts of about 64 k-chunks (at most 512 / tiles
// parts, workspace of at most 2^24 doubles), whose partial sums a recorded
// reduction adds in a fixed order.  A product with a small output and a
// long K (Kki of the residual: 16 x 16 outputs over K = Nv^2 No) would
// otherwise run its whole K in one workgroup.
static void tape_gemm(Engine &eng, int precision, GemmArgs g) {
  if (precision != SISI_FP64) precision = SISI_FP32;
  bool AKf, BNFf;
  plan_k4(g, precision, eng.cus, AKf, BNFf);
  const int64_t tb = int64_t(g.tiles_m) * g.tiles_n * g.batch;
  const int64_t MNb = int64_t(g.M) * g.N * g.batch;
  int64_t ks = 1;
  if (tb < 32 && g.K >= 2048) {
    ks = std::min<int64_t>((g.K + 1023) / 1024, std::max<int64_t>(1, 512 / tb));
    ks = std::min<int64_t>(ks, std::max<int64_t>(1, (int64_t(1) << 24) / MNb));
  }
  set_ksplit(g, int(ks), 16);
  SISI_REQUIRE(tb * g.ksplit < (int64_t(1) << 31), SISI_ERR_INVALID, "gemm: too many tiles");
  DBuf work;
  if (g.ksplit > 1) {
    work = DBuf(&eng.pool, size_t(g.ksplit) * size_t(MNb));
    g.work = work.get();
  }
  TapeOp op;
  op.kind = OpKind::Gemm;
  op.sub = tape_gemm_sub(precision, AKf, BNFf);
  op.gx = unsigned(g.tiles_m * g.tiles_n);
  op.gy = unsigned(g.ksplit);
  op.gz = unsigned(g.batch);
  op.g = g;
  eng.tape->ops.push_back(op);
  if (g.ksplit > 1) {
    TapeOp red;
    red.kind = OpKind::GemmReduce;
    red.gx = unsigned(std::min<int64_t>((int64_t(g.M) * g.N + 63) / 64, 1024));
    red.gz = unsigned(g.batch);
    red.g = g;
    eng.tape->ops.push_back(red);
  }
}

void gemm(Engine &eng, int precision, GemmArgs g, bool exact) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return;
  if (eng.tape) return tape_gemm(eng, precision, g);
  // split-BF16x3 splits both operands first (a bandwidth pass over each); a
  // product too small to amortise that runs on FP32 MFMA instead
  if ((precision == SISI_BF16X3 || precision == SISI_BF16) && !exact &&
      gemm_flops(g) < kB3MinFlops)
    precision = SISI_FP32;
  if ((precision == SISI_BF16X3 || precision == SISI_BF16) && g.batch == 1 && g.K > 0) {
    static const int env_b3 = getenv("SISI_B3_PRESPLIT") ? atoi(getenv("SISI_B3_PRESPLIT")) : 1;
    if (env_b3) {
      gemm_b3(eng, g, precision == SISI_BF16 ? 1 : 3);
      return;
    }
  }
  if (precision == SISI_BF16) precision = SISI_BF16X3;  // batched: the three-product kernel
  bool AKf, BNFf;
  plan_k4(g, precision, eng.cus, AKf, BNFf);
  DBuf work;
  if (g.ksplit > 1) {
    work = DBuf(&eng.pool, size_t(g.ksplit) * g.batch * g.M * g.N);
    g.work = work.get();
  }
  const Kind kind = precision == SISI_FP64 ? Kind::Gemm64
                    : precision == SISI_FP32 ? Kind::Gemm32 : Kind::GemmB3;
  {
    Timed t(*eng.timers, kind, gemm_flops(g));
    launch_planned(g, precision, AKf, BNFf, eng.stream);
  }
  if (g.ksplit > 1) gemm_reduce(eng, g);
}

void gemm_workspace(hipStream_t stream, int cus, int precision, GemmArgs g, double *work,
                    int64_t work_doubles, int ksplit) {
  SISI_REQUIRE(precision == SISI_FP64 || precision == SISI_FP32, SISI_ERR_INVALID,
               "gemm_workspace: FP64 or FP32 only");
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return;
  SISI_REQUIRE(ksplit >= 0, SISI_ERR_INVALID, "gemm_workspace: negative split");
  bool AKf, BNFf;
  plan_k4(g, precision, cus, AKf, BNFf);
  const int64_t per = int64_t(g.batch) * g.M * g.N;  // doubles per partial sum
  if (ksplit > 0) {
    set_ksplit(g, ksplit, 16);
    SISI_REQUIRE(g.ksplit == 1 || int64_t(g.ksplit) * per <= work_doubles, SISI_ERR_INVALID,
                 "gemm_workspace: " + std::to_string(g.ksplit) + " partial sums of " +
                     std::to_string(per) + " doubles exceed the workspace of " +
                     std::to_string(work_doubles) + " doubles");
  } else if (g.ksplit > 1 && int64_t(g.ksplit) * per > work_doubles) {
    set_ksplit(g, int(std::max<int64_t>(1, work_doubles / per)), 16);  // fewer partial sums
  }
  if (g.ksplit > 1) g.work = work;
  launch_planned(g, precision, AKf, BNFf, stream);
  if (g.ksplit > 1) launch_reduce(g, stream);
}

}  // namespace sisi
