// Segmented MFMA GEMM: C = alpha * sum_s scale_s A_s B_s + beta * C with up
// to four products of different K, operands and layouts accumulated in the
// same registers, so a sum of contractions into one output costs one launch
// and one epilogue pass instead of one read-modify-write of C per product.
//
// Same block tile, wave layout, LDS staging and MFMA forms as the FP64/FP32
// and on-the-fly BF16x3 kernels of gemm.hip; the operand load mapping
// (m- or k-fastest A, k- or n-fastest B) is chosen per segment at run time
// (uniform across the workgroup).  The k loop walks the segments' k-chunks
// in order, prefetching the next chunk (possibly of the next segment) while
// the current one is multiplied.
#include <cstdlib>

#include "engine.hpp"
#include "gemm.hpp"
#include "gemm_device.hpp"
#include "tune.hpp"

namespace sisi {

namespace {

__device__ __forceinline__ GemmSeg seg_at(const SegGemmArgs &a, int s) {
  return s == 0 ? a.seg[0] : s == 1 ? a.seg[1] : s == 2 ? a.seg[2] : a.seg[3];
}

__device__ __forceinline__ int seg_k(const SegGemmArgs &a, int s) {
  return s == 0 ? a.seg[0].K : s == 1 ? a.seg[1].K : s == 2 ? a.seg[2].K : a.seg[3].K;
}

// TBM x TBN block tile of WGM x WGN waves (each (TBM/WGM) x (TBN/WGN), a
// multiple of 16 in both directions: 128 x 128 as 2 x 2 waves of 64 x 64,
// and 80 x 128 / 128 x 80 as 1 x 4 / 4 x 1 waves of 80 x 32 / 32 x 80 for
// an output dimension of a multiple of 80 that 128-wide tiles would leave
// ragged, such as the 400 virtuals of the (T) blocks).
template <int PREC, int TBM, int TBN, int WGM, int WGN>
__global__ __launch_bounds__(256, 2) void gemm_seg_k4_kernel(SegGemmArgs a) {
  using LT = typename std::conditional<PREC == SISI_FP64, double, float>::type;
  using AT = typename std::conditional<PREC == SISI_FP64, d4, f4>::type;
  constexpr int BK = 16;
  constexpr int PAD = (PREC == SISI_FP64) ? 2 : 4;
  constexpr int LSTR = (PREC == SISI_FP64) ? 1 : 4;
  constexpr int RSTR = (PREC == SISI_FP64) ? 4 : 1;
  constexpr int WM = TBM / WGM, WN = TBN / WGN, TI = WM / 16, TJ = WN / 16;
  constexpr int NA = TBM * BK / 256, NB = TBN * BK / 256;  // loads per thread
  static_assert(WGM * WGN == 4 && TI * 16 == WM && TJ * 16 == WN, "four waves, 16-multiples");
  static_assert(NA * 256 == TBM * BK && NB * 256 == TBN * BK, "whole loads per thread");
  __shared__ LT As[BK][TBM + PAD];
  __shared__ LT Bs[BK][TBN + PAD];
  const GemmArgs &g = a.g;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar branches
  int tm, tn;
  tile_coords(g, blockIdx.x, tm, tn);
  const int m0 = tm * TBM, n0 = tn * TBN;

  double ra[NA], rb[NB];
  int lak = 0, lbnf = 0;
  // Out-of-range elements load from a clamped (in-range) address, so the
  // loads need no branch; at the LDS store they are multiplied by zero (the
  // segment's scale otherwise), one chunk later, so the loads stay in flight
  // across the MFMAs of the previous chunk.  Element q of a thread is
  // e = tid + 256 q of the tile, walked along the operand's unit-stride
  // index (m or n fastest: e % TB, e / TB; k fastest: e / 16, e % 16).
  const int mlast = g.M - 1, nlast = g.N - 1;
  int lk0 = 0, lK = 0;
  double lsc = 1.0;

  auto load = [&](int s, int k0) {
    const GemmSeg sg = seg_at(a, s);
    const int klast = sg.K - 1;
    lak = sg.ak;
    lbnf = sg.bnf;
    lk0 = k0;
    lK = sg.K;
    lsc = sg.scale;
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int e = tid + 256 * q;
      const int mm = sg.ak ? e / BK : e % TBM, kk = sg.ak ? e % BK : e / TBM;
      ra[q] = sg.A[int64_t(min(m0 + mm, mlast)) * sg.sam + int64_t(min(k0 + kk, klast)) * sg.sak];
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int e = tid + 256 * q;
      const int nn = sg.bnf ? e % TBN : e / BK, kk = sg.bnf ? e / TBN : e % BK;
      rb[q] = sg.B[int64_t(min(k0 + kk, klast)) * sg.sbk + int64_t(min(n0 + nn, nlast)) * sg.sbn];
    }
  };

  auto store = [&]() {
#pragma unroll
    for (int q = 0; q < NA; ++q) {
      const int e = tid + 256 * q;
      const int mm = lak ? e / BK : e % TBM, kk = lak ? e % BK : e / TBM;
      As[kk][mm] = LT(ra[q] * ((m0 + mm <= mlast && lk0 + kk < lK) ? lsc : 0.0));
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const int e = tid + 256 * q;
      const int nn = lbnf ? e % TBN : e / BK, kk = lbnf ? e / TBN : e % BK;
      Bs[kk][nn] = LT(rb[q] * ((n0 + nn <= nlast && lk0 + kk < lK) ? 1.0 : 0.0));
    }
  };

  AT acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = AT{0, 0, 0, 0};

  const int wm = (wave % WGM) * WM, wn = (wave / WGM) * WN;
  const int l16 = lane & 15, lk = lane >> 4;
  // a wave whose sub-tile lies wholly outside the output skips its MFMAs
  const bool active = (m0 + wm < g.M) && (n0 + wn < g.N);

  int s = 0, k0 = 0;
  bool have = a.nseg > 0;
  if (have) load(0, 0);
  while (have) {
    store();
    __syncthreads();
    int s1 = s, k1 = k0 + BK;
    if (k1 >= seg_k(a, s)) {
      s1 = s + 1;
      k1 = 0;
    }
    const bool more = s1 < a.nseg;
    if (more) load(s1, k1);
    if (active) {
#pragma unroll
      for (int kq = 0; kq < BK / 4; ++kq) {
        const int k = kq * 4 + lk;
        LT av[TI], bv[TJ];
#pragma unroll
        for (int i = 0; i < TI; ++i) av[i] = As[k][wm + i * 16 + l16];
#pragma unroll
        for (int j = 0; j < TJ; ++j) bv[j] = Bs[k][wn + j * 16 + l16];
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
          for (int j = 0; j < TJ; ++j) {
            if constexpr (PREC == SISI_FP64)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[j], av[i], acc[i][j], 0, 0, 0);
            else
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[j], av[i], acc[i][j], 0, 0, 0);
          }
      }
    }
    __syncthreads();
    s = s1;
    k0 = k1;
    have = more;
  }

  // row offsets once per row, each column offset once per column (the
  // output splits cost divisions)
  int64_t om[TI];
#pragma unroll
  for (int i = 0; i < TI; ++i) {
    const int m = m0 + wm + i * 16 + l16;
    om[i] = m < g.M ? out_off_m(g, m) : -1;
  }
#pragma unroll
  for (int j = 0; j < TJ; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = n0 + wn + j * 16 + LSTR * lk + RSTR * r;
      if (n >= g.N) continue;
      const int64_t on = out_off_n(g, n);
#pragma unroll
      for (int i = 0; i < TI; ++i) {
        if (om[i] < 0) continue;
        const int64_t o = om[i] + on;
        double v = g.alpha * double(acc[i][j][r]);
        if (g.beta != 0.0) v += g.beta * g.C[o];
        g.C[o] = v;
      }
    }
}

__global__ __launch_bounds__(256, 2) void gemm_seg_b3_kernel(SegGemmArgs a) {
  constexpr int BK = 32;
  constexpr int RS = BK + 8;  // row stride in bf16 (keeps 16-byte alignment)
  __shared__ __attribute__((aligned(16))) __bf16 Ah[BM * RS];
  __shared__ __attribute__((aligned(16))) __bf16 Al[BM * RS];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[BN * RS];
  __shared__ __attribute__((aligned(16))) __bf16 Bl[BN * RS];
  const GemmArgs &g = a.g;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar branches
  int tm, tn;
  tile_coords(g, blockIdx.x, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;

  double ra[16], rb[16];
  int lak = 0, lbnf = 0;
  // Out-of-range elements load from a clamped (in-range) address (no
  // branch) and are zeroed at the LDS store, one chunk later.
  const int mlast = g.M - 1, nlast = g.N - 1;
  int lk0 = 0, lK = 0;
  double lsc = 1.0;
  auto load = [&](int s, int k0) {
    const GemmSeg sg = seg_at(a, s);
    const int klast = sg.K - 1;
    lak = sg.ak;
    lbnf = sg.bnf;
    lk0 = k0;
    lK = sg.K;
    lsc = sg.scale;
    if (!sg.ak) {  // m fastest: 32 threads x 4 m, 8 k rows, 4 passes
      const int mm = (tid & 31) * 4, kk = tid >> 5;
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int k = k0 + kk + 8 * h;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int m = m0 + mm + j;
          ra[h * 4 + j] = sg.A[int64_t(min(m, mlast)) * sg.sam + int64_t(min(k, klast)) * sg.sak];
        }
      }
    } else {  // k fastest: 32 threads along k, m = t/32 + 8 j
      const int kk = tid & 31, mm = tid >> 5;
      const int k = k0 + kk;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int m = m0 + mm + 8 * j;
        ra[j] = sg.A[int64_t(min(m, mlast)) * sg.sam + int64_t(min(k, klast)) * sg.sak];
      }
    }
    if (sg.bnf) {
      const int nn = (tid & 31) * 4, kk = tid >> 5;
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int k = k0 + kk + 8 * h;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int n = n0 + nn + j;
          rb[h * 4 + j] = sg.B[int64_t(min(k, klast)) * sg.sbk + int64_t(min(n, nlast)) * sg.sbn];
        }
      }
    } else {
      const int kk = tid & 31, nn = tid >> 5;
      const int k = k0 + kk;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int n = n0 + nn + 8 * j;
        rb[j] = sg.B[int64_t(min(k, klast)) * sg.sbk + int64_t(min(n, nlast)) * sg.sbn];
      }
    }
  };
  auto store = [&]() {
    __bf16 h, l;
    if (!lak) {
      const int mm = (tid & 31) * 4, kk = tid >> 5;
#pragma unroll
      for (int hh = 0; hh < 4; ++hh)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          split_bf16(ra[hh * 4 + j] * ((m0 + mm + j <= mlast && lk0 + kk + 8 * hh < lK) ? lsc : 0.0), h, l);
          Ah[(mm + j) * RS + kk + 8 * hh] = h;
          Al[(mm + j) * RS + kk + 8 * hh] = l;
        }
    } else {
      const int kk = tid & 31, mm = tid >> 5;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        split_bf16(ra[j] * ((m0 + mm + 8 * j <= mlast && lk0 + kk < lK) ? lsc : 0.0), h, l);
        Ah[(mm + 8 * j) * RS + kk] = h;
        Al[(mm + 8 * j) * RS + kk] = l;
      }
    }
    if (lbnf) {
      const int nn = (tid & 31) * 4, kk = tid >> 5;
#pragma unroll
      for (int hh = 0; hh < 4; ++hh)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          split_bf16(rb[hh * 4 + j] * ((n0 + nn + j <= nlast && lk0 + kk + 8 * hh < lK) ? 1.0 : 0.0), h, l);
          Bh[(nn + j) * RS + kk + 8 * hh] = h;
          Bl[(nn + j) * RS + kk + 8 * hh] = l;
        }
    } else {
      const int kk = tid & 31, nn = tid >> 5;
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        split_bf16(rb[j] * ((n0 + nn + 8 * j <= nlast && lk0 + kk < lK) ? 1.0 : 0.0), h, l);
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

  int s = 0, k0 = 0;
  bool have = a.nseg > 0;
  if (have) load(0, 0);
  while (have) {
    store();
    __syncthreads();
    int s1 = s, k1 = k0 + BK;
    if (k1 >= seg_k(a, s)) {
      s1 = s + 1;
      k1 = 0;
    }
    const bool more = s1 < a.nseg;
    if (more) load(s1, k1);
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
    s = s1;
    k0 = k1;
    have = more;
  }

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
        if (g.beta != 0.0) v += g.beta * g.C[o];
        g.C[o] = v;
      }
    }
}

}  // namespace

// Tile variants of the FP64 / FP32 segmented kernel (ids of the tuning
// cache's family 64 / 32): 128 x 128, 80 x 128, 128 x 80.
struct SegVariant {
  int bm, bn;
  void (*launch)(const SegGemmArgs &, dim3, hipStream_t);
  const char *name;
};

template <int PREC, int TBM, int TBN, int WGM, int WGN>
static void launch_seg(const SegGemmArgs &a, dim3 grid, hipStream_t s) {
  hipLaunchKernelGGL((gemm_seg_k4_kernel<PREC, TBM, TBN, WGM, WGN>), grid, dim3(256), 0, s, a);
}

static const SegVariant kSeg64[] = {
    {128, 128, launch_seg<SISI_FP64, 128, 128, 2, 2>, "seg128x128"},
    {80, 128, launch_seg<SISI_FP64, 80, 128, 1, 4>, "seg80x128"},
    {128, 80, launch_seg<SISI_FP64, 128, 80, 4, 1>, "seg128x80"},
    {80, 192, launch_seg<SISI_FP64, 80, 192, 1, 4>, "seg80x192"},
    {256, 80, launch_seg<SISI_FP64, 256, 80, 4, 1>, "seg256x80"},
};
static const SegVariant kSeg32[] = {
    {128, 128, launch_seg<SISI_FP32, 128, 128, 2, 2>, "seg128x128"},
    {80, 128, launch_seg<SISI_FP32, 80, 128, 1, 4>, "seg80x128"},
    {128, 80, launch_seg<SISI_FP32, 128, 80, 4, 1>, "seg128x80"},
    {80, 192, launch_seg<SISI_FP32, 80, 192, 1, 4>, "seg80x192"},
    {256, 80, launch_seg<SISI_FP32, 256, 80, 4, 1>, "seg256x80"},
};
constexpr int kNumSeg = 5;

// Size rule when no tuning record exists: an 80-wide tile along a dimension
// that is a multiple of 80 but not of 128, when that leaves the other at 128.
static int default_seg_variant(const GemmArgs &g) {
  const bool m80 = g.M % 80 == 0 && g.M % 128 != 0, n80 = g.N % 80 == 0 && g.N % 128 != 0;
  if (m80 && !n80) return 1;
  if (n80 && !m80) return 2;
  return 0;
}

static void launch_seg_variant(Engine &eng, SegGemmArgs a, int precision, const SegVariant &v,
                               int xcd, double flops) {
  GemmArgs &g = a.g;
  g.tiles_m = (g.M + v.bm - 1) / v.bm;
  g.tiles_n = (g.N + v.bn - 1) / v.bn;
  g.xcd_map = xcd;
  const int64_t tiles = int64_t(g.tiles_m) * g.tiles_n;
  SISI_REQUIRE(tiles < (int64_t(1) << 31), SISI_ERR_INVALID, "gemm_segments: too many tiles");
  const Kind kind = precision == SISI_FP64 ? Kind::Gemm64 : Kind::Gemm32;
  Timed t(*eng.timers, kind, flops);
  v.launch(a, dim3(unsigned(tiles), 1u, 1u), eng.stream);
  SISI_HIP(hipGetLastError());
}

// Time every tile variant (both tile walks) of one segmented product class
// into a scratch output; the fastest is recorded for the class.
static TuneChoice tune_seg(Engine &eng, const SegGemmArgs &a0, int precision, const TuneKey &key,
                           double flops) {
  const SegVariant *table = precision == SISI_FP64 ? kSeg64 : kSeg32;
  SegGemmArgs a = a0;
  GemmArgs &g = a.g;
  DBuf scratch(&eng.pool, size_t(g.M) * g.N);
  g.C = scratch.get();
  g.cm0 = 1;
  g.cn0 = g.M;
  g.cm1 = g.cm2 = g.cm3 = g.cn1 = g.cn2 = g.cn3 = 0;
  g.dm0 = g.M;
  g.dn0 = g.N;
  g.dm1 = g.dn1 = 1;
  g.split_out = 0;
  g.out_cube = 0;  // dense M x N scratch, whatever the product writes
  g.out_planes = 0;
  g.alpha = 1.0;
  g.beta = 0.0;
  const bool stats = eng.timers->enabled;
  eng.timers->enabled = false;
  hipEvent_t t0, t1;
  SISI_HIP(hipEventCreate(&t0));
  SISI_HIP(hipEventCreate(&t1));
  TuneChoice best;
  double best_us = 1e300;
  for (int v = 0; v < kNumSeg; ++v)
    for (int xcd = 1; xcd >= 0; --xcd) {
      launch_seg_variant(eng, a, precision, table[v], xcd, flops);
      SISI_HIP(hipEventRecord(t0, eng.stream));
      for (int r = 0; r < 3; ++r) launch_seg_variant(eng, a, precision, table[v], xcd, flops);
      SISI_HIP(hipEventRecord(t1, eng.stream));
      SISI_HIP(hipEventSynchronize(t1));
      float ms = 0.f;
      SISI_HIP(hipEventElapsedTime(&ms, t0, t1));
      const double us = 1e3 * double(ms) / 3;
      if (us < best_us) {
        best_us = us;
        best.variant = v;
        best.xcd = xcd;
        best.us = us;
      }
    }
  SISI_HIP(hipEventDestroy(t0));
  SISI_HIP(hipEventDestroy(t1));
  eng.timers->enabled = stats;
  tune_record(key, best, table[best.variant].name);
  return best;
}

static bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

void gemm_segments(Engine &eng, int precision, SegGemmArgs a) {
  GemmArgs &g = a.g;
  if (g.M <= 0 || g.N <= 0) return;
  SISI_REQUIRE(a.nseg >= 1 && a.nseg <= kMaxSegs, SISI_ERR_INVALID,
               "gemm_segments: 1 to 4 segments");
  SISI_REQUIRE(precision == SISI_FP64 || precision == SISI_FP32 || precision == SISI_BF16X3,
               SISI_ERR_INVALID, "gemm_segments: unknown precision");
  double flops = 0.0;
  int64_t ktot = 0;
  for (int s = 0; s < a.nseg; ++s) {
    GemmSeg &sg = a.seg[s];
    SISI_REQUIRE(sg.K > 0 && sg.A && sg.B, SISI_ERR_INVALID,
                 "gemm_segments: every segment needs K > 0 and operands");
    // the flags only choose which index consecutive threads walk
    sg.ak = (sg.sak == 1 && sg.sam != 1) ? 1 : 0;
    sg.bnf = (sg.sbn == 1 && sg.sbk != 1) ? 1 : 0;
    flops += 2.0 * double(g.M) * double(g.N) * double(sg.K);
    ktot += sg.K;
  }
  g.batch = 1;
  g.ksplit = 1;
  g.kchunk = 0;
  g.split_out = !(g.dm0 >= g.M && g.dn0 >= g.N);
  {
    static const int env_group = getenv("SISI_GEMM_GROUP") ? atoi(getenv("SISI_GEMM_GROUP")) : 8;
    g.group = env_group > 0 ? env_group : 8;
  }
  if (precision == SISI_BF16X3) {
    g.tiles_m = (g.M + BM - 1) / BM;
    g.tiles_n = (g.N + BN - 1) / BN;
    g.xcd_map = 1;
    const int64_t tiles = int64_t(g.tiles_m) * g.tiles_n;
    SISI_REQUIRE(tiles < (int64_t(1) << 31), SISI_ERR_INVALID, "gemm_segments: too many tiles");
    Timed t(*eng.timers, Kind::GemmB3, flops);
    hipLaunchKernelGGL(gemm_seg_b3_kernel, dim3(unsigned(tiles), 1u, 1u), dim3(256), 0,
                       eng.stream, a);
    SISI_HIP(hipGetLastError());
    return;
  }
  const SegVariant *table = precision == SISI_FP64 ? kSeg64 : kSeg32;
  const TuneKey key = tune_key(precision == SISI_FP64 ? TuneFamily::Seg64 : TuneFamily::Seg32,
                               g.M, g.N, ktot);
  TuneChoice c = tune_lookup(key);
  if (c.variant < 0 && tune_enabled() && !capturing(eng.stream))
    c = tune_seg(eng, a, precision, key, flops);
  int vid = c.variant >= 0 && c.variant < kNumSeg ? c.variant : default_seg_variant(g);
  const int xcd = c.variant >= 0 ? c.xcd : 1;
  launch_seg_variant(eng, a, precision, table[vid], xcd, flops);
}

}  // namespace sisi
