// Split-BF16x3 GEMM on pre-split operands.
//
// Each FP64 operand is rounded to FP32 and split once into two bf16 planes,
// x ~ hi + lo, stored k-contiguous ([rows][Kp], rows padded to the block
// tile, Kp to 32, zero-filled).  The GEMM then streams the planes with
// 16-byte loads and evaluates a_hi*b_lo + a_lo*b_hi + a_hi*b_hi per tile with
// v_mfma_f32_16x16x32_bf16 (FP32 accumulation); the FP64 epilogue applies
// alpha/beta and the split output addressing of the FP64 kernels.
#include <cstdlib>
#include <atomic>
#include <cstring>
#include <tuple>

#include "blaslt.hpp"
#include "engine.hpp"
#include "gemm.hpp"

namespace sisi {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16;
typedef unsigned int U128 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// operand split: (R x K) strided FP64 view -> hi/lo [Rp][Kp] bf16
// ---------------------------------------------------------------------------

struct SplitArgs {
  const double *src;
  int64_t sr, sk;  // element (r,k) at src[r*sr + k*sk]
  int R, K, Rp, Kp;
  u16 *hi, *lo;
  int64_t bsrc = 0, bdst = 0;  // batch strides (grid z) of source and planes
  double scale = 1.0;          // applied before the split (exact for +-1)
  int64_t ld = 0;              // row stride of hi/lo/hi2 in bf16 (0: Kp)
  u16 *hi2 = nullptr;          // second copy of hi (tripled rows), or none
};

__device__ __forceinline__ u16 bf16_bits(float f) {
  __bf16 b = __bf16(f);
  return *reinterpret_cast<u16 *>(&b);
}

// 64 rows x 64 k per block: read along the unit-stride index into LDS, then
// each thread converts 16 consecutive k of one row.
__global__ __launch_bounds__(256) void split_bf16_kernel(SplitArgs a) {
  __shared__ float t[64][65];
  const int r0 = blockIdx.y * 64, k0 = blockIdx.x * 64;
  const int tid = threadIdx.x;
  a.src += int64_t(blockIdx.z) * a.bsrc;
  a.hi += int64_t(blockIdx.z) * a.bdst;
  a.lo += int64_t(blockIdx.z) * a.bdst;
  if (a.sk == 1) {
    for (int e = tid; e < 4096; e += 256) {
      const int kk = e & 63, rr = e >> 6;
      const int r = r0 + rr, k = k0 + kk;
      t[rr][kk] = (r < a.R && k < a.K) ? float(a.scale * a.src[int64_t(r) * a.sr + k]) : 0.f;
    }
  } else {
    for (int e = tid; e < 4096; e += 256) {
      const int rr = e & 63, kk = e >> 6;
      const int r = r0 + rr, k = k0 + kk;
      t[rr][kk] =
          (r < a.R && k < a.K) ? float(a.scale * a.src[int64_t(r) * a.sr + int64_t(k) * a.sk]) : 0.f;
    }
  }
  __syncthreads();
  const int rr = tid >> 2, kq = (tid & 3) * 16;
  const int r = r0 + rr, k = k0 + kq;
  if (r >= a.Rp || k >= a.Kp) return;
  u16 h[16], l[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const float x = t[rr][kq + j];
    h[j] = bf16_bits(x);
    __bf16 hb = *reinterpret_cast<__bf16 *>(&h[j]);
    l[j] = bf16_bits(x - float(hb));
  }
  const int64_t row = int64_t(r) * (a.ld ? a.ld : a.Kp) + k;
  U128 *ph = reinterpret_cast<U128 *>(a.hi + row);
  U128 *pl = reinterpret_cast<U128 *>(a.lo + row);
  ph[0] = *reinterpret_cast<U128 *>(&h[0]);
  ph[1] = *reinterpret_cast<U128 *>(&h[8]);
  pl[0] = *reinterpret_cast<U128 *>(&l[0]);
  pl[1] = *reinterpret_cast<U128 *>(&l[8]);
  if (a.hi2) {
    U128 *p2 = reinterpret_cast<U128 *>(a.hi2 + row);
    p2[0] = *reinterpret_cast<U128 *>(&h[0]);
    p2[1] = *reinterpret_cast<U128 *>(&h[8]);
  }
}

// ---------------------------------------------------------------------------
// GEMM on split planes
// ---------------------------------------------------------------------------

struct B3Args {
  const u16 *Ah, *Al, *Bh, *Bl;  // [Mp][Kp], [Np][Kp]
  int64_t Kp;
  GemmArgs g;  // sizes, output addressing, alpha/beta, split-K
};

__device__ __forceinline__ int64_t b3_off_m(const GemmArgs &g, int m) {
  if (!g.split_out) return int64_t(m) * g.cm0;
  int m0 = m % g.dm0, t = m / g.dm0, m1 = t % g.dm1, m2 = t / g.dm1;
  return int64_t(m0) * g.cm0 + int64_t(m1) * g.cm1 + int64_t(m2) * g.cm2;
}
__device__ __forceinline__ int64_t b3_off_n(const GemmArgs &g, int n) {
  if (!g.split_out) return int64_t(n) * g.cn0;
  int n0 = n % g.dn0, t = n / g.dn0, n1 = t % g.dn1, n2 = t / g.dn1;
  return int64_t(n0) * g.cn0 + int64_t(n1) * g.cn1 + int64_t(n2) * g.cn2;
}

// BM x BN block, 4 waves of WM x WN, k-chunk 32, planes staged through LDS
// with 16-byte stores.  NBUF register sets keep NBUF chunks of global loads in
// flight, so a chunk's load latency is covered by NBUF chunks of MFMA work.
// PASSES = 3: split-BF16x3 (hi*lo + lo*hi + hi*hi); PASSES = 1: the hi*hi
// product alone (plain bf16 operands, FP32 accumulation), lo planes unread.
template <int BM, int BN, int WM, int WN, int NBUF, int MINB, int PASSES>
__global__ __launch_bounds__(256, MINB) void gemm_b3p_kernel(B3Args p) {
  constexpr int BK = 32, RS = BK + 8;  // LDS row stride in bf16 (80 bytes)
  constexpr int NA = BM * 4 / 256, NB = BN * 4 / 256;  // 16-byte chunks per thread per plane
  static_assert((BM / WM) * (BN / WN) == 4, "four waves");
  __shared__ __attribute__((aligned(16))) u16 Ah[BM * RS];
  __shared__ __attribute__((aligned(16))) u16 Al[BM * RS];
  __shared__ __attribute__((aligned(16))) u16 Bh[BN * RS];
  __shared__ __attribute__((aligned(16))) u16 Bl[BN * RS];
  const GemmArgs &g = p.g;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // tile walk: XCD-contiguous ranges, groups of 8 m-tiles
  int pid = blockIdx.x;
  {
    const int tiles = g.tiles_m * g.tiles_n, per = tiles >> 3;
    if (g.xcd_map && pid < (per << 3)) pid = (pid & 7) * per + (pid >> 3);
  }
  const int per_group = 8 * g.tiles_n, gid = pid / per_group, first_m = gid * 8;
  const int gsize = min(g.tiles_m - first_m, 8), rr = pid % per_group;
  const int tm = first_m + rr % gsize, tn = rr / gsize;
  const int m0 = tm * BM, n0 = tn * BN;
  const int ks = blockIdx.y, bz = blockIdx.z;
  const int kbeg = ks * g.kchunk;
  const int kend = min(int(p.Kp), kbeg + g.kchunk);
  const int nk = kend > kbeg ? (kend - kbeg) / BK : 0;
  const u16 *Agh = p.Ah + int64_t(m0) * p.Kp, *Agl = p.Al + int64_t(m0) * p.Kp;
  const u16 *Bgh = p.Bh + int64_t(n0) * p.Kp, *Bgl = p.Bl + int64_t(n0) * p.Kp;

  U128 ra[NBUF][2 * NA], rb[NBUF][2 * NB];
  // The k argument is evaluated once, before the macro declares any names of
  // its own, so an argument naming an outer variable cannot be captured.
#define B3_LOAD(b, kexpr)                                                                  \
  do {                                                                                    \
    const int kb_ = (kexpr);                                                              \
    _Pragma("unroll") for (int i_ = 0; i_ < NA; ++i_) {                                   \
      const int q_ = tid + 256 * i_, row_ = q_ >> 2, kq_ = (q_ & 3) * 8;                  \
      ra[b][i_] = *reinterpret_cast<const U128 *>(Agh + int64_t(row_) * p.Kp + kb_ + kq_);    \
      if (PASSES == 3)                                                                    \
        ra[b][NA + i_] = *reinterpret_cast<const U128 *>(Agl + int64_t(row_) * p.Kp + kb_ + kq_); \
    }                                                                                     \
    _Pragma("unroll") for (int i_ = 0; i_ < NB; ++i_) {                                   \
      const int q_ = tid + 256 * i_, row_ = q_ >> 2, kq_ = (q_ & 3) * 8;                  \
      rb[b][i_] = *reinterpret_cast<const U128 *>(Bgh + int64_t(row_) * p.Kp + kb_ + kq_);    \
      if (PASSES == 3)                                                                    \
        rb[b][NB + i_] = *reinterpret_cast<const U128 *>(Bgl + int64_t(row_) * p.Kp + kb_ + kq_); \
    }                                                                                     \
  } while (0)
#define B3_STORE(b)                                                                        \
  do {                                                                                    \
    _Pragma("unroll") for (int i_ = 0; i_ < NA; ++i_) {                                   \
      const int q_ = tid + 256 * i_, row_ = q_ >> 2, kq_ = (q_ & 3) * 8;                  \
      *reinterpret_cast<U128 *>(&Ah[row_ * RS + kq_]) = ra[b][i_];                         \
      if (PASSES == 3) *reinterpret_cast<U128 *>(&Al[row_ * RS + kq_]) = ra[b][NA + i_];   \
    }                                                                                     \
    _Pragma("unroll") for (int i_ = 0; i_ < NB; ++i_) {                                   \
      const int q_ = tid + 256 * i_, row_ = q_ >> 2, kq_ = (q_ & 3) * 8;                  \
      *reinterpret_cast<U128 *>(&Bh[row_ * RS + kq_]) = rb[b][i_];                         \
      if (PASSES == 3) *reinterpret_cast<U128 *>(&Bl[row_ * RS + kq_]) = rb[b][NB + i_];   \
    }                                                                                     \
  } while (0)

  constexpr int TI = WM / 16, TJ = WN / 16;
  f4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = f4{0, 0, 0, 0};
  constexpr int WX = BN / WN;
  const int wm = (wave / WX) * WM, wn = (wave % WX) * WN;
  const int l16 = lane & 15, lk = lane >> 4;

  // B fragments stay in registers for the chunk; A fragments are read one
  // 16-row tile at a time.  Per row tile, three passes (hi*lo, lo*hi, hi*hi)
  // over the TJ column tiles, so consecutive MFMAs update different
  // accumulators and none waits on the previous result.
  auto compute = [&]() {
    bf8 bh[TJ], bl[TJ];
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      const int col = wn + j * 16 + l16;
      bh[j] = *reinterpret_cast<const bf8 *>(&Bh[col * RS + 8 * lk]);
      if (PASSES == 3) bl[j] = *reinterpret_cast<const bf8 *>(&Bl[col * RS + 8 * lk]);
    }
#pragma unroll
    for (int i = 0; i < TI; ++i) {
      const int row = wm + i * 16 + l16;
      const bf8 ah = *reinterpret_cast<const bf8 *>(&Ah[row * RS + 8 * lk]);
      if constexpr (PASSES == 3) {
        const bf8 al = *reinterpret_cast<const bf8 *>(&Al[row * RS + 8 * lk]);
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah, acc[i][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al, acc[i][j], 0, 0, 0);
      }
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah, acc[i][j], 0, 0, 0);
    }
  };


#pragma unroll
  for (int b = 0; b < NBUF; ++b)
    if (b < nk) B3_LOAD(b, kbeg + b * BK);
  for (int kc = 0; kc < nk; kc += NBUF) {
#pragma unroll
    for (int b = 0; b < NBUF; ++b) {
      const int chunk = kc + b;
      if (chunk < nk) {
        B3_STORE(b);
        __syncthreads();
        if (chunk + NBUF < nk) B3_LOAD(b, kbeg + (chunk + NBUF) * BK);
        compute();
        __syncthreads();
      }
    }
  }
#undef B3_LOAD
#undef B3_STORE

  if (g.ksplit == 1) {
    double *C = g.C + int64_t(bz) * g.bsc;
#pragma unroll
    for (int i = 0; i < TI; ++i) {
      const int m = m0 + wm + i * 16 + l16;
      if (m >= g.M) continue;
      const int64_t om = b3_off_m(g, m);
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = n0 + wn + j * 16 + 4 * lk + r;
          if (n < g.N) {
            const int64_t o = om + b3_off_n(g, n);
            double v = g.alpha * double(acc[i][j][r]);
            if (g.beta != 0.0) v += g.beta * C[o];
            C[o] = v;
          }
        }
    }
  } else {
    double *W = g.work + (int64_t(bz) * g.ksplit + ks) * int64_t(g.M) * g.N;
#pragma unroll
    for (int i = 0; i < TI; ++i) {
      const int m = m0 + wm + i * 16 + l16;
      if (m >= g.M) continue;
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = n0 + wn + j * 16 + 4 * lk + r;
          if (n < g.N) W[m + int64_t(n) * g.M] = double(acc[i][j][r]);
        }
    }
  }
}

// Segmented form of gemm_b3p_kernel<128,128,64,64,1>: the chunks of all
// segments are walked in order with the next chunk (possibly of the next
// segment) prefetched into registers while the current one is multiplied.
__device__ __forceinline__ B3Seg b3seg_at(const B3SegArgs &a, int s) {
  return s == 0 ? a.seg[0] : s == 1 ? a.seg[1] : s == 2 ? a.seg[2] : a.seg[3];
}

__global__ __launch_bounds__(256) void gemm_b3p_seg_kernel(B3SegArgs p) {
  constexpr int BM = 128, BN = 128, WM = 64, WN = 64;
  constexpr int BK = 32, RS = BK + 8;
  constexpr int NA = BM * 4 / 256, NB = BN * 4 / 256;
  __shared__ __attribute__((aligned(16))) u16 Ah[BM * RS];
  __shared__ __attribute__((aligned(16))) u16 Al[BM * RS];
  __shared__ __attribute__((aligned(16))) u16 Bh[BN * RS];
  __shared__ __attribute__((aligned(16))) u16 Bl[BN * RS];
  const GemmArgs &g = p.g;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  int pid = blockIdx.x;
  {
    const int tiles = g.tiles_m * g.tiles_n, per = tiles >> 3;
    if (g.xcd_map && pid < (per << 3)) pid = (pid & 7) * per + (pid >> 3);
  }
  const int per_group = 8 * g.tiles_n, gid = pid / per_group, first_m = gid * 8;
  const int gsize = min(g.tiles_m - first_m, 8), rr = pid % per_group;
  const int tm = first_m + rr % gsize, tn = rr / gsize;
  const int m0 = tm * BM, n0 = tn * BN;

  U128 ra[2 * NA], rb[2 * NB];
  auto load = [&](int s, int c) {
    const B3Seg sg = b3seg_at(p, s);
    const int64_t Kp = sg.Kp;
    const int kb = c * BK;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int q = tid + 256 * i, row = q >> 2, kq = (q & 3) * 8;
      const int64_t off = int64_t(m0 + row) * Kp + kb + kq;
      ra[i] = *reinterpret_cast<const U128 *>(sg.Ah + off);
      ra[NA + i] = *reinterpret_cast<const U128 *>(sg.Al + off);
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int q = tid + 256 * i, row = q >> 2, kq = (q & 3) * 8;
      const int64_t off = int64_t(n0 + row) * Kp + kb + kq;
      rb[i] = *reinterpret_cast<const U128 *>(sg.Bh + off);
      rb[NB + i] = *reinterpret_cast<const U128 *>(sg.Bl + off);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int q = tid + 256 * i, row = q >> 2, kq = (q & 3) * 8;
      *reinterpret_cast<U128 *>(&Ah[row * RS + kq]) = ra[i];
      *reinterpret_cast<U128 *>(&Al[row * RS + kq]) = ra[NA + i];
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int q = tid + 256 * i, row = q >> 2, kq = (q & 3) * 8;
      *reinterpret_cast<U128 *>(&Bh[row * RS + kq]) = rb[i];
      *reinterpret_cast<U128 *>(&Bl[row * RS + kq]) = rb[NB + i];
    }
  };

  constexpr int TI = WM / 16, TJ = WN / 16;
  f4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = f4{0, 0, 0, 0};
  const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
  const int l16 = lane & 15, lk = lane >> 4;
  auto compute = [&]() {
    bf8 bh[TJ], bl[TJ];
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      const int col = wn + j * 16 + l16;
      bh[j] = *reinterpret_cast<const bf8 *>(&Bh[col * RS + 8 * lk]);
      bl[j] = *reinterpret_cast<const bf8 *>(&Bl[col * RS + 8 * lk]);
    }
#pragma unroll
    for (int i = 0; i < TI; ++i) {
      const int row = wm + i * 16 + l16;
      const bf8 ah = *reinterpret_cast<const bf8 *>(&Ah[row * RS + 8 * lk]);
      const bf8 al = *reinterpret_cast<const bf8 *>(&Al[row * RS + 8 * lk]);
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah, acc[i][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al, acc[i][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah, acc[i][j], 0, 0, 0);
    }
  };

  int s = 0, c = 0;
  bool have = p.nseg > 0;
  if (have) load(0, 0);
  while (have) {
    store();
    __syncthreads();
    int s1 = s, c1 = c + 1;
    if (c1 * BK >= b3seg_at(p, s).Kp) {
      s1 = s + 1;
      c1 = 0;
    }
    const bool more = s1 < p.nseg;
    if (more) load(s1, c1);
    compute();
    __syncthreads();
    s = s1;
    c = c1;
    have = more;
  }

#pragma unroll
  for (int i = 0; i < TI; ++i) {
    const int m = m0 + wm + i * 16 + l16;
    if (m >= g.M) continue;
    const int64_t om = b3_off_m(g, m);
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn + j * 16 + 4 * lk + r;
        if (n < g.N) {
          const int64_t o = om + b3_off_n(g, n);
          double v = g.alpha * double(acc[i][j][r]);
          if (g.beta != 0.0) v += g.beta * g.C[o];
          g.C[o] = v;
        }
      }
  }
}

namespace {

struct Planes {
  DBuf store;  // the planes in one allocation (doubles as raw storage)
  const u16 *hi = nullptr, *lo = nullptr;
  const u16 *base = nullptr;  // start of row 0 (tripled layouts)
  int64_t ld = 0;             // row stride in bf16
};

// Plane layouts: 0 = separate [Rp][Kp] hi and lo planes (hand-written
// kernels); 1 = rows [hi | hi | lo] and 2 = rows [hi | lo | hi] of 3 Kp bf16,
// so that sum_k' A'[m,k'] B'[n,k'] over k' < 3 Kp is hi*hi + hi*lo + lo*hi
// with A in layout 1 and B in layout 2: one library GEMM with K tripled.
void plane_pointers(u16 *base, int64_t Rp, int64_t Kp, int layout, u16 *&hi, u16 *&lo,
                    u16 *&hi2, int64_t &ld) {
  hi = base;
  hi2 = nullptr;
  if (layout == 0) {
    lo = base + Rp * Kp;
    ld = Kp;
  } else {
    ld = 3 * Kp;
    lo = base + (layout == 1 ? 2 * Kp : Kp);
    hi2 = base + (layout == 1 ? Kp : 2 * Kp);
  }
}

// Split one (rows x K) operand view; frozen sources are split once per
// residual evaluation and reused.
const Planes &split_operand(Engine &e, const double *src, int64_t sr, int64_t sk, int R, int K,
                            int Rp, int Kp, Planes &local, int layout = 0,
                            bool allow_cache = true) {
  std::string key;
  const bool cached = allow_cache && e.is_frozen(src);
  const int64_t elems = int64_t(Rp) * Kp;  // bf16 per plane
  const int nplanes = layout == 0 ? 2 : 3;
  auto fill = [&](u16 *base) {
    u16 *hi, *lo, *hi2;
    int64_t ld;
    plane_pointers(base, Rp, Kp, layout, hi, lo, hi2, ld);
    local.base = base;
    local.hi = hi;
    local.lo = lo;
    local.ld = ld;
    return std::make_tuple(hi, lo, hi2, ld);
  };
  if (cached) {
    key = "b3:" + std::to_string(reinterpret_cast<uintptr_t>(src)) + ":" + std::to_string(sr) +
          ":" + std::to_string(sk) + ":" + std::to_string(R) + ":" + std::to_string(K) + ":" +
          std::to_string(Rp) + ":" + std::to_string(Kp) + ":" + std::to_string(layout);
    auto it = e.split_cache.find(key);
    if (it != e.split_cache.end()) {
      fill(reinterpret_cast<u16 *>(it->second.get()));
      return local;
    }
  }
  const size_t doubles = size_t((nplanes * elems * 2 + 7) / 8);
  local.store = DBuf(&e.pool, doubles);
  auto [hi, lo, hi2, ld] = fill(reinterpret_cast<u16 *>(local.store.get()));
  SplitArgs a{src, sr, sk, R, K, Rp, Kp, hi, lo};
  a.ld = ld;
  a.hi2 = hi2;
  {
    Timed t(*e.timers, Kind::Bandwidth, 8.0 * double(R) * K + 2.0 * nplanes * double(elems));
    dim3 grid(unsigned((Kp + 63) / 64), unsigned((Rp + 63) / 64));
    hipLaunchKernelGGL(split_bf16_kernel, grid, dim3(256), 0, e.stream, a);
    SISI_HIP(hipGetLastError());
  }
  if (cached) {
    e.split_cache[key] = std::move(local.store);
    fill(reinterpret_cast<u16 *>(e.split_cache[key].get()));
  }
  return local;
}

}  // namespace

// C[off(m, n)] = alpha * C32[m + M n] + beta * C[off(m, n)]: the FP64
// epilogue of a library product (same output addressing as the kernels).
__global__ void b3_epilogue_kernel(const float *C32, GemmArgs g) {
  const int64_t n = int64_t(g.M) * g.N;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int m = int(e % g.M), c = int(e / g.M);
    const int64_t o = b3_off_m(g, m) + b3_off_n(g, c);
    double v = g.alpha * double(C32[e]);
    if (g.beta != 0.0) v += g.beta * g.C[o];
    g.C[o] = v;
  }
}

// Library backend of one split-BF16x3 (passes = 3: A in layout 1, B in
// layout 2, one bf16 GEMM over 3 Kp) or plain bf16 (passes = 1: hi rows of
// any layout) product: FP32 accumulation through hipBLASLt, then the FP64
// epilogue.  False: no library algorithm for the shape.
static bool b3_library(Engine &e, const GemmArgs &g, const Planes &A, const Planes &B, int Kp,
                       int passes) {
  DBuf c32(&e.pool, (size_t(g.M) * g.N + 1) / 2);
  float *C = reinterpret_cast<float *>(c32.get());
  Timed t(*e.timers, Kind::GemmB3, gemm_flops(g));
  const int K = passes == 3 ? 3 * Kp : Kp;
  if (!blaslt_bf16_gemm(e, g.M, g.N, K, A.hi, A.ld, B.hi, B.ld, C, 1.0f, 0.0f)) return false;
  const int64_t n = int64_t(g.M) * g.N;
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 1 << 16));
  hipLaunchKernelGGL(b3_epilogue_kernel, dim3(blocks), dim3(256), 0, e.stream, C, g);
  SISI_HIP(hipGetLastError());
  return true;
}

// Backend of the batch-1 split-BF16x3 and plain bf16 products: the
// hand-written kernel (default; deterministic) or hipBLASLt on the same
// split planes (SISI_B3_BACKEND=library: 1.3x the kernel's rate on the
// ladder and ring shapes including the epilogue, run-to-run rounding not
// reproducible).  Segmented and batched products always use the
// hand-written kernels.
static std::atomic<int> &b3_backend() {
  static std::atomic<int> backend{[] {
    const char *v = getenv("SISI_B3_BACKEND");
    return v != nullptr && std::strcmp(v, "library") == 0 ? 1 : 0;
  }()};
  return backend;
}
static bool b3_use_library() { return b3_backend().load() == 1; }
// The plain bf16 products (the refinement residuals' changes) run on
// hipBLASLt by default: 17 % faster refinement iterations at No=40, Nv=400,
// NG=1600 with the same iteration count; their rounding, already at the bf16
// level and corrected by the FP64 anchors, is not reproducible bit for bit.
// SISI_B1_BACKEND=kernel keeps them on the hand-written kernel.
static bool b1_use_library() {
  static const bool lib = [] {
    const char *v = getenv("SISI_B1_BACKEND");
    return v == nullptr || std::strcmp(v, "kernel") != 0;
  }();
  return lib;
}

void set_b3_backend(int library) { b3_backend().store(library ? 1 : 0); }
int get_b3_backend() { return b3_backend().load(); }

template <int BM, int BN, int WM, int WN, int NBUF, int MINB = 1, int PASSES = 3>
static void run_b3(Engine &e, GemmArgs g, const B3FrozenA *fa = nullptr) {
  const int Kp = ((g.K + 31) / 32) * 32;
  const int Mp = ((g.M + BM - 1) / BM) * BM, Np = ((g.N + BN - 1) / BN) * BN;
  if (PASSES == 1 && fa && (b3_use_library() || b1_use_library())) {
    // frozen hi plane: rows of Kp bf16, the first M rows used
    g.split_out = !(g.dm0 >= g.M && g.dn0 >= g.N);
    Planes la, lb;
    la.hi = reinterpret_cast<const u16 *>(fa->store.get());
    la.ld = Kp;
    SISI_REQUIRE(fa->M == g.M && fa->Kp == Kp, SISI_ERR_INVALID, "frozen A planes do not match");
    const Planes &B = split_operand(e, g.B + 0, g.sbn, g.sbk, g.N, g.K, g.N, Kp, lb, 0);
    if (b3_library(e, g, la, B, Kp, 1)) return;
  }
  if ((b3_use_library() || (PASSES == 1 && b1_use_library())) && !fa) {
    g.split_out = !(g.dm0 >= g.M && g.dn0 >= g.N);
    Planes la, lb;
    const int lay_a = PASSES == 3 ? 1 : 0, lay_b = PASSES == 3 ? 2 : 0;
    const Planes &A = split_operand(e, g.A + 0, g.sam, g.sak, g.M, g.K, g.M, Kp, la, lay_a);
    const Planes &B = split_operand(e, g.B + 0, g.sbn, g.sbk, g.N, g.K, g.N, Kp, lb, lay_b);
    if (b3_library(e, g, A, B, Kp, PASSES)) return;
  }
  Planes pa, pb;
  if (fa) {  // planes split earlier in the layout split_operand gives (layout 0)
    SISI_REQUIRE(fa->store.get() && fa->M == g.M && fa->K == g.K && fa->Mp == Mp && fa->Kp == Kp,
                 SISI_ERR_INVALID, "frozen A planes do not match the product");
    pa.hi = reinterpret_cast<const u16 *>(fa->store.get());
    pa.lo = pa.hi + int64_t(Mp) * Kp;
    pa.base = pa.hi;
    pa.ld = Kp;
  }
  const Planes &A = fa ? pa : split_operand(e, g.A + 0, g.sam, g.sak, g.M, g.K, Mp, Kp, pa);
  const Planes &B = split_operand(e, g.B + 0, g.sbn, g.sbk, g.N, g.K, Np, Kp, pb);
  g.tiles_m = Mp / BM;
  g.tiles_n = Np / BN;
  const int64_t tiles = int64_t(g.tiles_m) * g.tiles_n;
  g.split_out = !(g.dm0 >= g.M && g.dn0 >= g.N);
  int ksplit = 1;
  if (Kp > 4 * 32 && tiles < 512) {
    const int64_t want = (512 + tiles - 1) / tiles;
    const int64_t smax = std::max<int64_t>(1, std::min<int64_t>(Kp / 256, 64));
    ksplit = round_filling_split(tiles, std::min(want, smax), smax, e.cus);
  }
  int kchunk = ((Kp + ksplit - 1) / ksplit + 31) / 32 * 32;
  ksplit = (Kp + kchunk - 1) / kchunk;
  g.ksplit = ksplit;
  g.kchunk = kchunk;
  DBuf work;
  if (ksplit > 1) {
    work = DBuf(&e.pool, size_t(ksplit) * g.M * g.N);
    g.work = work.get();
  }
  B3Args p{A.hi, A.lo, B.hi, B.lo, Kp, g};
  dim3 grid(unsigned(tiles), unsigned(ksplit), 1);
  {
    Timed t(*e.timers, Kind::GemmB3, gemm_flops(g));
    hipLaunchKernelGGL((gemm_b3p_kernel<BM, BN, WM, WN, NBUF, MINB, PASSES>), grid, dim3(256), 0,
                       e.stream, p);
    SISI_HIP(hipGetLastError());
  }
  if (ksplit > 1) gemm_reduce(e, g);
}

void split_planes(Engine &e, const double *src, int64_t sr, int64_t sk, int R, int K, int Rp,
                  int batch, int64_t bsrc, double scale, B3Planes &out) {
  SISI_REQUIRE(R > 0 && K > 0 && Rp >= R && batch > 0 && batch <= 65535, SISI_ERR_INVALID,
               "split_planes: bad shape");
  const int Kp = ((K + 31) / 32) * 32;
  const int64_t elems = int64_t(batch) * Rp * Kp;  // bf16 per plane
  out.store = DBuf(&e.pool, size_t((2 * elems * 2 + 7) / 8));
  u16 *hi = reinterpret_cast<u16 *>(out.store.get());
  u16 *lo = hi + elems;
  SplitArgs a{src, sr, sk, R, K, Rp, Kp, hi, lo};
  a.bsrc = bsrc;
  a.bdst = int64_t(Rp) * Kp;
  a.scale = scale;
  {
    Timed t(*e.timers, Kind::Bandwidth, double(batch) * (8.0 * double(R) * K + 4.0 * double(Rp) * Kp));
    dim3 grid(unsigned((Kp + 63) / 64), unsigned((Rp + 63) / 64), unsigned(batch));
    hipLaunchKernelGGL(split_bf16_kernel, grid, dim3(256), 0, e.stream, a);
    SISI_HIP(hipGetLastError());
  }
  out.hi = hi;
  out.lo = lo;
  out.Rp = Rp;
  out.Kp = Kp;
  out.batch = batch;
}

void gemm_b3_segments(Engine &e, B3SegArgs a) {
  GemmArgs &g = a.g;
  if (g.M <= 0 || g.N <= 0) return;
  SISI_REQUIRE(a.nseg >= 1 && a.nseg <= kMaxSegs, SISI_ERR_INVALID,
               "gemm_b3_segments: 1 to 4 segments");
  double flops = 0.0;
  for (int s = 0; s < a.nseg; ++s) {
    SISI_REQUIRE(a.seg[s].Kp > 0 && a.seg[s].Kp % 32 == 0 && a.seg[s].Ah && a.seg[s].Bh,
                 SISI_ERR_INVALID, "gemm_b3_segments: planes with Kp a multiple of 32");
    flops += 2.0 * double(g.M) * double(g.N) * double(a.seg[s].Kp);
  }
  g.batch = 1;
  g.ksplit = 1;
  g.tiles_m = (g.M + kB3SegTile - 1) / kB3SegTile;
  g.tiles_n = (g.N + kB3SegTile - 1) / kB3SegTile;
  g.split_out = !(g.dm0 >= g.M && g.dn0 >= g.N);
  g.xcd_map = 1;
  const int64_t tiles = int64_t(g.tiles_m) * g.tiles_n;
  SISI_REQUIRE(tiles < (int64_t(1) << 31), SISI_ERR_INVALID, "gemm_b3_segments: too many tiles");
  Timed t(*e.timers, Kind::GemmB3, flops);
  hipLaunchKernelGGL(gemm_b3p_seg_kernel, dim3(unsigned(tiles)), dim3(256), 0, e.stream, a);
  SISI_HIP(hipGetLastError());
}

void b3_freeze_a(Engine &e, const double *A, int64_t sam, int64_t sak, int M, int K,
                 B3FrozenA &out) {
  SISI_REQUIRE(A && M > 0 && K > 0, SISI_ERR_INVALID, "b3_freeze_a: empty operand");
  const int Kp = ((K + 31) / 32) * 32, Mp = ((M + 127) / 128) * 128;
  Planes local;
  split_operand(e, A, sam, sak, M, K, Mp, Kp, local, 0, false);
  out.store = std::move(local.store);
  out.M = M;
  out.K = K;
  out.Mp = Mp;
  out.Kp = Kp;
}

void gemm_b3_frozen_a(Engine &e, GemmArgs g, const B3FrozenA &A, int passes) {
  if (g.M <= 0 || g.N <= 0) return;
  SISI_REQUIRE(g.batch == 1, SISI_ERR_INVALID, "gemm_b3_frozen_a: batch 1 only");
  // the 128-row kernels whose plane layout b3_freeze_a produced
  if (passes == 1) run_b3<128, 128, 64, 64, 1, 1, 1>(e, g, &A);
  else run_b3<128, 128, 64, 64, 1>(e, g, &A);
}

void gemm_b3(Engine &e, GemmArgs g, int passes) {
  // 128x128 blocks of four 64x64 waves, two workgroups per CU, measured
  // fastest on the engine's shapes; the others are kept for measurement
  static const int cfg = getenv("SISI_B3_CFG") ? atoi(getenv("SISI_B3_CFG")) : 2;
  if (passes == 1) {
    run_b3<128, 128, 64, 64, 1, 1, 1>(e, g);
    return;
  }
  switch (cfg) {
    case 0: run_b3<256, 128, 128, 64, 2>(e, g); break;
    case 1: run_b3<256, 128, 128, 64, 1>(e, g); break;
    case 3: run_b3<128, 128, 64, 64, 2>(e, g); break;
    case 5: run_b3<128, 128, 64, 64, 1, 3>(e, g); break;
    default: run_b3<128, 128, 64, 64, 1>(e, g); break;
  }
}

}  // namespace sisi
