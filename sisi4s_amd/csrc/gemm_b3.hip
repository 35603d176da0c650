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
#include "gemm_device.hpp"
#include "tune.hpp"

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

// Epilogue store of one output element: FP64 (alpha, beta), or the bf16
// hi/lo planes of float(alpha * acc) (out_planes).
__device__ __forceinline__ void b3_store(const GemmArgs &g, double *C, int64_t o, float acc) {
  if (g.out_planes) {
    const float x = float(g.alpha * double(acc));
    const u16 h = bf16_bits(x);
    const __bf16 hb = *reinterpret_cast<const __bf16 *>(&h);
    u16 *P = reinterpret_cast<u16 *>(C);
    P[o] = h;
    if (g.out_planes == 1) P[o + g.plane_lo] = bf16_bits(x - float(hb));
  } else {
    double v = g.alpha * double(acc);
    if (g.beta != 0.0) v += g.beta * C[o];
    C[o] = v;
  }
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
  ph[0] = *reinterpret_cast<U128 *>(&h[0]);
  ph[1] = *reinterpret_cast<U128 *>(&h[8]);
  if (a.lo) {  // hi-only planes (plain-bf16 consumers) have no lo plane
    U128 *pl = reinterpret_cast<U128 *>(a.lo + row);
    pl[0] = *reinterpret_cast<U128 *>(&l[0]);
    pl[1] = *reinterpret_cast<U128 *>(&l[8]);
  }
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
  if (g.out_cube) return cube_coord(m % g.dm0, g.cm0, g.cm1) + cube_coord(m / g.dm0, g.cm2, g.cm3);
  if (!g.split_out) return int64_t(m) * g.cm0;
  int m0 = m % g.dm0, t = m / g.dm0, m1 = t % g.dm1, m2 = t / g.dm1;
  return int64_t(m0) * g.cm0 + int64_t(m1) * g.cm1 + int64_t(m2) * g.cm2;
}
__device__ __forceinline__ int64_t b3_off_n(const GemmArgs &g, int n) {
  if (g.out_cube) return cube_coord(n % g.dn0, g.cn0, g.cn1) + cube_coord(n / g.dn0, g.cn2, g.cn3);
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
    // runs of four consecutive columns per lane: offsets hoisted as in the
    // 256-row kernel's epilogue
    const bool run4 = (!g.split_out && !g.out_cube) || (g.dn0 & 3) == 0;
    int64_t offn[TJ];
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      const int n = n0 + wn + j * 16 + 4 * lk;
      offn[j] = run4 && n < g.N ? b3_off_n(g, n) : 0;
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
          if (n < g.N)
            b3_store(g, C, om + (run4 ? offn[j] + int64_t(r) * g.cn0 : b3_off_n(g, n)),
                     acc[i][j][r]);
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

// ---------------------------------------------------------------------------
// Large-tile GEMM on split planes with direct global-to-LDS loads
// ---------------------------------------------------------------------------
//
// BM x BN block of WGM x WGN wave64s, k-chunks of 32 bf16 (64-byte rows),
// STAGES chunks resident in LDS as a ring.  Each chunk is fetched with
// global_load_lds_dwordx4: one instruction moves 16 rows x 64 bytes (64
// lanes x 16 bytes) of one plane straight into LDS, no registers involved,
// so a chunk's fetch runs while the MFMAs of earlier chunks execute.  Lane l
// of an instruction writes 16-byte slot l & 3 of row l >> 2; it reads the
// k-chunk (l & 3) ^ swz(row), so the rows are stored XOR-swizzled and the
// fragment reads (32 rows of one 16-byte chunk per half wave) spread over
// the LDS banks.
//
// Per chunk: wait for this thread's loads of the chunk (vmcnt), barrier
// (every wave's loads landed, every wave done with the slot about to be
// refilled), issue the loads of chunk c + STAGES - 1 into the slot freed by
// chunk c - 1, multiply chunk c.  One barrier per chunk.
//
// MFMA: v_mfma_f32_32x32x16_bf16 with the B plane in the instruction's A
// role and the A plane in its B role, so lane l, register r of a 32x32 tile
// holds C(m = l % 32, n = 8 (r / 4) + 4 (l / 32) + r % 4): 32 lanes store
// 32 consecutive m (the output's unit-stride index).

typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

#define SISI_LDS_PTR(p) ((__attribute__((address_space(3))) void *)(p))

// s_waitcnt with vmcnt(V), lgkmcnt(0), expcnt untouched (gfx9 encoding:
// vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
template <int V>
__device__ __forceinline__ void wait_vm() {
  static_assert(V >= 0 && V < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((V & 15) | ((V >> 4) << 14) | (7 << 4));
}

// Workgroup barrier without the memory fence of __syncthreads (whose release
// half waits for every outstanding global load, including the prefetches);
// the asm statements keep the compiler from moving memory accesses across.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// 16 bytes per lane from global memory straight into LDS: lane l's bytes land
// at dst + 16 l (dst uniform across the wave).
__device__ __forceinline__ void dma16(const u16 *src, unsigned char *dst) {
  __builtin_amdgcn_global_load_lds(src, SISI_LDS_PTR(dst), 16, 0, 0);
}

template <int BM, int BN, int PASSES, int STAGES>
constexpr int b3d_lds_bytes() {
  return STAGES * (PASSES == 3 ? 2 : 1) * (BM + BN) * 64;
}

template <int BM, int BN, int WGM, int WGN, int PASSES, int STAGES, int SWZ, int SCHED>
__global__ __launch_bounds__(64 * WGM * WGN, 1) void gemm_b3d_kernel(B3Args p) {
  constexpr int NW = WGM * WGN;
  constexpr int WM = BM / WGM, WN = BN / WGN, TI = WM / 32, TJ = WN / 32;
  constexpr int PL = PASSES == 3 ? 2 : 1;
  constexpr int A_BYTES = BM * 64, B_BYTES = BN * 64;
  constexpr int STAGE_BYTES = PL * (A_BYTES + B_BYTES);
  constexpr int IA = BM / (16 * NW), IB = BN / (16 * NW);
  static_assert(IA * 16 * NW == BM && IB * 16 * NW == BN, "16-row blocks per wave");
  static_assert(TI * 32 == WM && TJ * 32 == WN, "32x32 tiles per wave");
  constexpr int LPS = PL * (IA + IB);  // loads per chunk per thread
  static_assert(LPS * (STAGES - 2 > 0 ? STAGES - 2 : 0) < 64, "vmcnt range");
  extern __shared__ __attribute__((aligned(1024))) unsigned char smem[];

  const GemmArgs &g = p.g;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
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
  const int nk = kend > kbeg ? (kend - kbeg) / 32 : 0;

  // this lane's source row (within a 16-row block) and k-chunk
  const int lr = lane >> 2;
  const int lc = (lane & 3) ^ ((lr >> SWZ) & 3);
  const int64_t Kp = p.Kp;
  const u16 *srcA[PL], *srcB[PL];
  srcA[0] = p.Ah + (int64_t(m0) + lr) * Kp + 8 * lc;
  srcB[0] = p.Bh + (int64_t(n0) + lr) * Kp + 8 * lc;
  if constexpr (PL == 2) {
    srcA[PL - 1] = p.Al + (int64_t(m0) + lr) * Kp + 8 * lc;
    srcB[PL - 1] = p.Bl + (int64_t(n0) + lr) * Kp + 8 * lc;
  }

  auto fetch = [&](int chunk, int slot) {
    const int kb = kbeg + chunk * 32;
    unsigned char *st = smem + slot * STAGE_BYTES;
#pragma unroll
    for (int q = 0; q < PL; ++q) {
#pragma unroll
      for (int i = 0; i < IA; ++i) {
        const int row = 16 * (wave + NW * i);
        dma16(srcA[q] + int64_t(row) * Kp + kb, st + q * A_BYTES + row * 64);
      }
#pragma unroll
      for (int i = 0; i < IB; ++i) {
        const int row = 16 * (wave + NW * i);
        dma16(srcB[q] + int64_t(row) * Kp + kb, st + PL * A_BYTES + q * B_BYTES + row * 64);
      }
    }
  };

  f16v acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  const int wm = (wave / WGN) * WM, wn = (wave % WGN) * WN;
  const int l32 = lane & 31, lh = lane >> 5;

  // byte offset of (row, 16-byte chunk c) in a plane
  auto at = [](int row, int c) { return row * 64 + ((c ^ ((row >> SWZ) & 3)) << 4); };

  // The fragments of both k16 steps of a chunk are read from LDS before its
  // MFMAs, so the reads of the second step are in flight while the first
  // step's MFMAs issue (the compiler waits only for the set it uses).
  auto compute = [&](int slot) {
    const unsigned char *st = smem + slot * STAGE_BYTES;
    bf8v bh[2][TJ], bl[2][PASSES == 3 ? TJ : 1], ah[2][TI], al[2][PASSES == 3 ? TI : 1];
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const int c = 2 * kk + lh;
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        const int o = at(wn + 32 * j + l32, c);
        bh[kk][j] = *reinterpret_cast<const bf8v *>(st + PL * A_BYTES + o);
        if constexpr (PASSES == 3)
          bl[kk][j] = *reinterpret_cast<const bf8v *>(st + PL * A_BYTES + B_BYTES + o);
      }
#pragma unroll
      for (int i = 0; i < TI; ++i) {
        const int o = at(wm + 32 * i + l32, c);
        ah[kk][i] = *reinterpret_cast<const bf8v *>(st + o);
        if constexpr (PASSES == 3) al[kk][i] = *reinterpret_cast<const bf8v *>(st + A_BYTES + o);
      }
    }
    if constexpr (SCHED == 2 && PASSES == 3) {
      // pass-outer order: TI * TJ independent accumulations between two
      // MFMAs into the same accumulator
#pragma unroll
      for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
          for (int j = 0; j < TJ; ++j)
            acc[i][j] =
                __builtin_amdgcn_mfma_f32_32x32x16_bf16(bl[kk][j], ah[kk][i], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
          for (int j = 0; j < TJ; ++j)
            acc[i][j] =
                __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[kk][j], al[kk][i], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
          for (int j = 0; j < TJ; ++j)
            acc[i][j] =
                __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[kk][j], ah[kk][i], acc[i][j], 0, 0, 0);
      }
    } else {
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
      for (int i = 0; i < TI; ++i) {
        if constexpr (PASSES == 3) {
#pragma unroll
          for (int j = 0; j < TJ; ++j)
            acc[i][j] =
                __builtin_amdgcn_mfma_f32_32x32x16_bf16(bl[kk][j], ah[kk][i], acc[i][j], 0, 0, 0);
#pragma unroll
          for (int j = 0; j < TJ; ++j)
            acc[i][j] =
                __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[kk][j], al[kk][i], acc[i][j], 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          acc[i][j] =
              __builtin_amdgcn_mfma_f32_32x32x16_bf16(bh[kk][j], ah[kk][i], acc[i][j], 0, 0, 0);
      }
    }
    // schedule: the first step's reads, then the second step's reads
    // interleaved one to one with the first step's MFMAs, then the rest
    // (SCHED = 1; SCHED = 0 leaves the order to the compiler)
    if constexpr (SCHED == 1) {
      constexpr int R = (TI + TJ) * (PASSES == 3 ? 2 : 1), Q = TI * TJ * PASSES;
      __builtin_amdgcn_sched_group_barrier(0x100, R, 0);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x008, 2 * Q - R, 0);
    }
  };

#pragma unroll
  for (int s = 0; s < STAGES - 1; ++s)
    if (s < nk) fetch(s, s);
  for (int c = 0; c < nk; ++c) {
    // loads issued after chunk c's that may stay in flight
    const int ahead = min(STAGES - 2, nk - 1 - c);
    if constexpr (STAGES >= 4) {
      if (ahead >= 2) wait_vm<2 * LPS>();
      else if (ahead == 1) wait_vm<LPS>();
      else wait_vm<0>();
    } else if constexpr (STAGES == 3) {
      if (ahead >= 1) wait_vm<LPS>();
      else wait_vm<0>();
    } else {
      wait_vm<0>();
    }
    lds_barrier();
    if (c + STAGES - 1 < nk) fetch(c + STAGES - 1, (c + STAGES - 1) % STAGES);
    compute(c % STAGES);
  }

  if (g.ksplit == 1) {
    double *C = g.C + int64_t(bz) * g.bsc;
    // the lane's columns come in aligned runs of four; when the output split
    // of n keeps such a run together (dn0 a multiple of 4, or no split) the
    // run's offsets are the first one plus multiples of cn0, computed once
    // here instead of per row and element
    const bool run4 = (!g.split_out && !g.out_cube) || (g.dn0 & 3) == 0;
    int64_t offn[TJ][4];
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int n = n0 + wn + 32 * j + 8 * q + 4 * lh;
        offn[j][q] = run4 && n < g.N ? b3_off_n(g, n) : 0;
      }
#pragma unroll
    for (int i = 0; i < TI; ++i) {
      const int m = m0 + wm + 32 * i + l32;
      if (m >= g.M) continue;
      const int64_t om = b3_off_m(g, m);
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int n = n0 + wn + 32 * j + 8 * (r >> 2) + 4 * lh + (r & 3);
          if (n < g.N)
            b3_store(g, C,
                     om + (run4 ? offn[j][r >> 2] + int64_t(r & 3) * g.cn0 : b3_off_n(g, n)),
                     acc[i][j][r]);
        }
    }
  } else {
    double *W = g.work + (int64_t(bz) * g.ksplit + ks) * int64_t(g.M) * g.N;
#pragma unroll
    for (int i = 0; i < TI; ++i) {
      const int m = m0 + wm + 32 * i + l32;
      if (m >= g.M) continue;
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int n = n0 + wn + 32 * j + 8 * (r >> 2) + 4 * lh + (r & 3);
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

  // row offsets once per row, each column offset once per column (the
  // cube-blocked output splits cost divisions)
  int64_t om[TI];
#pragma unroll
  for (int i = 0; i < TI; ++i) {
    const int m = m0 + wm + i * 16 + l16;
    om[i] = m < g.M ? b3_off_m(g, m) : -1;
  }
#pragma unroll
  for (int j = 0; j < TJ; ++j)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = n0 + wn + j * 16 + 4 * lk + r;
      if (n >= g.N) continue;
      const int64_t on = b3_off_n(g, n);
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
// Plane layouts: 0 hi plane then lo plane; 1 / 2 rows of (hi, hi, lo) /
// (hi, lo, hi) for the library path; kHiOnly the hi plane alone.
constexpr int kHiOnly = 3;

void plane_pointers(u16 *base, int64_t Rp, int64_t Kp, int layout, u16 *&hi, u16 *&lo,
                    u16 *&hi2, int64_t &ld) {
  hi = base;
  hi2 = nullptr;
  if (layout == 0) {
    lo = base + Rp * Kp;
    ld = Kp;
  } else if (layout == kHiOnly) {  // plain-bf16 consumers read the hi plane alone
    lo = nullptr;
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
  const int nplanes = layout == 0 ? 2 : layout == kHiOnly ? 1 : 3;
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
// The plain bf16 products (the refinement residuals' changes) run on the
// hand-written kernels; SISI_B1_BACKEND=library sends them to hipBLASLt on
// the same planes (kept for comparison; rounding not reproducible).
static bool b1_use_library() {
  static const bool lib = [] {
    const char *v = getenv("SISI_B1_BACKEND");
    return v != nullptr && std::strcmp(v, "library") == 0;
  }();
  return lib;
}

void set_b3_backend(int library) { b3_backend().store(library ? 1 : 0); }
int get_b3_backend() { return b3_backend().load(); }

// ---------------------------------------------------------------------------
// Kernel variants of the batch-1 products and their launchers
// ---------------------------------------------------------------------------

template <int BM, int BN, int WM, int WN, int NBUF, int MINB, int PASSES>
static void launch_b3p(const B3Args &p, dim3 grid, hipStream_t s) {
  hipLaunchKernelGGL((gemm_b3p_kernel<BM, BN, WM, WN, NBUF, MINB, PASSES>), grid, dim3(256), 0, s,
                     p);
}

template <int BM, int BN, int WGM, int WGN, int PASSES, int STAGES, int SWZ, int SCHED = 0>
static void launch_b3d(const B3Args &p, dim3 grid, hipStream_t s) {
  constexpr int lds = b3d_lds_bytes<BM, BN, PASSES, STAGES>();
  static_assert(lds <= 160 * 1024, "LDS of one CU");
  static const hipError_t attr = hipFuncSetAttribute(
      reinterpret_cast<const void *>(
          &gemm_b3d_kernel<BM, BN, WGM, WGN, PASSES, STAGES, SWZ, SCHED>),
      hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  SISI_HIP(attr);
  hipLaunchKernelGGL((gemm_b3d_kernel<BM, BN, WGM, WGN, PASSES, STAGES, SWZ, SCHED>), grid,
                     dim3(64 * WGM * WGN), lds, s, p);
}

// Every variant reads planes whose rows are padded to kB3RowPad, so planes
// split once (frozen or cached) serve all of them.
constexpr int kB3RowPad = 256;

struct B3Variant {
  int bm, bn;
  void (*launch)(const B3Args &, dim3, hipStream_t);
  const char *name;
};

// Variant tables per pass count; the index is the variant id of
// sisi_gemm_set_variant, the tuning cache and the environment overrides
// SISI_B3_VARIANT / SISI_B1_VARIANT.
static const B3Variant kB3Variants[] = {
    {128, 128, launch_b3p<128, 128, 64, 64, 1, 1, 3>, "lds128x128"},
    {128, 128, launch_b3p<128, 128, 64, 64, 1, 3, 3>, "lds128x128_occ3"},
    {256, 256, launch_b3d<256, 256, 2, 4, 3, 2, 2>, "dma256x256_w8_s2_x2"},
    {256, 256, launch_b3d<256, 256, 2, 4, 3, 2, 1>, "dma256x256_w8_s2_x1"},
    {256, 128, launch_b3d<256, 128, 2, 2, 3, 3, 2>, "dma256x128_w4_s3_x2"},
    {256, 256, launch_b3d<256, 256, 2, 2, 3, 2, 2>, "dma256x256_w4_s2_x2"},
    {256, 256, launch_b3d<256, 256, 2, 4, 3, 2, 2, 1>, "dma256x256_w8_s2_x2_il"},
    {256, 256, launch_b3d<256, 256, 2, 4, 3, 2, 2, 2>, "dma256x256_w8_s2_x2_po"},
};
static const B3Variant kB1Variants[] = {
    {128, 128, launch_b3p<128, 128, 64, 64, 1, 1, 1>, "lds128x128"},
    {128, 128, launch_b3p<128, 128, 64, 64, 1, 2, 1>, "lds128x128_occ2"},
    {256, 256, launch_b3d<256, 256, 2, 4, 1, 4, 2>, "dma256x256_w8_s4_x2"},
    {256, 256, launch_b3d<256, 256, 2, 4, 1, 4, 1>, "dma256x256_w8_s4_x1"},
    {256, 128, launch_b3d<256, 128, 2, 2, 1, 4, 2>, "dma256x128_w4_s4_x2"},
    {256, 256, launch_b3d<256, 256, 2, 2, 1, 4, 2>, "dma256x256_w4_s4_x2"},
    {256, 256, launch_b3d<256, 256, 2, 4, 1, 4, 2, 1>, "dma256x256_w8_s4_x2_il"},
};
constexpr int kNumB3Variants = int(sizeof(kB3Variants) / sizeof(kB3Variants[0]));
constexpr int kNumB1Variants = int(sizeof(kB1Variants) / sizeof(kB1Variants[0]));

static std::atomic<int> &forced_variant(int passes) {
  static std::atomic<int> v3{[] {
    const char *v = getenv("SISI_B3_VARIANT");
    return v != nullptr && *v ? atoi(v) : -1;
  }()};
  static std::atomic<int> v1{[] {
    const char *v = getenv("SISI_B1_VARIANT");
    return v != nullptr && *v ? atoi(v) : -1;
  }()};
  return passes == 1 ? v1 : v3;
}

int b3_variant_count(int passes) { return passes == 1 ? kNumB1Variants : kNumB3Variants; }
const char *b3_variant_name(int passes, int v) {
  if (v < 0 || v >= b3_variant_count(passes)) return "";
  return (passes == 1 ? kB1Variants : kB3Variants)[v].name;
}
void set_b3_variant(int passes, int v) {
  SISI_REQUIRE(v >= -1 && v < b3_variant_count(passes), SISI_ERR_INVALID, "unknown GEMM variant");
  forced_variant(passes).store(v);
}

// Default choice: the 256-row DMA kernel when its tile grid fills the chip,
// else the 128-row kernel (with split-K where the grid stays small).
static int default_b3_variant(const GemmArgs &g, int passes) {
  const int64_t t256 = int64_t((g.M + 255) / 256) * ((g.N + 255) / 256);
  (void)passes;
  return t256 >= 240 ? 2 : 0;
}

// Tile grid, split-K and workspace of one variant; launches the kernel (and
// the split-K reduction) on the engine's stream.
static void launch_variant(Engine &e, GemmArgs g, const B3Variant &v, int xcd, const Planes &A,
                           const Planes &B, int Kp) {
  g.tiles_m = (g.M + v.bm - 1) / v.bm;
  g.tiles_n = (g.N + v.bn - 1) / v.bn;
  g.xcd_map = xcd;
  const int64_t tiles = int64_t(g.tiles_m) * g.tiles_n;
  g.split_out = !(g.dm0 >= g.M && g.dn0 >= g.N);
  int ksplit = 1;
  const int64_t fill = v.bm >= 256 ? 256 : 512;  // workgroups that fill the chip
  if (Kp > 4 * 32 && tiles < fill && !g.out_planes) {
    const int64_t want = (fill + tiles - 1) / tiles;
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
    v.launch(p, grid, e.stream);
    SISI_HIP(hipGetLastError());
  }
  if (ksplit > 1) gemm_reduce(e, g);
}

// Time every variant (both tile walks) of one product class on its real
// planes into a scratch output (FP64, or dense bf16 planes for a
// plane-output product); the fastest is recorded for the class.
static TuneChoice tune_b3(Engine &e, const GemmArgs &g0, int passes, const Planes &A,
                          const Planes &B, int Kp, const TuneKey &key) {
  const B3Variant *table = passes == 1 ? kB1Variants : kB3Variants;
  const int count = b3_variant_count(passes);
  GemmArgs g = g0;
  const int64_t mn = int64_t(g.M) * g.N;
  const int nplanes = g.out_planes == 1 ? 2 : 1;
  DBuf scratch(&e.pool, g.out_planes ? size_t((nplanes * mn * 2 + 7) / 8) : size_t(mn));
  g.C = scratch.get();
  g.cm0 = 1;
  g.cn0 = g.M;
  g.cm1 = g.cm2 = g.cn1 = g.cn2 = 0;
  g.dm0 = g.M;
  g.dn0 = g.N;
  g.dm1 = g.dn1 = 1;
  g.alpha = 1.0;
  g.beta = 0.0;
  g.plane_lo = g.out_planes == 1 ? mn : 0;
  g.out_cube = 0;
  g.cm3 = g.cn3 = 0;
  const bool stats = e.timers->enabled;
  e.timers->enabled = false;
  hipEvent_t t0, t1;
  SISI_HIP(hipEventCreate(&t0));
  SISI_HIP(hipEventCreate(&t1));
  TuneChoice best;
  double best_us = 1e300;
  for (int v = 0; v < count; ++v) {
    for (int xcd = 1; xcd >= 0; --xcd) {
      launch_variant(e, g, table[v], xcd, A, B, Kp);  // warm-up
      const int reps = 3;
      SISI_HIP(hipEventRecord(t0, e.stream));
      for (int r = 0; r < reps; ++r) launch_variant(e, g, table[v], xcd, A, B, Kp);
      SISI_HIP(hipEventRecord(t1, e.stream));
      SISI_HIP(hipEventSynchronize(t1));
      float ms = 0.f;
      SISI_HIP(hipEventElapsedTime(&ms, t0, t1));
      const double us = 1e3 * double(ms) / reps;
      if (us < best_us) {
        best_us = us;
        best.variant = v;
        best.xcd = xcd;
        best.us = us;
      }
    }
  }
  SISI_HIP(hipEventDestroy(t0));
  SISI_HIP(hipEventDestroy(t1));
  e.timers->enabled = stats;
  tune_record(key, best, table[best.variant].name);
  return best;
}

static bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

static void run_b3(Engine &e, GemmArgs g, int passes, const B3FrozenA *fa = nullptr) {
  const int Kp = ((g.K + 31) / 32) * 32;
  const int Mp = ((g.M + kB3RowPad - 1) / kB3RowPad) * kB3RowPad;
  const int Np = ((g.N + kB3RowPad - 1) / kB3RowPad) * kB3RowPad;
  SISI_REQUIRE(!g.out_planes || (g.beta == 0.0 && g.batch == 1), SISI_ERR_INVALID,
               "gemm_b3: plane output needs beta 0 and batch 1");
  if (passes == 1 && fa && (b3_use_library() || b1_use_library()) && !g.out_planes) {
    // frozen hi plane: rows of Kp bf16, the first M rows used
    g.split_out = !(g.dm0 >= g.M && g.dn0 >= g.N);
    Planes la, lb;
    la.hi = reinterpret_cast<const u16 *>(fa->hi());
    la.ld = Kp;
    SISI_REQUIRE(fa->M == g.M && fa->Kp == Kp, SISI_ERR_INVALID, "frozen A planes do not match");
    const Planes &B = split_operand(e, g.B + 0, g.sbn, g.sbk, g.N, g.K, g.N, Kp, lb, 0);
    if (b3_library(e, g, la, B, Kp, 1)) return;
  }
  if ((b3_use_library() || (passes == 1 && b1_use_library())) && !fa && !g.out_planes) {
    g.split_out = !(g.dm0 >= g.M && g.dn0 >= g.N);
    Planes la, lb;
    const int lay_a = passes == 3 ? 1 : 0, lay_b = passes == 3 ? 2 : 0;
    const Planes &A = split_operand(e, g.A + 0, g.sam, g.sak, g.M, g.K, g.M, Kp, la, lay_a);
    const Planes &B = split_operand(e, g.B + 0, g.sbn, g.sbk, g.N, g.K, g.N, Kp, lb, lay_b);
    if (b3_library(e, g, A, B, Kp, passes)) return;
  }
  Planes pa, pb;
  if (fa) {  // planes split earlier in the layout split_operand gives (layout 0)
    SISI_REQUIRE(fa->valid() && fa->M == g.M && fa->K == g.K && fa->Mp == Mp && fa->Kp == Kp,
                 SISI_ERR_INVALID, "frozen A planes do not match the product");
    pa.hi = reinterpret_cast<const u16 *>(fa->hi());
    pa.lo = pa.hi + fa->lo_offset();
    pa.base = pa.hi;
    pa.ld = Kp;
  }
  const int lay = passes == 1 ? kHiOnly : 0;  // plain bf16 reads no lo plane
  const Planes &A = fa ? pa : split_operand(e, g.A + 0, g.sam, g.sak, g.M, g.K, Mp, Kp, pa, lay);
  const Planes &B = split_operand(e, g.B + 0, g.sbn, g.sbk, g.N, g.K, Np, Kp, pb, lay);
  const int count = b3_variant_count(passes);
  int vid = forced_variant(passes).load(), xcd = 1;
  if (vid < 0) {
    const TuneFamily fam = g.out_planes ? (passes == 1 ? TuneFamily::B1P : TuneFamily::B3P)
                                        : (passes == 1 ? TuneFamily::B1 : TuneFamily::B3);
    const TuneKey key = tune_key(fam, g.M, g.N, g.K);
    TuneChoice c = tune_lookup(key);
    if (c.variant < 0 && tune_enabled() && !stream_capturing(e.stream))
      c = tune_b3(e, g, passes, A, B, Kp, key);
    if (c.variant >= 0 && c.variant < count) {
      vid = c.variant;
      xcd = c.xcd;
    }
  }
  if (vid < 0) vid = default_b3_variant(g, passes);
  const B3Variant &v = (passes == 1 ? kB1Variants : kB3Variants)[vid];
  SISI_REQUIRE(int64_t((g.M + v.bm - 1) / v.bm) * v.bm <= Mp &&
                   int64_t((g.N + v.bn - 1) / v.bn) * v.bn <= Np,
               SISI_ERR_INVALID, "gemm_b3: tile grid exceeds the planes");
  launch_variant(e, g, v, xcd, A, B, Kp);
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
  const int Kp = ((K + 31) / 32) * 32, Mp = ((M + kB3RowPad - 1) / kB3RowPad) * kB3RowPad;
  Planes local;
  split_operand(e, A, sam, sak, M, K, Mp, Kp, local, 0, false);
  out.store = std::move(local.store);
  out.M = M;
  out.K = K;
  out.Mp = Mp;
  out.Kp = Kp;
}

void b3_split_into(Engine &e, const double *A, int64_t sam, int64_t sak, int M, int K, int Mp,
                   int Kp, void *hi, void *lo) {
  SISI_REQUIRE(A && hi && lo && M > 0 && K > 0 && Mp >= M && Kp >= K && Kp % 32 == 0,
               SISI_ERR_INVALID, "b3_split_into: bad shape");
  SplitArgs a{A, sam, sak, M, K, Mp, Kp, reinterpret_cast<u16 *>(hi), reinterpret_cast<u16 *>(lo)};
  Timed t(*e.timers, Kind::Bandwidth, 8.0 * double(M) * K + 4.0 * double(Mp) * Kp);
  dim3 grid(unsigned((Kp + 63) / 64), unsigned((Mp + 63) / 64));
  hipLaunchKernelGGL(split_bf16_kernel, grid, dim3(256), 0, e.stream, a);
  SISI_HIP(hipGetLastError());
}

void gemm_b3_frozen_a(Engine &e, GemmArgs g, const B3FrozenA &A, int passes) {
  if (g.M <= 0 || g.N <= 0) return;
  SISI_REQUIRE(g.batch == 1, SISI_ERR_INVALID, "gemm_b3_frozen_a: batch 1 only");
  run_b3(e, g, passes == 1 ? 1 : 3, &A);
}

void gemm_b3(Engine &e, GemmArgs g, int passes) { run_b3(e, g, passes == 1 ? 1 : 3); }

}  // namespace sisi
