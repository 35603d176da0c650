// Bandwidth kernels.  All are HBM-bound; they are written for coalesced
// access (unit-stride index across consecutive threads on both the read and
// the write side, using an LDS tile when the two unit strides differ).
#include "tensor_ops.hpp"
#include "tape.hpp"

#include <algorithm>

namespace sisi {


// Kernel bodies take their block index and block count explicitly, so the
// grouped launches of the batched solver run them on virtual blocks.
__device__ __forceinline__ void copy_generic_body(const CopyArgs &a, unsigned bx, unsigned gx) {
  for (int64_t e = int64_t(bx) * blockDim.x + threadIdx.x; e < a.total;
       e += int64_t(gx) * blockDim.x) {
    int64_t r = e, od = 0, os = 0;
#pragma unroll
    for (int k = 0; k < kMaxOrder; ++k) {
      if (k < a.nd) {
        const int64_t x = r % a.len[k];
        r /= a.len[k];
        od += x * a.ds[k];
        os += x * a.ss[k];
      }
    }
    double v = (a.alpha != 0.0) ? a.alpha * a.src[os] : 0.0;
    if (a.beta != 0.0) v += a.beta * a.dst[od];
    a.dst[od] = v;
  }
}

__global__ void copy_generic_kernel(CopyArgs a) { copy_generic_body(a, blockIdx.x, gridDim.x); }

// 32 x 32 tile over (dim 0, dim q): read along q (src unit stride), write
// along 0 (dst unit stride).
__device__ __forceinline__ void copy_tiled_body(const CopyArgs &a, unsigned bx, unsigned by,
                                                unsigned gy) {
  __shared__ double tile[32][33];
  const int64_t n0 = a.len[0], nq = a.len[a.q];
  const int64_t t0 = (n0 + 31) / 32;
  const int64_t tx = bx % t0, tq = bx / t0;
  for (int64_t rr = by; rr < a.rest; rr += gy) {
    // decode the other dims
    int64_t r = rr, od = 0, os = 0;
    for (int k = 1; k < a.nd; ++k) {
      if (k == a.q) continue;
      const int64_t x = r % a.len[k];
      r /= a.len[k];
      od += x * a.ds[k];
      os += x * a.ss[k];
    }
    // read: threadIdx.x along q
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
      const int64_t x0 = tx * 32 + j;
      const int64_t xq = tq * 32 + threadIdx.x;
      if (x0 < n0 && xq < nq)
        tile[j][threadIdx.x] = (a.alpha != 0.0) ? a.src[os + x0 * a.ss[0] + xq * a.ss[a.q]] : 0.0;
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
      const int64_t x0 = tx * 32 + threadIdx.x;
      const int64_t xq = tq * 32 + j;
      if (x0 < n0 && xq < nq) {
        const int64_t o = od + x0 * a.ds[0] + xq * a.ds[a.q];
        double v = a.alpha * tile[threadIdx.x][j];
        if (a.beta != 0.0) v += a.beta * a.dst[o];
        a.dst[o] = v;
      }
    }
    __syncthreads();
  }
}

__global__ void copy_tiled_kernel(CopyArgs a) { copy_tiled_body(a, blockIdx.x, blockIdx.y, gridDim.y); }

void copy(Engine &e, const View &dst, const View &src, double alpha, double beta) {
  SISI_REQUIRE(dst.order == src.order, SISI_ERR_INVALID, "copy: order mismatch");
  CopyArgs a{};
  a.dst = dst.p;
  a.src = src.p;
  a.alpha = alpha;
  a.beta = beta;
  // drop unit-length dims
  int nd = 0;
  for (int k = 0; k < dst.order; ++k) {
    SISI_REQUIRE(dst.len[k] == src.len[k], SISI_ERR_INVALID, "copy: length mismatch");
    if (dst.len[k] == 1) continue;
    a.len[nd] = dst.len[k];
    a.ds[nd] = dst.stride[k];
    a.ss[nd] = src.stride[k];
    nd++;
  }
  if (nd == 0) {
    a.len[0] = 1;
    a.ds[0] = a.ss[0] = 0;
    nd = 1;
  }
  // put the dst unit-stride dim first
  int d0 = -1;
  for (int k = 0; k < nd; ++k)
    if (a.ds[k] == 1) {
      d0 = k;
      break;
    }
  if (d0 > 0) {
    std::swap(a.len[0], a.len[d0]);
    std::swap(a.ds[0], a.ds[d0]);
    std::swap(a.ss[0], a.ss[d0]);
  }
  a.nd = nd;
  a.total = 1;
  for (int k = 0; k < nd; ++k) a.total *= a.len[k];
  if (a.total == 0) return;
  int q = -1;
  if (a.ss[0] != 1 && a.ds[0] == 1)
    for (int k = 1; k < nd; ++k)
      if (a.ss[k] == 1) {
        q = k;
        break;
      }
  Timed t(*e.timers, Kind::Bandwidth, (beta != 0.0 ? 24.0 : 16.0) * double(a.total));
  if (q > 0 && a.len[0] >= 8 && a.len[q] >= 8) {
    a.q = q;
    a.rest = a.total / (a.len[0] * a.len[q]);
    const int64_t tiles = ((a.len[0] + 31) / 32) * ((a.len[q] + 31) / 32);
    SISI_REQUIRE(tiles < (int64_t(1) << 31), SISI_ERR_INVALID, "copy: too many tiles");
    dim3 grid(unsigned(tiles), unsigned(std::min<int64_t>(a.rest, e.tape ? 64 : 65535)));
    a.gx = grid.x;
    if (e.tape) {
      TapeOp op;
      op.kind = OpKind::CopyTiled;
      op.gx = grid.x;
      op.gy = grid.y;
      op.c = a;
      e.tape->ops.push_back(op);
      return;
    }
    hipLaunchKernelGGL(copy_tiled_kernel, grid, dim3(32, 8), 0, e.stream, a);
  } else {
    const int blocks =
        int(std::min<int64_t>((a.total + 255) / 256, e.tape ? 256 : int64_t(1) << 16));
    if (e.tape) {
      TapeOp op;
      op.kind = OpKind::CopyGeneric;
      op.gx = unsigned(blocks);
      op.c = a;
      e.tape->ops.push_back(op);
      return;
    }
    hipLaunchKernelGGL(copy_generic_kernel, dim3(blocks), dim3(256), 0, e.stream, a);
  }
  SISI_HIP(hipGetLastError());
}

__device__ __forceinline__ void fill_body(const ElemArgs &a, unsigned bx, unsigned gx) {
  for (int64_t i = int64_t(bx) * blockDim.x + threadIdx.x; i < a.n; i += int64_t(gx) * blockDim.x)
    a.y[i] = a.a;
}

__global__ void fill_kernel(double *p, int64_t n, double v) {
  ElemArgs a;
  a.y = p;
  a.n = n;
  a.a = v;
  fill_body(a, blockIdx.x, gridDim.x);
}

// Records an element-wise launch on the engine's tape; false without a tape.
static bool tape_elem(Engine &e, OpKind kind, const ElemArgs &a, int64_t work) {
  if (!e.tape) return false;
  TapeOp op;
  op.kind = kind;
  op.gx = unsigned(std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, 256)));
  op.e = a;
  e.tape->ops.push_back(op);
  return true;
}

void zero(Engine &e, double *p, int64_t n) {
  if (n <= 0) return;
  ElemArgs ta;
  ta.y = p;
  ta.n = n;
  if (tape_elem(e, OpKind::Fill, ta, n)) return;
  Timed t(*e.timers, Kind::Bandwidth, 8.0 * double(n));
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 1 << 16));
  hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, e.stream, p, n, 0.0);
  SISI_HIP(hipGetLastError());
}

// p[i] = a value in [-1, 1) from an integer hash of (i, seed): operands for
// timing runs (no denormals, no structure a kernel could exploit).
__global__ void fill_pattern_kernel(double *p, int64_t n, uint32_t seed) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    uint64_t x = uint64_t(i) * 0x9E3779B97F4A7C15ull + seed;
    x ^= x >> 31;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 29;
    p[i] = double(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

void fill_pattern(Engine &e, double *p, int64_t n, uint32_t seed) {
  if (n <= 0) return;
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 16384));
  hipLaunchKernelGGL(fill_pattern_kernel, dim3(blocks), dim3(256), 0, e.stream, p, n, seed);
  SISI_HIP(hipGetLastError());
}

void zero(Engine &e, const View &dst) {
  // alpha = 0, beta = 0 writes zeros through the copy kernel's addressing
  copy(e, dst, dst, 0.0, 0.0);
}

// C[R,S,ij] *= V[R0 + R, S0 + S] over planar complex C (Rx x Ry x nij per
// plane) and V (ldV rows, plane size nV); real == 1: real parts only.
__global__ void scale_pairs_kernel(double *C, int64_t Rx, int64_t Ry, int64_t nij, const double *V,
                                   int64_t ldV, int64_t nV, int64_t R0, int64_t S0, int real) {
  const int64_t n = Rx * Ry * nij;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int64_t R = e % Rx, S = (e / Rx) % Ry;
    const int64_t w = (R0 + R) + ldV * (S0 + S);
    const double vr = V[w];
    if (real) {
      C[e] *= vr;
    } else {
      const double vi = V[nV + w], cr = C[e], ci = C[n + e];
      C[e] = cr * vr - ci * vi;
      C[n + e] = cr * vi + ci * vr;
    }
  }
}

void scale_pairs(Engine &e, double *C, int64_t Rx, int64_t Ry, int64_t nij, const double *V,
                 int64_t ldV, int64_t nV, int64_t R0, int64_t S0, bool real) {
  const int64_t n = Rx * Ry * nij;
  if (n <= 0) return;
  Timed t(*e.timers, Kind::Bandwidth, (real ? 16.0 : 32.0) * double(n));
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 1 << 16));
  hipLaunchKernelGGL(scale_pairs_kernel, dim3(blocks), dim3(256), 0, e.stream, C, Rx, Ry, nij, V,
                     ldV, nV, R0, S0, real ? 1 : 0);
  SISI_HIP(hipGetLastError());
}

__device__ __forceinline__ void axpby_body(const ElemArgs &p, unsigned bx, unsigned gx) {
  for (int64_t i = int64_t(bx) * blockDim.x + threadIdx.x; i < p.n;
       i += int64_t(gx) * blockDim.x) {
    double v = p.a * p.x0[i];
    if (p.b != 0.0) v += p.b * p.y[i];
    p.y[i] = v;
  }
}

__global__ void axpby_kernel(int64_t n, double a, const double *x, double b, double *y) {
  ElemArgs p;
  p.n = n;
  p.a = a;
  p.x0 = x;
  p.b = b;
  p.y = y;
  axpby_body(p, blockIdx.x, gridDim.x);
}

void axpby(Engine &e, int64_t n, double a, const double *x, double b, double *y) {
  if (n <= 0) return;
  {
    ElemArgs ta;
    ta.n = n;
    ta.a = a;
    ta.x0 = x;
    ta.b = b;
    ta.y = y;
    if (tape_elem(e, OpKind::Axpby, ta, n)) return;
  }
  Timed t(*e.timers, Kind::Bandwidth, (b != 0.0 ? 24.0 : 16.0) * double(n));
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 1 << 16));
  hipLaunchKernelGGL(axpby_kernel, dim3(blocks), dim3(256), 0, e.stream, n, a, x, b, y);
  SISI_HIP(hipGetLastError());
}

__device__ __forceinline__ void add_t1t1_body(const ElemArgs &p, unsigned bx, unsigned gx) {
  double *X = p.y;
  const double *T2 = p.x0, *T1 = p.x1, *T1b = p.x2;
  const double f = p.a;
  const int Nv = p.i0, No = p.i1;
  const int64_t n = int64_t(Nv) * Nv * No * No;
  for (int64_t e = int64_t(bx) * blockDim.x + threadIdx.x; e < n;
       e += int64_t(gx) * blockDim.x) {
    int64_t r = e;
    const int a = int(r % Nv);
    r /= Nv;
    const int b = int(r % Nv);
    r /= Nv;
    const int i = int(r % No);
    const int j = int(r / No);
    X[e] = T2[e] + f * T1[a + int64_t(Nv) * i] * T1b[b + int64_t(Nv) * j];
  }
}

__global__ void add_t1t1_kernel(double *X, const double *T2, const double *T1,
                                const double *T1b, double f, int Nv, int No) {
  ElemArgs p;
  p.y = X;
  p.x0 = T2;
  p.x1 = T1;
  p.x2 = T1b;
  p.a = f;
  p.i0 = Nv;
  p.i1 = No;
  add_t1t1_body(p, blockIdx.x, gridDim.x);
}

void add_t1t1(Engine &e, double *X, const double *T2, const double *T1, double f, int Nv,
              int No, const double *T1b) {
  const int64_t n = int64_t(Nv) * Nv * No * No;
  {
    ElemArgs ta;
    ta.y = X;
    ta.x0 = T2;
    ta.x1 = T1;
    ta.x2 = T1b ? T1b : T1;
    ta.a = f;
    ta.i0 = Nv;
    ta.i1 = No;
    if (tape_elem(e, OpKind::AddT1T1, ta, n)) return;
  }
  Timed t(*e.timers, Kind::Bandwidth, 16.0 * double(n));
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 1 << 16));
  hipLaunchKernelGGL(add_t1t1_kernel, dim3(blocks), dim3(256), 0, e.stream, X, T2, T1,
                     T1b ? T1b : T1, f, Nv, No);
  SISI_HIP(hipGetLastError());
}

// Two-pass deterministic reduction: fixed grid of partial sums, then one
// block adds the partials in index order.
constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[256];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ void dot_partial_body(const ElemArgs &p, unsigned bx, unsigned gx) {
  double s = 0.0;
  for (int64_t i = int64_t(bx) * blockDim.x + threadIdx.x; i < p.n;
       i += int64_t(gx) * blockDim.x)
    s += p.x0[i] * p.x1[i];
  s = block_sum(s);
  if (threadIdx.x == 0) p.y[bx] = s;
}

__global__ void dot_partial_kernel(int64_t n, const double *x, const double *y, double *part) {
  ElemArgs p;
  p.n = n;
  p.x0 = x;
  p.x1 = y;
  p.y = part;
  dot_partial_body(p, blockIdx.x, gridDim.x);
}

// y[c] = sum_b x0[b * i1 + c] over b < i0, for column c = bx (< i2 blocks)
__device__ __forceinline__ void sum_partials_body(const ElemArgs &p, unsigned bx) {
  double s = 0.0;
  for (int b = threadIdx.x; b < p.i0; b += blockDim.x) s += p.x0[int64_t(b) * p.i1 + bx];
  s = block_sum(s);
  if (threadIdx.x == 0) p.y[bx] = s;
}

__global__ void sum_partials_kernel(const double *part, int n, int stride, double *out) {
  ElemArgs p;
  p.x0 = part;
  p.i0 = n;
  p.i1 = stride;
  p.y = out;
  sum_partials_body(p, blockIdx.x);
}

double dot(Engine &e, int64_t n, const double *x, const double *y) {
  if (n <= 0) return 0.0;
  DBuf part(&e.pool, kRedBlocks + 1);
  {
    Timed t(*e.timers, Kind::Bandwidth, 16.0 * double(n));
    hipLaunchKernelGGL(dot_partial_kernel, dim3(kRedBlocks), dim3(256), 0, e.stream, n, x, y,
                       part.get());
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, e.stream, part.get(),
                       kRedBlocks, 1, part.get() + kRedBlocks);
    SISI_HIP(hipGetLastError());
  }
  double r = 0.0;
  SISI_HIP(hipMemcpyAsync(&r, part.get() + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost,
                          e.stream));
  SISI_HIP(hipStreamSynchronize(e.stream));
  return r;
}

__device__ __forceinline__ void estimate_body(const ElemArgs &p, unsigned bx, unsigned gx) {
  const int order = p.i0, Nv = p.i1, No = p.i2;
  double *R = p.y;
  const double *T = p.x0, *eps_i = p.x2, *eps_a = p.x3;
  double *dT = const_cast<double *>(p.x1);
  const double shift = p.a;
  const int64_t n = p.n, nelem = p.nelem;
  for (int64_t e = int64_t(bx) * blockDim.x + threadIdx.x; e < n;
       e += int64_t(gx) * blockDim.x) {
    double D;
    const int64_t x = e % nelem;  // element within its (real or imaginary) plane
    if (order == 2) {
      const int a = int(x % Nv);
      const int i = int(x / Nv);
      D = eps_a[a] - eps_i[i];
    } else {
      int64_t r = x;
      const int a = int(r % Nv);
      r /= Nv;
      const int b = int(r % Nv);
      r /= Nv;
      const int i = int(r % No);
      const int j = int(r / No);
      D = eps_a[a] - eps_i[i] + eps_a[b] - eps_i[j];
    }
    const double t = T[e];
    const double est = -(R[e] - shift * t) / (D + shift);
    R[e] = est;
    dT[e] = est - t;
  }
}

__global__ void estimate_kernel(int order, int Nv, int No, double *R, const double *T,
                                double *dT, const double *eps_i, const double *eps_a,
                                double shift, int64_t n, int64_t nelem) {
  ElemArgs p;
  p.i0 = order;
  p.i1 = Nv;
  p.i2 = No;
  p.y = R;
  p.x0 = T;
  p.x1 = dT;
  p.x2 = eps_i;
  p.x3 = eps_a;
  p.a = shift;
  p.n = n;
  p.nelem = nelem;
  estimate_body(p, blockIdx.x, gridDim.x);
}

// One block per tile pair {T[32ta.., 32tb.., i, j], T[32tb.., 32ta.., j, i]},
// taken by the block with i < j, or i == j and ta <= tb (the others exit):
// both tiles are staged in LDS, then each element and its partner get the
// same average (an addition of the same two values in either order).
__global__ __launch_bounds__(256) void symmetrize_doubles_kernel(double *T, int Nv, int No) {
  const int ta = int(blockIdx.x), tb = int(blockIdx.y);
  const int x = int(threadIdx.x) & 31, y0 = int(threadIdx.x) >> 5;
  const int64_t nv2 = int64_t(Nv) * Nv;
  __shared__ double s0[32][33], s1[32][33];
  // (i, j) pairs strided by the grid's z extent (block-uniform skips)
  for (int ij = int(blockIdx.z); ij < No * No; ij += int(gridDim.z)) {
    const int i = ij % No, j = ij / No;
    if (i > j || (i == j && ta > tb)) continue;
    double *P = T + nv2 * (i + int64_t(No) * j);  // [a, b] of (i, j) at a + Nv b
    double *Q = T + nv2 * (j + int64_t(No) * i);  // [a, b] of (j, i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int y = y0 + 8 * r;
      const int pa = 32 * ta + x, pb = 32 * tb + y, qa = 32 * tb + x, qb = 32 * ta + y;
      s0[y][x] = (pa < Nv && pb < Nv) ? P[pa + int64_t(Nv) * pb] : 0.0;
      s1[y][x] = (qa < Nv && qb < Nv) ? Q[qa + int64_t(Nv) * qb] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int y = y0 + 8 * r;
      const int pa = 32 * ta + x, pb = 32 * tb + y, qa = 32 * tb + x, qb = 32 * ta + y;
      // partners: P[pa, pb] <-> Q[pb, pa] = s1[x][y], Q[qa, qb] <-> P[qb, qa] = s0[x][y]
      if (pa < Nv && pb < Nv) P[pa + int64_t(Nv) * pb] = 0.5 * (s0[y][x] + s1[x][y]);
      if (qa < Nv && qb < Nv) Q[qa + int64_t(Nv) * qb] = 0.5 * (s1[y][x] + s0[x][y]);
    }
    __syncthreads();  // the tiles are refilled for the next (i, j)
  }
}

void symmetrize_doubles(Engine &e, double *T2, int Nv, int No) {
  SISI_REQUIRE(Nv > 0 && No > 0 && int64_t(No) * No < (int64_t(1) << 31), SISI_ERR_INVALID,
               "symmetrize_doubles: bad shape");
  const unsigned tiles = unsigned((Nv + 31) / 32);
  const unsigned gz = unsigned(std::min<int64_t>(int64_t(No) * No, 65535));
  Timed t(*e.timers, Kind::Bandwidth, 16.0 * double(Nv) * Nv * No * No);
  hipLaunchKernelGGL(symmetrize_doubles_kernel, dim3(tiles, tiles, gz), dim3(256), 0, e.stream,
                     T2, Nv, No);
  SISI_HIP(hipGetLastError());
}

void estimate_amplitudes(Engine &e, int order, double *R, const double *T, double *dT,
                         double shift) {
  const int64_t n = order == 2 ? int64_t(e.n1()) : int64_t(e.n2());
  const int64_t nelem = order == 2 ? int64_t(e.n1e()) : int64_t(e.n2e());
  {
    ElemArgs ta;
    ta.i0 = order;
    ta.i1 = e.Nv;
    ta.i2 = e.No;
    ta.y = R;
    ta.x0 = T;
    ta.x1 = dT;
    ta.x2 = e.eps_i.get();
    ta.x3 = e.eps_a.get();
    ta.a = shift;
    ta.n = n;
    ta.nelem = nelem;
    if (tape_elem(e, OpKind::Estimate, ta, n)) return;
  }
  Timed t(*e.timers, Kind::Bandwidth, 32.0 * double(n));
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 1 << 16));
  hipLaunchKernelGGL(estimate_kernel, dim3(blocks), dim3(256), 0, e.stream, order, e.Nv, e.No,
                     R, T, dT, e.eps_i.get(), e.eps_a.get(), shift, n, nelem);
  SISI_HIP(hipGetLastError());
}

__device__ __forceinline__ void energy_partial_body(const ElemArgs &p, unsigned bx,
                                                    unsigned gx) {
  const int Nv = p.i0, No = p.i1;
  const double *T1 = p.x0, *T2 = p.x1, *V = p.x2, *Vx = p.x3;
  double *part = p.y;
  const int64_t n = int64_t(Nv) * Nv * No * No;
  double sd = 0.0, sx = 0.0;
  for (int64_t e = int64_t(bx) * blockDim.x + threadIdx.x; e < n;
       e += int64_t(gx) * blockDim.x) {
    int64_t r = e;
    const int a = int(r % Nv);
    r /= Nv;
    const int b = int(r % Nv);
    r /= Nv;
    const int i = int(r % No);
    const int j = int(r / No);
    const double tau = T2[e] + T1[a + int64_t(Nv) * i] * T1[b + int64_t(Nv) * j];
    sd += tau * V[e];
    sx += tau * Vx[e];
  }
  sd = block_sum(sd);
  sx = block_sum(sx);
  if (threadIdx.x == 0) {
    part[2 * bx] = sd;
    part[2 * bx + 1] = sx;
  }
}

__global__ void energy_partial_kernel(int Nv, int No, const double *T1, const double *T2,
                                      const double *V, const double *Vx, double *part) {
  ElemArgs p;
  p.i0 = Nv;
  p.i1 = No;
  p.x0 = T1;
  p.x1 = T2;
  p.x2 = V;
  p.x3 = Vx;
  p.y = part;
  energy_partial_body(p, blockIdx.x, gridDim.x);
}

void energy_terms(Engine &e, const double *T1, const double *T2, double &direct,
                  double &exchange) {
  DBuf part(&e.pool, 2 * kRedBlocks + 2);
  {
    Timed t(*e.timers, Kind::Bandwidth, 24.0 * double(e.n2()));
    hipLaunchKernelGGL(energy_partial_kernel, dim3(kRedBlocks), dim3(256), 0, e.stream, e.Nv,
                       e.No, T1, T2, e.Vabij.get(), e.Vbaij.get(), part.get());
    hipLaunchKernelGGL(sum_partials_kernel, dim3(2), dim3(256), 0, e.stream, part.get(),
                       kRedBlocks, 2, part.get() + 2 * kRedBlocks);
    SISI_HIP(hipGetLastError());
  }
  double r[2];
  SISI_HIP(hipMemcpyAsync(r, part.get() + 2 * kRedBlocks, 2 * sizeof(double),
                          hipMemcpyDeviceToHost, e.stream));
  SISI_HIP(hipStreamSynchronize(e.stream));
  direct = 2.0 * r[0];
  exchange = -1.0 * r[1];
}

// Taped deterministic reductions of the batched solver: the sums land in
// device memory (out) and are read back once for all systems.
void dot_into(Engine &e, int64_t n, const double *x, const double *y, double *part,
              double *out) {
  SISI_REQUIRE(e.tape, SISI_ERR_STATE, "dot_into: taped engines only");
  ElemArgs a;
  a.n = n;
  a.x0 = x;
  a.x1 = y;
  a.y = part;
  TapeOp op;
  op.kind = OpKind::DotPartial;
  op.gx = kTapeRedBlocks;
  op.e = a;
  e.tape->ops.push_back(op);
  ElemArgs b;
  b.x0 = part;
  b.i0 = kTapeRedBlocks;
  b.i1 = 1;
  b.i2 = 1;
  b.y = out;
  op.kind = OpKind::SumPartials;
  op.gx = 1;
  op.e = b;
  e.tape->ops.push_back(op);
}

void energy_into(Engine &e, const double *T1, const double *T2, double *part, double *out) {
  SISI_REQUIRE(e.tape, SISI_ERR_STATE, "energy_into: taped engines only");
  ElemArgs a;
  a.i0 = e.Nv;
  a.i1 = e.No;
  a.x0 = T1;
  a.x1 = T2;
  a.x2 = e.Vabij.get();
  a.x3 = e.Vbaij.get();
  a.y = part;
  TapeOp op;
  op.kind = OpKind::EnergyPartial;
  op.gx = kTapeRedBlocks;
  op.e = a;
  e.tape->ops.push_back(op);
  ElemArgs b;
  b.x0 = part;
  b.i0 = kTapeRedBlocks;
  b.i1 = 2;
  b.i2 = 2;
  b.y = out;
  op.kind = OpKind::SumPartials;
  op.gx = 2;
  op.e = b;
  e.tape->ops.push_back(op);
}

// Record p of a grouped launch: the largest p with start[p] <= block.
__device__ __forceinline__ int group_record(const unsigned *start, int n, unsigned block) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= block) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void copy_generic_grouped(const CopyArgs *const *args, const unsigned *start, int n) {
  const int p = group_record(start, n, blockIdx.x);
  const CopyArgs a = *args[p];
  copy_generic_body(a, blockIdx.x - start[p], start[p + 1] - start[p]);
}

__global__ void copy_tiled_grouped(const CopyArgs *const *args, const unsigned *start, int n) {
  const int p = group_record(start, n, blockIdx.x);
  // the record is read in place: a copy indexed by dimension went to scratch
  const CopyArgs &a = *args[p];
  const unsigned local = blockIdx.x - start[p], gx = a.gx;
  const unsigned gy = (start[p + 1] - start[p]) / gx;
  copy_tiled_body(a, local % gx, local / gx, gy);
}

template <int KIND>
__global__ __launch_bounds__(256) void elem_grouped(const ElemArgs *const *args,
                                                    const unsigned *start, int n) {
  const int p = group_record(start, n, blockIdx.x);
  const unsigned bx = blockIdx.x - start[p], gx = start[p + 1] - start[p];
  const ElemArgs a = *args[p];
  if constexpr (KIND == int(OpKind::Fill)) fill_body(a, bx, gx);
  else if constexpr (KIND == int(OpKind::Axpby)) axpby_body(a, bx, gx);
  else if constexpr (KIND == int(OpKind::AddT1T1)) add_t1t1_body(a, bx, gx);
  else if constexpr (KIND == int(OpKind::Estimate)) estimate_body(a, bx, gx);
  else if constexpr (KIND == int(OpKind::DotPartial)) dot_partial_body(a, bx, gx);
  else if constexpr (KIND == int(OpKind::SumPartials)) sum_partials_body(a, bx);
  else if constexpr (KIND == int(OpKind::EnergyPartial)) energy_partial_body(a, bx, gx);
}

void launch_copy_grouped(OpKind kind, const CopyArgs *const *args, const unsigned *start, int n,
                         unsigned total, hipStream_t s) {
  if (kind == OpKind::CopyTiled)
    hipLaunchKernelGGL(copy_tiled_grouped, dim3(total), dim3(32, 8), 0, s, args, start, n);
  else
    hipLaunchKernelGGL(copy_generic_grouped, dim3(total), dim3(256), 0, s, args, start, n);
  SISI_HIP(hipGetLastError());
}

void launch_elem_grouped(OpKind kind, const ElemArgs *const *args, const unsigned *start, int n,
                         unsigned total, hipStream_t s) {
  switch (kind) {
#define SISI_ELEM_CASE(K)                                                                      \
  case OpKind::K:                                                                            \
    hipLaunchKernelGGL(elem_grouped<int(OpKind::K)>, dim3(total), dim3(256), 0, s, args, start, \
                       n);                                                                   \
    break;
    SISI_ELEM_CASE(Fill)
    SISI_ELEM_CASE(Axpby)
    SISI_ELEM_CASE(AddT1T1)
    SISI_ELEM_CASE(Estimate)
    SISI_ELEM_CASE(DotPartial)
    SISI_ELEM_CASE(SumPartials)
    SISI_ELEM_CASE(EnergyPartial)
#undef SISI_ELEM_CASE
    default: throw Error(SISI_ERR_INVALID, "grouped launch: not an element-wise kind");
  }
  SISI_HIP(hipGetLastError());
}

}  // namespace sisi
