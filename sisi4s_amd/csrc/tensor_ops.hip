// Bandwidth kernels.  All are HBM-bound; they are written for coalesced
// access (unit-stride index across consecutive threads on both the read and
// the write side, using an LDS tile when the two unit strides differ).
#include "tensor_ops.hpp"

#include <algorithm>

namespace sisi {

struct CopyArgs {
  double *dst;
  const double *src;
  int nd;
  int64_t len[kMaxOrder];
  int64_t ds[kMaxOrder];
  int64_t ss[kMaxOrder];
  int64_t total;
  double alpha, beta;
  // tiled path: tile dims (0 = dst unit-stride dim, q = src unit-stride dim)
  int q;
  int64_t rest;  // product of the other dims
};

__global__ void copy_generic_kernel(CopyArgs a) {
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < a.total;
       e += int64_t(gridDim.x) * blockDim.x) {
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

// 32 x 32 tile over (dim 0, dim q): read along q (src unit stride), write
// along 0 (dst unit stride).
__global__ void copy_tiled_kernel(CopyArgs a) {
  __shared__ double tile[32][33];
  const int64_t n0 = a.len[0], nq = a.len[a.q];
  const int64_t t0 = (n0 + 31) / 32;
  const int64_t tx = blockIdx.x % t0, tq = blockIdx.x / t0;
  for (int64_t rr = blockIdx.y; rr < a.rest; rr += gridDim.y) {
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
    dim3 grid(unsigned(tiles), unsigned(std::min<int64_t>(a.rest, 65535)));
    hipLaunchKernelGGL(copy_tiled_kernel, grid, dim3(32, 8), 0, e.stream, a);
  } else {
    const int blocks = int(std::min<int64_t>((a.total + 255) / 256, 1 << 16));
    hipLaunchKernelGGL(copy_generic_kernel, dim3(blocks), dim3(256), 0, e.stream, a);
  }
  SISI_HIP(hipGetLastError());
}

__global__ void fill_kernel(double *p, int64_t n, double v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

void zero(Engine &e, double *p, int64_t n) {
  if (n <= 0) return;
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

__global__ void axpby_kernel(int64_t n, double a, const double *x, double b, double *y) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    double v = a * x[i];
    if (b != 0.0) v += b * y[i];
    y[i] = v;
  }
}

void axpby(Engine &e, int64_t n, double a, const double *x, double b, double *y) {
  if (n <= 0) return;
  Timed t(*e.timers, Kind::Bandwidth, (b != 0.0 ? 24.0 : 16.0) * double(n));
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 1 << 16));
  hipLaunchKernelGGL(axpby_kernel, dim3(blocks), dim3(256), 0, e.stream, n, a, x, b, y);
  SISI_HIP(hipGetLastError());
}

__global__ void add_t1t1_kernel(double *X, const double *T2, const double *T1,
                                const double *T1b, double f, int Nv, int No) {
  const int64_t n = int64_t(Nv) * Nv * No * No;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * blockDim.x) {
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

void add_t1t1(Engine &e, double *X, const double *T2, const double *T1, double f, int Nv,
              int No, const double *T1b) {
  const int64_t n = int64_t(Nv) * Nv * No * No;
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

__global__ void dot_partial_kernel(int64_t n, const double *x, const double *y, double *part) {
  double s = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    s += x[i] * y[i];
  s = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void sum_partials_kernel(const double *part, int n, int stride, double *out) {
  // out[c] = sum_b part[b*stride + c] for c = blockIdx.x
  double s = 0.0;
  for (int b = threadIdx.x; b < n; b += blockDim.x) s += part[int64_t(b) * stride + blockIdx.x];
  s = block_sum(s);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
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

__global__ void estimate_kernel(int order, int Nv, int No, double *R, const double *T,
                                double *dT, const double *eps_i, const double *eps_a,
                                double shift, int64_t n, int64_t nelem) {
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * blockDim.x) {
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

void estimate_amplitudes(Engine &e, int order, double *R, const double *T, double *dT,
                         double shift) {
  const int64_t n = order == 2 ? int64_t(e.n1()) : int64_t(e.n2());
  const int64_t nelem = order == 2 ? int64_t(e.n1e()) : int64_t(e.n2e());
  Timed t(*e.timers, Kind::Bandwidth, 32.0 * double(n));
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 1 << 16));
  hipLaunchKernelGGL(estimate_kernel, dim3(blocks), dim3(256), 0, e.stream, order, e.Nv, e.No,
                     R, T, dT, e.eps_i.get(), e.eps_a.get(), shift, n, nelem);
  SISI_HIP(hipGetLastError());
}

__global__ void energy_partial_kernel(int Nv, int No, const double *T1, const double *T2,
                                      const double *V, const double *Vx, double *part) {
  const int64_t n = int64_t(Nv) * Nv * No * No;
  double sd = 0.0, sx = 0.0;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * blockDim.x) {
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
    part[2 * blockIdx.x] = sd;
    part[2 * blockIdx.x + 1] = sx;
  }
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

}  // namespace sisi
