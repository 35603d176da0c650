// Device helpers shared by the MFMA GEMM kernels: vector types, the block
// tile, output addressing of split row/column indices, and the XCD-aware
// tile walk.
#pragma once

#include "gemm.hpp"

namespace sisi {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BN = 128;

// Cube-blocked coordinate: (x & 7) * lo + (x >> 3) * hi.
__device__ __forceinline__ int64_t cube_coord(int x, int64_t lo, int64_t hi) {
  return int64_t(x & 7) * lo + int64_t(x >> 3) * hi;
}

__device__ __forceinline__ int64_t out_off_m(const GemmArgs &g, int m) {
  if (g.out_cube) return cube_coord(m % g.dm0, g.cm0, g.cm1) + cube_coord(m / g.dm0, g.cm2, g.cm3);
  if (!g.split_out) return int64_t(m) * g.cm0;
  int m0 = m % g.dm0;
  int t = m / g.dm0;
  int m1 = t % g.dm1;
  int m2 = t / g.dm1;
  return int64_t(m0) * g.cm0 + int64_t(m1) * g.cm1 + int64_t(m2) * g.cm2;
}

__device__ __forceinline__ int64_t out_off_n(const GemmArgs &g, int n) {
  if (g.out_cube) return cube_coord(n % g.dn0, g.cn0, g.cn1) + cube_coord(n / g.dn0, g.cn2, g.cn3);
  if (!g.split_out) return int64_t(n) * g.cn0;
  int n0 = n % g.dn0;
  int t = n / g.dn0;
  int n1 = t % g.dn1;
  int n2 = t / g.dn1;
  return int64_t(n0) * g.cn0 + int64_t(n1) * g.cn1 + int64_t(n2) * g.cn2;
}

// Tile coordinates.  Workgroups are dealt to the 8 XCDs round-robin
// (workgroup id mod 8), and each XCD has its own L2.  With xcd_map the id is
// remapped so XCD x walks the contiguous tile range [x*T/8, (x+1)*T/8); inside
// a range, groups of `group` m-tiles walk the n direction together, so the
// workgroups resident on one XCD form a compact block of tiles that share A
// row panels and B column panels in that XCD's L2.
__device__ __forceinline__ void tile_coords(const GemmArgs &g, int pid, int &tm, int &tn) {
  if (g.xcd_map) {
    const int tiles = g.tiles_m * g.tiles_n;
    const int per = tiles >> 3;
    if (pid < (per << 3)) pid = (pid & 7) * per + (pid >> 3);
  }
  const int group = g.group;
  int per_group = group * g.tiles_n;
  int gid = pid / per_group;
  int first_m = gid * group;
  int gsize = min(g.tiles_m - first_m, group);
  int r = pid % per_group;
  tm = first_m + r % gsize;
  tn = r / gsize;
}

__device__ __forceinline__ void split_bf16(double x, __bf16 &hi, __bf16 &lo) {
  float xf = float(x);
  hi = __bf16(xf);
  lo = __bf16(xf - float(hi));
}

}  // namespace sisi
