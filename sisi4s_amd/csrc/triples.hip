// (T) triples (CcsdPerturbativeTriples.cxx:119-248) on the device.
//
// For each sorted block i = (i0 <= i1 <= i2) the reference forms, for the six
// permutations pi of (i0,i1,i2), with (x,y,z) = i.pi,
//     W(x,y,z)_abc = sum_d T_{ad,xy} V_{bdcz} - sum_l T_{ab,xl} V_{yzlc}
// and reads them back with "abc"*pi (cxx:159-181).  Here
//     W[x] = sum_{distinct pi} W(i.pi)[x_pi(0), x_pi(1), x_pi(2)]
// is built by three segmented GEMMs, one per output index j that plays the
// GEMM row: the particle terms of the permutations with pi(0) = j (K = Nv
// each) and the hole terms of those with pi(2) = j (K = No each) are summed in
// the same registers, the two orders of the remaining pair (p, q) coming from
// transposed copies of V and T2.  The reference energy (cxx:183-214),
//     E = sum_{distinct pi} sum_x DV[x] sum_sigma sf(sigma) Z_pi[("abc" sigma pi)(x)],
// DV = Wsym / D, Wsym the sum over all six permutations, Z_pi = W(i.pi) + S(i.pi),
// is rewritten by relabelling the summation index as
//     E = sum_x F[x] Z[x],   F[x] = (1/D[x]) sum_sigma sf'(sigma) W[x.sigma],
//     Z[x] = W[x] + sum_{distinct pi} S(i.pi)[x.pi],
// with sf = +8 (identity), -4 (transpositions), +2 (3-cycles).  When all three
// indices differ every permutation is distinct, W = Wsym and sf' = sf.  When
// two are equal, Wsym[x] = W[x] + W[x.tau] with tau the swap of the equal
// positions, so sf'(sigma) = sf(sigma) + sf(sigma.tau) = +4 for sigma in
// {id, tau}, -2 otherwise; when all three are equal sf' = 0 and the block
// contributes nothing (as in the reference, whose i=j=k terms cancel).
// W is stored cube-blocked (each 8x8x8 cube of indices contiguous, written so
// by the GEMM epilogues).  The energy kernel walks orbits of cubes under
// index permutations: the (up to) six permuted cubes of an orbit are staged
// in LDS once and every cube of the orbit is evaluated from them, so W is
// read once, in whole 4 KB cubes; the kernel is persistent and fetches the
// next orbit while it evaluates the current one; the sum is reduced in FP64.
#include "triples.hpp"

#include "contract.hpp"
#include "tensor_ops.hpp"
#include "ztensor.hpp"

namespace sisi {

namespace {

// Permutation<3>(p) of math/Permutation.hpp:52-62 (recursive construction).
std::array<int, 3> permutation3(int p) {
  // Permutation<2>(q): q = 0 -> (0,1), q = 1 -> (1,0)
  const int q = p / 3;
  const int sub[2] = {q == 0 ? 0 : 1, q == 0 ? 1 : 0};
  std::array<int, 3> im{};
  const int pos = p % 3;
  int i = 0;
  for (; i < pos; ++i) im[i] = sub[i] + 1;
  im[pos] = 0;
  for (++i; i < 3; ++i) im[i] = sub[i - 1] + 1;
  return im;
}

struct PermTerm {
  int pi[3];
  int idx[3];
};

struct EnergyArgs {
  const double *W;
  const double *T1;  // [Nv, No]
  const double *V;   // Vabij [Nv,Nv,No,No]
  // optional second singles product sb * T1b V_b (complex amplitudes: the
  // Re/Im part of T1 V needs two products of planes); T1b == nullptr: none
  const double *T1b;
  const double *Vb;
  double sb;
  const double *eps_a;
  const int *orbits;  // X0 | X1 << 10 | X2 << 20, X0 <= X1 <= X2
  double *part;
  double eps_ijk;
  double sfp[6];
  int Nv, No, norbits;
  int tmask;          // bit p: permutation p (kPerm3[p]) is a distinct singles term
  PermTerm terms[6];  // by permutation index p (pi = kPerm3[p], idx = i.pi)
};

// Permutation<3>(p) for p = 0..5 (math/Permutation.hpp:52-62), the order of
// permutation3(); the energy kernel unrolls over it.
constexpr int kPerm3[6][3] = {{0, 1, 2}, {1, 0, 2}, {1, 2, 0}, {0, 2, 1}, {2, 0, 1}, {2, 1, 0}};

// sigma as reads (x_s0, x_s1, x_s2): identity, three transpositions, two 3-cycles
constexpr int hSig[6][3] = {{0, 1, 2}, {1, 0, 2}, {2, 1, 0}, {0, 2, 1}, {1, 2, 0}, {2, 0, 1}};
constexpr double hSf[6] = {8.0, -4.0, -4.0, -4.0, 2.0, 2.0};

// index r of the composition sig[r][k] = sig[a][sig[b][k]]
constexpr int compose_idx(int a, int b) {
  int r = 0;
  for (; r < 6; ++r) {
    bool ok = true;
    for (int k = 0; k < 3; ++k) ok = ok && (hSig[r][k] == hSig[a][hSig[b][k]]);
    if (ok) break;
  }
  return r;
}

__device__ __forceinline__ double block_sum_t(double v) {
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

// LDS layout of one 8x8x8 cube: (a, b, c) at 64 c + 8 b + (a ^ ((b + c) & 7)).
// A wave holds one l2 plane of 8 x 8 elements; reading the cube permuted puts
// the lane-varying indices at strides 1, 8 or 64, and the XOR spreads the
// strided ones over the banks (the pair (2k, 2k+1) stays a pair, so the two
// doubles a lane loads land in one aligned 16-byte slot).
__device__ __forceinline__ int sw_off(int a, int b, int c) {
  return 64 * c + 8 * b + (a ^ ((b + c) & 7));
}

// composition sig[q] o sig[t] as an index of hSig (folded at compile time
// for unrolled q, t)
__host__ __device__ constexpr int comp_idx(int q, int t) {
  int r = 0;
  for (; r < 6; ++r) {
    bool ok = true;
    for (int k = 0; k < 3; ++k) ok = ok && (hSig[r][k] == hSig[q][hSig[t][k]]);
    if (ok) break;
  }
  return r;
}

__device__ __forceinline__ int sel3(int a, int b, int c, int i) {
  return i == 0 ? a : (i == 1 ? b : c);
}

// Persistent kernel over the orbits of 8x8x8 cubes under index permutations
// (workgroup b takes orbits b, b + grid, ...).  W is cube-blocked, so each of
// the six permuted cubes X.sigma_r of an orbit is one contiguous 4 KB read,
// 16 bytes per lane.  The loads of the next orbit (the cubes, and for each
// cube Y = X.sigma_q and singles term t the factors st[q][t][l] =
// T1[8 Y_pi0 + l, i_pi0] / 2 and sv[q][t][8 l1 + l2] = V[8 Y_pi1 + l1,
// 8 Y_pi2 + l2, i_pi1, i_pi2], whose product is S(i.pi)[x.pi], and the
// particle energies of the orbit's cube ranges) are issued into registers
// before the current orbit is evaluated from LDS, and written to LDS after
// it.  Every element of every distinct cube: F = (sum_sigma sf'(sigma)
// W[x.sigma]) / D times Z = W + sum_pi S_pi.  TWO: a second singles product
// (complex amplitudes).  Partial sums per workgroup in a fixed order.
template <bool TWO>
__global__ __launch_bounds__(256) void triples_energy_kernel(EnergyArgs g) {
  __shared__ double tile[6][512];
  // V blocks as 8 rows of 9 (one pad double per row: reads along the rows'
  // first index hit distinct banks)
  __shared__ double sv[6][6][72], st[6][6][8];
  __shared__ double sv2[TWO ? 6 : 1][6][72], st2[TWO ? 6 : 1][6][8];
  __shared__ double ea[3][8];
  const int Nv = g.Nv, tid = threadIdx.x, norbits = g.norbits, tmask = g.tmask;
  const int C = (Nv + 7) >> 3;
  const int64_t s1 = Nv, s2 = int64_t(Nv) * Nv;
  // registers of the orbit being fetched; the singles factors are fetched per
  // (cube q, permutation t) pair, pair q * 6 + t = wave + 4 k for k < 9, so
  // that the pair's indices are uniform across the wave (scalar operands)
  double rt[6][2];
  double rsv[9], rst[9], rsv2[TWO ? 9 : 1], rst2[TWO ? 9 : 1], rea = 0.0;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  // the orbit-independent part of pair k's singles loads: which orbit index
  // X_c gives each of the three cube coordinates (c = sigma_q(pi_k)), the
  // occupied-index offsets into V and T1, and whether the term is used
  // (packed: scode = c0 | c1 << 2 | c2 << 4 | used << 6)
  int scode[9], stoff[9];
  int64_t svoff[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const int pair = wave + 4 * k, q = pair / 6, t = pair - 6 * q;
    const PermTerm &p = g.terms[t];
    scode[k] = hSig[q][p.pi[0]] | (hSig[q][p.pi[1]] << 2) | (hSig[q][p.pi[2]] << 4) |
               (((tmask >> t) & 1) << 6);
    svoff[k] = s2 * (p.idx[1] + int64_t(g.No) * p.idx[2]);
    stoff[k] = Nv * p.idx[0];
  }

  auto decode = [](int code, int &X0, int &X1, int &X2) {
    X0 = code & 1023;
    X1 = (code >> 10) & 1023;
    X2 = code >> 20;
  };
  auto fetch = [&](int code) {
    int X0, X1, X2;
    decode(code, X0, X1, X2);
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      const int64_t base = 512 * (int64_t(sel3(X0, X1, X2, hSig[r][0])) +
                                  int64_t(C) * (sel3(X0, X1, X2, hSig[r][1]) +
                                                int64_t(C) * sel3(X0, X1, X2, hSig[r][2])));
      const double2 w2 = *reinterpret_cast<const double2 *>(g.W + base + 2 * tid);
      rt[r][0] = w2.x;
      rt[r][1] = w2.y;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      double x = 0.0, x2 = 0.0, y = 0.0, y2 = 0.0;
      if ((scode[k] >> 6) & 1) {
        const int l1 = lane >> 3, l2 = lane & 7;
        const int u1 = sel3(X0, X1, X2, (scode[k] >> 2) & 3) * 8 + l1;
        const int u2 = sel3(X0, X1, X2, (scode[k] >> 4) & 3) * 8 + l2;
        if (u1 < Nv && u2 < Nv) {
          const int64_t vv = u1 + s1 * u2 + svoff[k];
          x = g.V[vv];
          if (TWO) x2 = g.Vb[vv];
        }
        const int u0 = sel3(X0, X1, X2, scode[k] & 3) * 8 + lane;
        if (lane < 8 && u0 < Nv) {
          const int64_t t1 = u0 + int64_t(stoff[k]);
          y = 0.5 * g.T1[t1];
          if (TWO) y2 = 0.5 * g.sb * g.T1b[t1];
        }
      }
      rsv[k] = x;
      rst[k] = y;
      if (TWO) {
        rsv2[k] = x2;
        rst2[k] = y2;
      }
    }
    if (tid < 24) {
      const int k = tid >> 3, l = tid & 7, u = sel3(X0, X1, X2, k) * 8 + l;
      rea = u < Nv ? g.eps_a[u] : 0.0;
    }
  };
  auto stash = [&]() {
    const int e0 = 2 * tid, a = e0 & 7, b = (e0 >> 3) & 7, c = e0 >> 6;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      tile[r][sw_off(a, b, c)] = rt[r][0];
      tile[r][sw_off(a + 1, b, c)] = rt[r][1];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int pair = wave + 4 * k, q = pair / 6, t = pair - 6 * q;
      const int pos = 9 * (lane >> 3) + (lane & 7);
      sv[q][t][pos] = rsv[k];
      if (TWO) sv2[q][t][pos] = rsv2[k];
      if (lane < 8) {
        st[q][t][lane] = rst[k];
        if (TWO) st2[q][t][lane] = rst2[k];
      }
    }
    if (tid < 24) ea[tid >> 3][tid & 7] = rea;
  };

  // this thread's elements e = tid + 256 h of a cube and, per sigma_t, the
  // LDS offset of the element x.sigma_t (the same for every orbit)
  int lx[2][3], off[2][6];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int e = tid + 256 * h;
    lx[h][0] = e & 7;
    lx[h][1] = (e >> 3) & 7;
    lx[h][2] = e >> 6;
#pragma unroll
    for (int t = 0; t < 6; ++t)
      off[h][t] = sw_off(lx[h][hSig[t][0]], lx[h][hSig[t][1]], lx[h][hSig[t][2]]);
  }
  double acc = 0.0;
  int o = blockIdx.x;
  if (o < norbits) fetch(g.orbits[o]);
  while (o < norbits) {
    const int code = g.orbits[o];
    __syncthreads();  // the previous orbit's reads of LDS are done
    stash();
    __syncthreads();
    const int nxt = o + int(gridDim.x);
    if (nxt < norbits) fetch(g.orbits[nxt]);  // in flight while this orbit is evaluated
    int X0, X1, X2;
    decode(code, X0, X1, X2);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      // the cube Y = X.sigma_q, once per distinct cube of the orbit (uniform)
      const int Y0 = sel3(X0, X1, X2, hSig[q][0]), Y1 = sel3(X0, X1, X2, hSig[q][1]),
                Y2 = sel3(X0, X1, X2, hSig[q][2]);
      bool dup = false;
#pragma unroll
      for (int p = 0; p < q; ++p)
        dup = dup || (sel3(X0, X1, X2, hSig[p][0]) == Y0 && sel3(X0, X1, X2, hSig[p][1]) == Y1 &&
                      sel3(X0, X1, X2, hSig[p][2]) == Y2);
      if (dup) continue;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int l0 = lx[h][0], l1 = lx[h][1], l2 = lx[h][2];
        if (Y0 * 8 + l0 >= Nv || Y1 * 8 + l1 >= Nv || Y2 * 8 + l2 >= Nv) continue;
        double f = 0.0;
#pragma unroll
        for (int t = 0; t < 6; ++t) f += g.sfp[t] * tile[comp_idx(q, t)][off[h][t]];
        // axis k of Y is axis sigma_q(k) of X
        const double D =
            g.eps_ijk - ea[hSig[q][0]][l0] - ea[hSig[q][1]][l1] - ea[hSig[q][2]][l2];
        double z = tile[q][off[h][0]];
#pragma unroll
        for (int t = 0; t < 6; ++t) {
          if (!((tmask >> t) & 1)) continue;
          const int a = lx[h][kPerm3[t][0]];
          const int b = 9 * lx[h][kPerm3[t][1]] + lx[h][kPerm3[t][2]];
          z += st[q][t][a] * sv[q][t][b];
          if (TWO) z += st2[q][t][a] * sv2[q][t][b];
        }
        acc += f * z / D;
      }
    }
    o = nxt;
  }
  acc = block_sum_t(acc);
  if (tid == 0) g.part[blockIdx.x] = acc;
}

__global__ void sum_into_kernel(const double *part, int n, double *out) {
  double s = 0.0;
  for (int b = threadIdx.x; b < n; b += blockDim.x) s += part[b];
  s = block_sum_t(s);
  if (threadIdx.x == 0) *out = s;
}

// Operands of one (T) block, all device pointers.
struct TriplesOps {
  const double *T2, *T2t;      // T2[a,b,i,j] and T2t[a,b,i,j] = T2[b,a,i,j]
  const double *Vt, *Vt2;      // Vt[d,b,c,k] = V_bdck, Vt2[d,b,c,k] = Vt[d,c,b,k]
  const double *VP;            // VP[l,c,j,k] = V_jklc
};

// Output addressing of the W products for output index j as the GEMM row
// (j = 0: rows x0, columns x1 + v x2; else rows x_p + v x_q, columns x_j):
// W is stored as contiguous 8 x 8 x 8 cubes (w_cube_size), element (x0, x1,
// x2) at 512 (X0 + C X1 + C^2 X2) + l0 + 8 l1 + 64 l2 with X = x >> 3,
// l = x & 7, C = ceil(v / 8) cubes per index.
void w_cube_addressing(GemmArgs &g, int v, int j, int p, int q) {
  const int64_t C = (v + 7) / 8;
  const int64_t lo[3] = {1, 8, 64}, hi[3] = {512, 512 * C, 512 * C * C};
  g.out_cube = 1;
  g.dm0 = v;
  g.dn0 = v;
  if (j == 0) {
    g.cm0 = lo[0];
    g.cm1 = hi[0];
    g.cn0 = lo[1];
    g.cn1 = hi[1];
    g.cn2 = lo[2];
    g.cn3 = hi[2];
  } else {
    g.cm0 = lo[p];
    g.cm1 = hi[p];
    g.cm2 = lo[q];
    g.cm3 = hi[q];
    g.cn0 = lo[j];
    g.cn1 = hi[j];
  }
}

// Doubles of a cube-blocked W of v^3 elements.
int64_t w_cube_size(int64_t v) {
  const int64_t C = (v + 7) / 8;
  return 512 * C * C * C;
}

// W[x] = sum over the given permutations of W(i.pi)[x.pi], three launches.
// The first factor of every product (T2 of the particle terms, VP of the hole
// terms) comes from opA, the second from opB, the products are scaled by
// `sign`, and W is overwritten unless `accumulate` is set (complex amplitudes
// combine planes this way; the real path uses opA = opB, +1, overwrite).
void build_w(Engine &e, int prec, const TriplesOps &opA, const TriplesOps &opB,
             const std::array<int, 3> &ib, const std::array<std::array<int, 3>, 6> &perms,
             const bool *use, double *W, double sign = 1.0, bool accumulate = false) {
  const int64_t o = e.No, v = e.Nv;
  bool first = !accumulate;
  for (int j = 0; j < 3; ++j) {
    const int p = j == 0 ? 1 : 0, q = j == 2 ? 1 : 2;  // the other two positions, p < q
    SegGemmArgs a;
    for (int n = 0; n < 6; ++n) {
      if (!use[n]) continue;
      const auto &pi = perms[n];
      const int x = ib[pi[0]], y = ib[pi[1]], z = ib[pi[2]];
      if (pi[0] == j) {
        // + sum_d T2[x_j, d, x, y] V[d, x_pi1, x_pi2, z]
        GemmSeg &sg = a.seg[a.nseg++];
        sg.A = opA.T2 + (x + o * y) * v * v;
        sg.sam = 1;
        sg.sak = v;
        sg.B = (pi[1] == p ? opB.Vt : opB.Vt2) + z * v * v * v;
        sg.sbk = 1;
        sg.sbn = v;
        sg.K = int(v);
        sg.scale = sign;
      }
      if (pi[2] == j) {
        // - sum_l V_P[l, x_j, y, z] T2[x_pi0, x_pi1, x, l]
        GemmSeg &sg = a.seg[a.nseg++];
        sg.A = opA.VP + (y + o * z) * o * v;
        sg.sam = o;
        sg.sak = 1;
        sg.B = (pi[0] == p ? opB.T2 : opB.T2t) + x * v * v;
        sg.sbk = v * v * o;
        sg.sbn = 1;
        sg.K = int(o);
        sg.scale = -sign;
      }
    }
    if (a.nseg == 0) continue;
    GemmArgs &g = a.g;
    g.C = W;
    g.alpha = 1.0;
    g.beta = first ? 0.0 : 1.0;
    first = false;
    if (j == 0) {
      // rows x0, columns (x1, x2) as n = x1 + v x2
      g.M = int(v);
      g.N = int(v * v);
    } else {
      // rows and columns exchanged so that lanes run along x0: C^T = B^T A^T
      for (int s = 0; s < a.nseg; ++s) {
        GemmSeg &sg = a.seg[s];
        std::swap(sg.A, sg.B);
        const int64_t sam = sg.sbn, sak = sg.sbk, sbk = sg.sak, sbn = sg.sam;
        sg.sam = sam;
        sg.sak = sak;
        sg.sbk = sbk;
        sg.sbn = sbn;
      }
      g.M = int(v * v);
      g.N = int(v);
    }
    w_cube_addressing(g, int(v), j, p, q);
    gemm_segments(e, prec, a);
  }
}

// The operands of the block GEMMs as split-BF16x3 planes.
struct TriplesPlanes {
  const B3Planes *V, *V2;  // rows (b,c) of V_bdck, k = d: [d,b,c] and [d,c,b] orders
  const B3Planes *VP;      // rows c, k = l of -V_jklc, one matrix per (j,k)
  B3Planes T2p;            // rows a, k = d of T2[a,d,x,y], one matrix per (x,y)
  B3Planes T2h, T2th;      // rows (a,b), k = l of T2[a,b,x,l] and T2[b,a,x,l], per x
};

// build_w on pre-split planes (split-BF16x3, FP32 accumulation).
void build_w_b3(Engine &e, const TriplesPlanes &pl, const std::array<int, 3> &ib,
                const std::array<std::array<int, 3>, 6> &perms, const bool *use, double *W) {
  const int64_t o = e.No, v = e.Nv;
  bool first = true;
  for (int j = 0; j < 3; ++j) {
    const int p = j == 0 ? 1 : 0, q = j == 2 ? 1 : 2;
    B3SegArgs a;
    for (int n = 0; n < 6; ++n) {
      if (!use[n]) continue;
      const auto &pi = perms[n];
      const int x = ib[pi[0]], y = ib[pi[1]], z = ib[pi[2]];
      if (pi[0] == j) {  // + T2[x_j, d, x, y] V[d, x_pi1, x_pi2, z]
        B3Seg &sg = a.seg[a.nseg++];
        const B3Planes &Vz = pi[1] == p ? *pl.V : *pl.V2;
        sg.Ah = pl.T2p.hi_of(x + o * y);
        sg.Al = pl.T2p.lo_of(x + o * y);
        sg.Bh = Vz.hi_of(z);
        sg.Bl = Vz.lo_of(z);
        sg.Kp = int(pl.T2p.Kp);
      }
      if (pi[2] == j) {  // - V_P[l, x_j, y, z] T2[x_pi0, x_pi1, x, l] (sign in the planes)
        B3Seg &sg = a.seg[a.nseg++];
        const B3Planes &Tx = pi[0] == p ? pl.T2h : pl.T2th;
        sg.Ah = pl.VP->hi_of(y + o * z);
        sg.Al = pl.VP->lo_of(y + o * z);
        sg.Bh = Tx.hi_of(x);
        sg.Bl = Tx.lo_of(x);
        sg.Kp = int(pl.VP->Kp);
      }
    }
    if (a.nseg == 0) continue;
    GemmArgs &g = a.g;
    g.C = W;
    g.alpha = 1.0;
    g.beta = first ? 0.0 : 1.0;
    first = false;
    if (j == 0) {
      g.M = int(v);
      g.N = int(v * v);
    } else {  // C^T = B^T A^T so lanes run along x0
      for (int s = 0; s < a.nseg; ++s) {
        B3Seg &sg = a.seg[s];
        std::swap(sg.Ah, sg.Bh);
        std::swap(sg.Al, sg.Bl);
      }
      g.M = int(v * v);
      g.N = int(v);
    }
    w_cube_addressing(g, int(v), j, p, q);
    gemm_b3_segments(e, a);
  }
}

// The permutations of a block: tup[p] = i.pi_p and whether pi_p gives a
// tuple no earlier permutation gave (CcsdPerturbativeTriples.cxx:161-176).
void block_perms(const std::array<int, 3> &ib, const std::array<std::array<int, 3>, 6> &perms,
                 std::array<int, 3> tup[6], bool distinct[6]) {
  for (int p = 0; p < 6; ++p) {
    for (int k = 0; k < 3; ++k) tup[p][k] = ib[perms[p][k]];  // i * pi (Permutation.hpp:93-98)
    int q = 0;
    for (; q < p; ++q)
      if (tup[q] == tup[p]) break;
    distinct[p] = (q == p);
  }
}

// Block-dependent fields of the energy kernel: eps_i+eps_j+eps_k, the
// sign factors sf' and the singles terms of the distinct permutations.
void block_energy_args(EnergyArgs &g, const std::array<int, 3> &ib,
                       const std::array<std::array<int, 3>, 6> &perms,
                       const std::array<int, 3> tup[6], const bool distinct[6],
                       const std::vector<double> &eps_i) {
  g.eps_ijk = eps_i[ib[0]] + eps_i[ib[1]] + eps_i[ib[2]];
  // tau: the swap of the equal positions (sig index 1 swaps 0,1; 3 swaps 1,2)
  const int tau = ib[0] == ib[1] ? 1 : ib[1] == ib[2] ? 3 : -1;
  for (int s = 0; s < 6; ++s) g.sfp[s] = hSf[s] + (tau >= 0 ? hSf[compose_idx(s, tau)] : 0.0);
  g.tmask = 0;
  for (int p = 0; p < 6; ++p) {
    PermTerm &t = g.terms[p];
    for (int k = 0; k < 3; ++k) {
      SISI_REQUIRE(perms[p][k] == kPerm3[p][k], SISI_ERR_STATE, "triples: permutation table");
      t.pi[k] = perms[p][k];
      t.idx[k] = tup[p][k];
    }
    if (distinct[p]) g.tmask |= 1 << p;
  }
}

std::vector<std::array<int, 3>> checked_blocks(const Engine &e,
                                               const std::vector<std::array<int, 3>> &in) {
  const int o = e.No;
  std::vector<std::array<int, 3>> blocks = in;
  if (blocks.empty())
    for (int a = 0; a < o; ++a)
      for (int b = a; b < o; ++b)
        for (int c = b; c < o; ++c) blocks.push_back({a, b, c});
  for (auto &b : blocks)
    SISI_REQUIRE(b[0] >= 0 && b[0] <= b[1] && b[1] <= b[2] && b[2] < o, SISI_ERR_INVALID,
                 "triples: blocks must be sorted i0 <= i1 <= i2 < No");
  return blocks;
}

// Workgroups of the persistent energy kernel: as many as are resident at
// once (three per CU, two with the second singles product's LDS), at most
// one per orbit.
int energy_grid(const Engine &e, int norbits, bool two) {
  return std::max(1, std::min(norbits, (two ? 2 : 3) * e.cus));
}

// Cube orbits X0 <= X1 <= X2 of the energy kernel, uploaded.
DBuf orbit_codes(Engine &e, int &norbits) {
  const int ncube = int((e.Nv + 7) / 8);
  std::vector<int> codes;
  for (int a = 0; a < ncube; ++a)
    for (int b = a; b < ncube; ++b)
      for (int c = b; c < ncube; ++c) codes.push_back(a | (b << 10) | (c << 20));
  norbits = int(codes.size());
  DBuf buf(&e.pool, (norbits + 1) / 2);
  SISI_HIP(hipMemcpyAsync(buf.get(), codes.data(), sizeof(int) * norbits, hipMemcpyHostToDevice,
                          e.stream));
  return buf;
}

// CcsdPerturbativeTriplesComplex::Calculator<complex>::calculate
// (CcsdPerturbativeTriplesComplex.cxx:172-279) on planar complex tensors.
// W = sum_d T2 V_t - sum_l T2 V_P is complex (four plane passes of build_w);
// the energy sum_x conj(DV/D)[x] Tabc[x] has real part E(Re) + E(Im), the
// real-path energy functional evaluated on the Re planes of W and the
// singles term plus the same on the Im planes, since every step between W
// and the energy is linear with real coefficients.
double ztriples_energy(Engine &e, const double *T1, const double *T2, int prec,
                       const std::vector<std::array<int, 3>> &blocks_in) {
  const int64_t o = e.No, v = e.Nv, G = e.NG;
  const int64_t n2 = v * v * o * o, nvt = v * v * v * o, nvp = o * v * o * o;
  std::vector<std::array<int, 3>> blocks = checked_blocks(e, blocks_in);
  if (!e.Vt.get()) {
    // V_t[d,b,c,k] = sum_F conj(Gamma_Fdb) Gamma_Fck (cxx:373-375)
    e.Vt = DBuf(&e.pool, 2 * nvt);
    zctr(e, 1.0, "Fdb,Fck->dbck", zdense(e.Gab.get(), {G, v, v}), true,
         zdense(e.Gai.get(), {G, v, o}), false, 0.0, zdense(e.Vt.get(), {v, v, v, o}), SISI_FP64);
    e.Vt2 = DBuf(&e.pool, 2 * nvt);
    zsum(e, 1.0, zdense(e.Vt.get(), {v, v, v, o}), "dbck", false, 0.0,
         zdense(e.Vt2.get(), {v, v, v, o}), "dcbk");
    // hole term V_PHHH[c,l,k,j] (cxx:157-162) as VP[l,c,j,k]
    e.VP = DBuf(&e.pool, 2 * nvp);
    zsum(e, 1.0, zdense(e.Vaijk.get(), {v, o, o, o}), "clkj", false, 0.0,
         zdense(e.VP.get(), {o, v, o, o}), "lcjk");
  }
  DBuf T2t(&e.pool, 2 * n2);
  zsum(e, 1.0, zdense(const_cast<double *>(T2), {v, v, o, o}), "abij", false, 0.0,
       zdense(T2t.get(), {v, v, o, o}), "baij");
  std::vector<double> eps_i(o);
  e.download(eps_i.data(), e.eps_i, size_t(o));
  const TriplesOps re{T2, T2t.get(), e.Vt.get(), e.Vt2.get(), e.VP.get()};
  const TriplesOps im{T2 + n2, T2t.get() + n2, e.Vt.get() + nvt, e.Vt2.get() + nvt,
                      e.VP.get() + nvp};
  int norbits = 0;
  DBuf orbits = orbit_codes(e, norbits);
  const int64_t v3 = w_cube_size(v);
  DBuf W(&e.pool, 2 * v3);
  DBuf part(&e.pool, norbits), Eb(&e.pool, 2 * blocks.size());
  zero(e, Eb.get(), int64_t(2 * blocks.size()));
  std::array<std::array<int, 3>, 6> perms;
  for (int p = 0; p < 6; ++p) perms[p] = permutation3(p);
  const double *Vr = e.Vabij.get(), *Vi = e.Vabij.get() + n2;
  const double *T1r = T1, *T1i = T1 + v * o;
  for (size_t bi = 0; bi < blocks.size(); ++bi) {
    const auto &ib = blocks[bi];
    if (ib[0] == ib[2]) continue;  // i0 = i1 = i2: sf' = 0
    std::array<int, 3> tup[6];
    bool distinct[6];
    block_perms(ib, perms, tup, distinct);
    // Re W = Re T Re V - Im T Im V, Im W = Re T Im V + Im T Re V (both terms)
    build_w(e, prec, re, re, ib, perms, distinct, W.get(), 1.0, false);
    build_w(e, prec, im, im, ib, perms, distinct, W.get(), -1.0, true);
    build_w(e, prec, re, im, ib, perms, distinct, W.get() + v3, 1.0, false);
    build_w(e, prec, im, re, ib, perms, distinct, W.get() + v3, 1.0, true);
    for (int plane = 0; plane < 2; ++plane) {
      EnergyArgs g{};
      g.W = W.get() + plane * v3;
      // singles: Re(T1 V) = T1r Vr - T1i Vi, Im(T1 V) = T1r Vi + T1i Vr
      g.T1 = T1r;
      g.V = plane ? Vi : Vr;
      g.T1b = T1i;
      g.Vb = plane ? Vr : Vi;
      g.sb = plane ? 1.0 : -1.0;
      g.eps_a = e.eps_a.get();
      g.orbits = reinterpret_cast<const int *>(orbits.get());
      g.part = part.get();
      g.Nv = int(v);
      g.No = int(o);
      block_energy_args(g, ib, perms, tup, distinct, eps_i);
      g.norbits = norbits;
      const int grid = energy_grid(e, norbits, true);
      Timed tm(*e.timers, Kind::Bandwidth, 8.0 * double(v * v * v));
      hipLaunchKernelGGL(triples_energy_kernel<true>, dim3(grid), dim3(256), 0, e.stream, g);
      hipLaunchKernelGGL(sum_into_kernel, dim3(1), dim3(256), 0, e.stream, part.get(), grid,
                         Eb.get() + 2 * bi + plane);
      SISI_HIP(hipGetLastError());
    }
  }
  std::vector<double> eb(2 * blocks.size());
  e.download(eb.data(), Eb, eb.size());
  double E = 0.0;
  for (double x : eb) E += x;
  return E;
}

}  // namespace

double triples_energy(Engine &e, const double *T1, const double *T2, int prec,
                      const std::vector<std::array<int, 3>> &blocks_in) {
  SISI_REQUIRE(e.have_integrals, SISI_ERR_STATE, "triples: integrals not built");
  SISI_REQUIRE(prec == SISI_FP64 || prec == SISI_FP32 || prec == SISI_BF16X3, SISI_ERR_INVALID,
               "triples: unknown precision");
  e.thaw();
  SISI_REQUIRE(e.Nv <= 8 * 1023, SISI_ERR_INVALID, "triples: Nv above 8184");
  SISI_REQUIRE(e.Gab.get() && e.Gai.get() && e.Vabij.get() && (e.cplx ? e.Vaijk.get() : e.Vijka.get()),
               SISI_ERR_STATE, "triples: needs the vertex and its integrals");
  if (e.cplx) return ztriples_energy(e, T1, T2, prec, blocks_in);
  const int64_t o = e.No, v = e.Nv, G = e.NG;
  std::vector<std::array<int, 3>> blocks = checked_blocks(e, blocks_in);
  if (!e.Vt.get()) {
    // V_bdck = sum_F Gamma_Fbd Gamma_Fck stored as Vt[d,b,c,k] (cxx:88-91), FP64
    e.Vt = DBuf(&e.pool, v * v * v * o);
    ctr(e, 1.0, "Fbd,Fck->dbck", dense(e.Gab.get(), {G, v, v}), dense(e.Gai.get(), {G, v, o}),
        0.0, dense(e.Vt.get(), {v, v, v, o}), SISI_FP64);
    e.Vt2 = DBuf(&e.pool, v * v * v * o);
    sum(e, 1.0, dense(e.Vt.get(), {v, v, v, o}), "dbck", 0.0, dense(e.Vt2.get(), {v, v, v, o}),
        "dcbk");
    // V_jklc (HHHP) as [l, c, j, k] so each (j,k) slice is a [l, c] matrix
    e.VP = DBuf(&e.pool, o * v * o * o);
    sum(e, 1.0, dense(e.Vijka.get(), {o, o, o, v}), "jklc", 0.0, dense(e.VP.get(), {o, v, o, o}),
        "lcjk");
  }
  // T2 with a and b exchanged, for the permutations that put b first
  DBuf T2t(&e.pool, v * v * o * o);
  sum(e, 1.0, dense(const_cast<double *>(T2), {v, v, o, o}), "abij", 0.0,
      dense(T2t.get(), {v, v, o, o}), "baij");
  std::vector<double> eps_i(o);
  e.download(eps_i.data(), e.eps_i, size_t(o));
  const TriplesOps op{T2, T2t.get(), e.Vt.get(), e.Vt2.get(), e.VP.get()};
  TriplesPlanes pl{};
  if (prec == SISI_BF16X3) {
    auto pad = [](int64_t r) { return int(((r + kB3SegTile - 1) / kB3SegTile) * kB3SegTile); };
    if (!e.Vt_b3.hi) {
      split_planes(e, e.Vt.get(), v, 1, int(v * v), int(v), pad(v * v), int(o), v * v * v, 1.0,
                   e.Vt_b3);
      split_planes(e, e.Vt2.get(), v, 1, int(v * v), int(v), pad(v * v), int(o), v * v * v, 1.0,
                   e.Vt2_b3);
      split_planes(e, e.VP.get(), o, 1, int(v), int(o), pad(v), int(o * o), o * v, -1.0, e.VP_b3);
    }
    pl.V = &e.Vt_b3;
    pl.V2 = &e.Vt2_b3;
    pl.VP = &e.VP_b3;
    split_planes(e, T2, 1, v, int(v), int(v), pad(v), int(o * o), v * v, 1.0, pl.T2p);
    split_planes(e, T2, 1, v * v * o, int(v * v), int(o), pad(v * v), int(o), v * v, 1.0, pl.T2h);
    split_planes(e, T2t.get(), 1, v * v * o, int(v * v), int(o), pad(v * v), int(o), v * v, 1.0,
                 pl.T2th);
  }

  int norbits = 0;
  DBuf orbits_buf = orbit_codes(e, norbits);
  DBuf W(&e.pool, w_cube_size(v));
  DBuf part(&e.pool, norbits), Eb(&e.pool, blocks.size());
  zero(e, Eb.get(), blocks.size());
  std::array<std::array<int, 3>, 6> perms;
  for (int p = 0; p < 6; ++p) perms[p] = permutation3(p);

  for (size_t bi = 0; bi < blocks.size(); ++bi) {
    const auto &ib = blocks[bi];
    if (ib[0] == ib[2]) continue;  // i0 = i1 = i2: sf' = 0
    bool distinct[6];
    std::array<int, 3> tup[6];
    block_perms(ib, perms, tup, distinct);
    if (prec == SISI_BF16X3) build_w_b3(e, pl, ib, perms, distinct, W.get());
    else build_w(e, prec, op, op, ib, perms, distinct, W.get());

    EnergyArgs g{};
    g.W = W.get();
    g.T1 = T1;
    g.V = e.Vabij.get();
    g.eps_a = e.eps_a.get();
    g.orbits = reinterpret_cast<const int *>(orbits_buf.get());
    g.part = part.get();
    g.Nv = int(v);
    g.No = int(o);
    block_energy_args(g, ib, perms, tup, distinct, eps_i);
    g.norbits = norbits;
    {
      const int grid = energy_grid(e, norbits, false);
      Timed tm(*e.timers, Kind::Bandwidth, 8.0 * double(v * v * v));
      hipLaunchKernelGGL(triples_energy_kernel<false>, dim3(grid), dim3(256), 0, e.stream, g);
      hipLaunchKernelGGL(sum_into_kernel, dim3(1), dim3(256), 0, e.stream, part.get(), grid,
                         Eb.get() + bi);
      SISI_HIP(hipGetLastError());
    }
  }
  std::vector<double> eb(blocks.size());
  e.download(eb.data(), Eb, blocks.size());
  double E = 0.0;
  for (double x : eb) E += x;
  return E;
}

}  // namespace sisi
