// (T) triples (CcsdPerturbativeTriples.cxx:119-248) on the device.
//
// For each sorted block i = (i0 <= i1 <= i2) the reference forms, for the six
// permutations pi of (i0,i1,i2), W(i.pi)_abc = sum_d T_{ad,x y} V_{bdcz}
// - sum_l T_{ab,x l} V_{yzlc} and reads them back with "abc"*pi (cxx:159-181).
// Here the GEMM epilogue writes W(i.pi) straight into the permuted position,
// so the six GEMM pairs accumulate
//     Wsym[x] = sum_pi W(i.pi)[x_pi(0), x_pi(1), x_pi(2)].
// The reference energy (cxx:183-214),
//     E = sum_{distinct pi} sum_x DV[x] sum_sigma sf(sigma) Z_pi[("abc" sigma pi)(x)],
// DV = Wsym / D, Z_pi = W(i.pi) + S(i.pi), is rewritten by relabelling the
// summation index as
//     E = sum_x F[x] Zsym[x],
//     F[x] = (1/D[x]) sum_sigma sf(sigma) Wsym[x_sigma(0), x_sigma(1), x_sigma(2)],
//     Zsym[x] = sum_{distinct pi} Z_pi[x_pi(0), x_pi(1), x_pi(2)],
// with sf = +8 (identity), -4 (transpositions), +2 (3-cycles).  For blocks
// with three different indices every pi is distinct and Zsym = Wsym + Ssym;
// degenerate blocks accumulate the distinct W(i.pi) separately.  The energy
// kernel evaluates F and Zsym on 8x8x8 cubes whose six permuted partner cubes
// are staged in LDS, and reduces in FP64.
#include "triples.hpp"

#include "contract.hpp"
#include "tensor_ops.hpp"

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
  const double *W;   // Wsym
  const double *Z;   // W part of Zsym
  const double *T1;  // [Nv, No]
  const double *V;   // Vabij [Nv,Nv,No,No]
  const double *eps_a;
  double *part;
  double eps_ijk;
  int Nv, No, ncube, nterms;
  PermTerm terms[6];
};

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

__global__ __launch_bounds__(256) void triples_energy_kernel(EnergyArgs g) {
  __shared__ double tile[6][512];
  const int nc = g.ncube;
  const int cb = blockIdx.x;
  const int X[3] = {cb % nc, (cb / nc) % nc, cb / (nc * nc)};
  const int Nv = g.Nv;
  const int64_t s1 = Nv, s2 = int64_t(Nv) * Nv;
  // sigma as reads (x_s0, x_s1, x_s2); sf by fixed points
  const int sig[6][3] = {{0, 1, 2}, {1, 0, 2}, {2, 1, 0}, {0, 2, 1}, {1, 2, 0}, {2, 0, 1}};
  const double sf[6] = {8.0, -4.0, -4.0, -4.0, 2.0, 2.0};
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int b0 = X[sig[s][0]] * 8, b1 = X[sig[s][1]] * 8, b2 = X[sig[s][2]] * 8;
    for (int e = threadIdx.x; e < 512; e += 256) {
      const int u0 = b0 + (e & 7), u1 = b1 + ((e >> 3) & 7), u2 = b2 + (e >> 6);
      tile[s][e] = (u0 < Nv && u1 < Nv && u2 < Nv) ? g.W[u0 + u1 * s1 + u2 * s2] : 0.0;
    }
  }
  __syncthreads();
  double acc = 0.0;
  for (int e = threadIdx.x; e < 512; e += 256) {
    const int l[3] = {e & 7, (e >> 3) & 7, e >> 6};
    const int x[3] = {X[0] * 8 + l[0], X[1] * 8 + l[1], X[2] * 8 + l[2]};
    if (x[0] >= Nv || x[1] >= Nv || x[2] >= Nv) continue;
    double f = 0.0;
#pragma unroll
    for (int s = 0; s < 6; ++s)
      f += sf[s] * tile[s][l[sig[s][0]] + 8 * l[sig[s][1]] + 64 * l[sig[s][2]]];
    const double D = g.eps_ijk - g.eps_a[x[0]] - g.eps_a[x[1]] - g.eps_a[x[2]];
    f /= D;
    double z = g.Z[x[0] + x[1] * s1 + x[2] * s2];
    for (int t = 0; t < g.nterms; ++t) {
      const PermTerm &p = g.terms[t];
      // S(i.pi)[u] = 0.5 T1[u0, i_pi0] V[u1, u2, i_pi1, i_pi2] read at u_k = x_pi(k)
      const int u0 = x[p.pi[0]], u1 = x[p.pi[1]], u2 = x[p.pi[2]];
      z += 0.5 * g.T1[u0 + int64_t(Nv) * p.idx[0]] *
           g.V[u1 + s1 * u2 + s2 * (p.idx[1] + int64_t(g.No) * p.idx[2])];
    }
    acc += f * z;
  }
  acc = block_sum_t(acc);
  if (threadIdx.x == 0) g.part[blockIdx.x] = acc;
}

__global__ void sum_into_kernel(const double *part, int n, double *out) {
  double s = 0.0;
  for (int b = threadIdx.x; b < n; b += blockDim.x) s += part[b];
  s = block_sum_t(s);
  if (threadIdx.x == 0) *out = s;
}

// W(x,y,z)[u0,u1,u2] accumulated (beta = 1) at Wp with output strides st.
// The epilogue's lanes walk the output's unit-stride index when it is the
// leading index of the GEMM's row or column group, so the operand orders are
// chosen per permutation: V_bdck as [d,c,b] when c lands on the unit stride,
// T2 as [b,a,x,l] when b does (two copies of the same tensors, exact).
void add_w(Engine &e, int prec, const double *T2, const double *T2t, const double *Vt,
           const double *Vt2, const double *VijkaP, int x, int y, int z, double *Wp,
           const int64_t st[3]) {
  const int64_t o = e.No, v = e.Nv;
  View C;
  C.p = Wp;
  C.order = 3;
  for (int k = 0; k < 3; ++k) {
    C.len[k] = v;
    C.stride[k] = st[k];
  }
  View Txy = dense(const_cast<double *>(T2) + (x + o * y) * v * v, {v, v});
  if (st[2] == 1) {
    View Vz = dense(const_cast<double *>(Vt2) + z * v * v * v, {v, v, v});
    ctr(e, 1.0, "ad,dcb->abc", Txy, Vz, 1.0, C, prec);  // sum_d T_ad,xy V_bdcz
  } else {
    View Vz = dense(const_cast<double *>(Vt) + z * v * v * v, {v, v, v});
    ctr(e, 1.0, "ad,dbc->abc", Txy, Vz, 1.0, C, prec);
  }
  View Tx;
  Tx.p = const_cast<double *>(st[1] == 1 ? T2t : T2) + x * v * v;
  Tx.order = 3;
  Tx.len[0] = v;
  Tx.len[1] = v;
  Tx.len[2] = o;
  Tx.stride[0] = 1;
  Tx.stride[1] = v;
  Tx.stride[2] = v * v * o;
  View Vyz;
  Vyz.p = const_cast<double *>(VijkaP) + (y + o * z) * o * v;
  Vyz.order = 2;
  Vyz.len[0] = o;
  Vyz.len[1] = v;
  Vyz.stride[0] = 1;
  Vyz.stride[1] = o;
  // - sum_l T_ab,xl V_yzlc (T2t[b,a,x,l] = T2[a,b,x,l])
  ctr(e, -1.0, st[1] == 1 ? "bal,lc->abc" : "abl,lc->abc", Tx, Vyz, 1.0, C, prec);
}

}  // namespace

double triples_energy(Engine &e, const double *T1, const double *T2, int prec,
                      const std::vector<std::array<int, 3>> &blocks_in) {
  SISI_REQUIRE(e.have_integrals, SISI_ERR_STATE, "triples: integrals not built");
  e.thaw();
  const int64_t o = e.No, v = e.Nv, G = e.NG;
  std::vector<std::array<int, 3>> blocks = blocks_in;
  if (blocks.empty())
    for (int a = 0; a < o; ++a)
      for (int b = a; b < o; ++b)
        for (int c = b; c < o; ++c) blocks.push_back({a, b, c});
  for (auto &b : blocks)
    SISI_REQUIRE(b[0] >= 0 && b[0] <= b[1] && b[1] <= b[2] && b[2] < o, SISI_ERR_INVALID,
                 "triples: blocks must be sorted i0 <= i1 <= i2 < No");
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
  DBuf &Vt = e.Vt, &Vt2 = e.Vt2, &VP = e.VP;
  // T2 with a and b exchanged, for the permutations that put b on the unit stride
  DBuf T2t(&e.pool, v * v * o * o);
  sum(e, 1.0, dense(const_cast<double *>(T2), {v, v, o, o}), "abij", 0.0,
      dense(T2t.get(), {v, v, o, o}), "baij");
  std::vector<double> eps_i(o);
  e.download(eps_i.data(), e.eps_i, size_t(o));

  DBuf Wsym(&e.pool, v * v * v), Wdist(&e.pool, v * v * v), Wtmp(&e.pool, 6 * v * v * v);
  const int ncube = int((v + 7) / 8);
  const int ncubes = ncube * ncube * ncube;
  DBuf part(&e.pool, ncubes), Eb(&e.pool, blocks.size());
  const int64_t s[3] = {1, v, v * v};
  std::array<std::array<int, 3>, 6> perms;
  for (int p = 0; p < 6; ++p) perms[p] = permutation3(p);

  for (size_t bi = 0; bi < blocks.size(); ++bi) {
    const auto &ib = blocks[bi];
    bool distinct[6];
    int equiv[6];
    std::array<int, 3> tup[6];
    for (int p = 0; p < 6; ++p) {
      for (int k = 0; k < 3; ++k) tup[p][k] = ib[perms[p][k]];  // i * pi (Permutation.hpp:93-98)
      int q = 0;
      for (; q < p; ++q)
        if (tup[q] == tup[p]) break;
      distinct[p] = (q == p);
      equiv[p] = q;
    }
    const bool degenerate = (ib[0] == ib[1]) || (ib[1] == ib[2]);
    zero(e, Wsym.get(), v * v * v);
    if (!degenerate) {
      for (int p = 0; p < 6; ++p) {
        const int64_t st[3] = {s[perms[p][0]], s[perms[p][1]], s[perms[p][2]]};
        add_w(e, prec, T2, T2t.get(), Vt.get(), Vt2.get(), VP.get(), tup[p][0], tup[p][1], tup[p][2],
              Wsym.get(), st);
      }
    } else {
      zero(e, Wdist.get(), v * v * v);
      for (int p = 0; p < 6; ++p) {
        const int src = distinct[p] ? p : equiv[p];
        double *Wp = Wtmp.get() + src * v * v * v;
        if (distinct[p]) {
          zero(e, Wp, v * v * v);
          const int64_t st[3] = {1, v, v * v};
          add_w(e, prec, T2, T2t.get(), Vt.get(), Vt2.get(), VP.get(), tup[p][0], tup[p][1],
                tup[p][2], Wp, st);
        }
        View Wv = dense(Wp, {v, v, v});
        View Ws, Wd;
        Ws.p = Wsym.get();
        Wd.p = Wdist.get();
        Ws.order = Wd.order = 3;
        for (int k = 0; k < 3; ++k) {
          Ws.len[k] = Wd.len[k] = v;
          Ws.stride[k] = Wd.stride[k] = s[perms[p][k]];
        }
        copy(e, Ws, Wv, 1.0, 1.0);
        if (distinct[p]) copy(e, Wd, Wv, 1.0, 1.0);
      }
    }
    EnergyArgs g{};
    g.W = Wsym.get();
    g.Z = degenerate ? Wdist.get() : Wsym.get();
    g.T1 = T1;
    g.V = e.Vabij.get();
    g.eps_a = e.eps_a.get();
    g.part = part.get();
    g.eps_ijk = eps_i[ib[0]] + eps_i[ib[1]] + eps_i[ib[2]];
    g.Nv = int(v);
    g.No = int(o);
    g.ncube = ncube;
    g.nterms = 0;
    for (int p = 0; p < 6; ++p) {
      if (!distinct[p]) continue;
      PermTerm &t = g.terms[g.nterms++];
      for (int k = 0; k < 3; ++k) {
        t.pi[k] = perms[p][k];
        t.idx[k] = tup[p][k];
      }
    }
    {
      Timed tm(*e.timers, Kind::Bandwidth, 8.0 * double(v * v * v) * (degenerate ? 8.0 : 7.0));
      hipLaunchKernelGGL(triples_energy_kernel, dim3(ncubes), dim3(256), 0, e.stream, g);
      hipLaunchKernelGGL(sum_into_kernel, dim3(1), dim3(256), 0, e.stream, part.get(), ncubes,
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
