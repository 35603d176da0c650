// CCSD hot path on the device.  Every contraction below names the reference
// line it evaluates; operands are FP64 in HBM and the contraction precision
// is the engine's (FP64, FP32, or BF16x3 for the ladder).
#include "ccsd.hpp"

#include <cmath>

#include "contract.hpp"
#include "gemm.hpp"
#include "tensor_ops.hpp"

namespace sisi {

namespace {
inline View V2(double *p, int64_t a, int64_t b) { return dense(p, {a, b}); }
inline View V3(double *p, int64_t a, int64_t b, int64_t c) { return dense(p, {a, b, c}); }
inline View V4(double *p, int64_t a, int64_t b, int64_t c, int64_t d) {
  return dense(p, {a, b, c, d});
}
inline View V1(double *p, int64_t a) { return dense(p, {a}); }
}  // namespace

void set_precision(Engine &e, int precision) {
  SISI_REQUIRE(is_precision(precision), SISI_ERR_INVALID, "unknown precision");
  // BF16X3: every contraction in split-BF16x3 (the GEMM family itself keeps
  // FP32 for products too small to pay for the operand split)
  e.precision = precision;
  e.prec_ladder = precision;
}

void integrals_from_vertex(Engine &e) {
  if (e.cplx) return zintegrals_from_vertex(e);
  e.require_inputs();
  const int64_t o = e.No, v = e.Nv, G = e.NG, P = e.Np;
  // slices (CoulombIntegralsFromVertex.cxx:121-136; holes are p < No)
  View gam = V3(e.gamma.get(), G, P, P);
  e.Gij = DBuf(&e.pool, G * o * o);
  e.Gai = DBuf(&e.pool, G * v * o);
  e.Gab = DBuf(&e.pool, G * v * v);
  copy(e, V3(e.Gij.get(), G, o, o), subview(gam, {0, 0, 0}, {G, o, o}));
  copy(e, V3(e.Gai.get(), G, v, o), subview(gam, {0, o, 0}, {G, P, o}));
  copy(e, V3(e.Gab.get(), G, v, v), subview(gam, {0, o, o}, {G, P, P}));
  View Gij = V3(e.Gij.get(), G, o, o), Gai = V3(e.Gai.get(), G, v, o),
       Gab = V3(e.Gab.get(), G, v, v);
  e.Vaibj = DBuf(&e.pool, v * o * v * o);
  e.Vabij = DBuf(&e.pool, v * v * o * o);
  e.Vijkl = DBuf(&e.pool, o * o * o * o);
  e.Vijka = DBuf(&e.pool, o * o * o * v);
  ctr(e, 1.0, "Gab,Gij->aibj", Gab, Gij, 0.0, V4(e.Vaibj.get(), v, o, v, o), SISI_FP64);  // :395
  ctr(e, 1.0, "Gai,Gbj->abij", Gai, Gai, 0.0, V4(e.Vabij.get(), v, v, o, o), SISI_FP64);  // :402
  ctr(e, 1.0, "Gik,Gjl->ijkl", Gij, Gij, 0.0, V4(e.Vijkl.get(), o, o, o, o), SISI_FP64);  // :409
  ctr(e, 1.0, "Gik,Gaj->ijka", Gij, Gai, 0.0, V4(e.Vijka.get(), o, o, o, v), SISI_FP64);  // :416
  // constants of the residual and the energy
  View V = V4(e.Vabij.get(), v, v, o, o);
  e.Vbaij = DBuf(&e.pool, v * v * o * o);
  e.Uabij = DBuf(&e.pool, v * v * o * o);
  sum(e, 1.0, V, "baij", 0.0, V4(e.Vbaij.get(), v, v, o, o), "abij");
  // U = 2 V_abij - V_baij: the combination (2 V[cdkl] - V[dckl]) of
  // CcsdEnergyFromCoulombIntegrals.cxx:146-147, 172-173, 462-463
  axpby(e, v * v * o * o, 2.0, e.Vabij.get(), 0.0, e.Uabij.get());
  axpby(e, v * v * o * o, -1.0, e.Vbaij.get(), 1.0, e.Uabij.get());
  e.drop_triples_cache();
  e.drop_residual_graph();
  e.have_integrals = true;
  e.sync();
}

void ccsd_residual(Engine &e, int it, bool initial_doubles, int slice_size, const double *T1c,
                   const double *T2c, double *R1p, double *R2p) {
  if (e.formulation == SISI_FORMULATION_DRCCD)
    return drccd_residual(e, it, initial_doubles, T2c, R1p, R2p);
  if (e.formulation == SISI_FORMULATION_UCCSD)
    return uccsd_residual(e, it, initial_doubles, T1c, T2c, R1p, R2p);
  if (e.cplx) return zccsd_residual(e, it, initial_doubles, slice_size, T1c, T2c, R1p, R2p);
  if (e.formulation != SISI_FORMULATION_VERTEX)
    return ccsd_residual_full(e, it, initial_doubles, T1c, T2c, R1p, R2p);
  SISI_REQUIRE(e.have_integrals, SISI_ERR_STATE, "residual: integrals not built");
  const int64_t o = e.No, v = e.Nv, G = e.NG;
  const int64_t n2 = v * v * o * o;
  double *T1p = const_cast<double *>(T1c);
  double *T2p = const_cast<double *>(T2c);
  View T1 = V2(T1p, v, o), T2 = V4(T2p, v, v, o, o);
  View R1 = V2(R1p, v, o), R2 = V4(R2p, v, v, o, o);
  View V = V4(e.Vabij.get(), v, v, o, o), U = V4(e.Uabij.get(), v, v, o, o);
  View Vaibj = V4(e.Vaibj.get(), v, o, v, o), Vijkl = V4(e.Vijkl.get(), o, o, o, o),
       Vijka = V4(e.Vijka.get(), o, o, o, v);
  View Gij = V3(e.Gij.get(), G, o, o), Gai = V3(e.Gai.get(), G, v, o),
       Gab = V3(e.Gab.get(), G, v, v);
  zero(e, R1p, v * o);
  if (it == 0 && !initial_doubles) {  // MP2 start (cxx:47-51)
    copy(e, R2, V);
    return;
  }
  zero(e, R2p, n2);
  SISI_REQUIRE(e.Gij.get() && e.Gai.get() && e.Gab.get() && e.Vaibj.get() && e.Vijkl.get() &&
                   e.Vijka.get(),
               SISI_ERR_STATE, "the vertex residual needs the vertex and its integrals");
  const int prec = e.precision;
  e.thaw();
  e.frozen = {T1p, T2p, e.Vabij.get(), e.Uabij.get(), e.Vaibj.get(), e.Vijka.get(),
              e.Gij.get(), e.Gai.get(), e.Gab.get()};

  DBuf Xb(&e.pool, n2);
  View X = V4(Xb.get(), v, v, o, o);
  add_t1t1(e, Xb.get(), T2p, T1p, 1.0, int(v), int(o));  // cxx:136-138
  e.frozen.push_back(Xb.get());

  DBuf Kacb(&e.pool, v * v), Lacb(&e.pool, v * v), Kkib(&e.pool, o * o), Lkib(&e.pool, o * o);
  View Kac = V2(Kacb.get(), v, v), Lac = V2(Lacb.get(), v, v), Kki = V2(Kkib.get(), o, o),
       Lki = V2(Lkib.get(), o, o);
  {
    // Kac = -2 V[cdkl] X[adkl] + V[dckl] X[adkl] = -U[cdkl] X[adkl] (cxx:146-147)
    ctr(e, -1.0, "cdkl,adkl->ac", U, X, 0.0, Kac, prec);
    // distinguishable: Lac = -V[cdkl] Y[adkl] + V[dckl] Y[adkl] / 2
    // = -U[cdkl] Y[adkl] / 2 with Y = T2 + 2 T1 T1 (cxx:150-160)
    DBuf Ydb;
    if (e.distinguishable) {
      Ydb = DBuf(&e.pool, n2);
      add_t1t1(e, Ydb.get(), T2p, T1p, 2.0, int(v), int(o));
      ctr(e, -0.5, "cdkl,adkl->ac", U, V4(Ydb.get(), v, v, o, o), 0.0, Lac, prec);
    } else {
      copy(e, Lac, Kac);  // cxx:151
    }
    // + 2 Gamma_Gca Gamma_Gdk T_dk (cxx:162-165)
    DBuf Zb(&e.pool, G);
    ctr(e, 1.0, "Gdk,dk->G", Gai, T1, 0.0, V1(Zb.get(), G), prec);
    ctr(e, 2.0, "Gca,G->ac", Gab, V1(Zb.get(), G), 1.0, Lac, prec);
    // - Gamma_Gck Gamma_Gda T_dk (cxx:166-169)
    DBuf Yb(&e.pool, G * v * o);
    ctr(e, 1.0, "Gda,dk->Gak", Gab, T1, 0.0, V3(Yb.get(), G, v, o), prec);
    ctr(e, -1.0, "Gck,Gak->ac", Gai, V3(Yb.get(), G, v, o), 1.0, Lac, prec);
    // Kki = 2 V[cdkl] X[cdil] - V[dckl] X[cdil] = U[cdkl] X[cdil] (cxx:172-173)
    ctr(e, 1.0, "cdkl,cdil->ki", U, X, 0.0, Kki, prec);
    if (e.distinguishable)  // Lki = U[cdkl] Y[cdil] / 2 (cxx:176-186)
      ctr(e, 0.5, "cdkl,cdil->ki", U, V4(Ydb.get(), v, v, o, o), 0.0, Lki, prec);
    else
      copy(e, Lki, Kki);  // cxx:177
    ctr(e, 2.0, "klic,cl->ki", Vijka, T1, 1.0, Lki, prec);   // cxx:188
    ctr(e, -1.0, "lkic,cl->ki", Vijka, T1, 1.0, Lki, prec);  // cxx:189
    e.record("Kac", Kacb.get(), v * v);
    e.record("Lac", Lacb.get(), v * v);
    e.record("Kki", Kkib.get(), o * o);
    e.record("Lki", Lkib.get(), o * o);
    ctr(e, 1.0, "ac,cbij->abij", Lac, T2, 1.0, R2, prec);    // cxx:192
    ctr(e, -1.0, "ki,abkj->abij", Lki, T2, 1.0, R2, prec);   // cxx:195
  }
  {
    // dressed Gamma_ai and its terms (cxx:198-216)
    DBuf dGb(&e.pool, G * v * o), Qb(&e.pool, G * v * o);
    View dGai = V3(dGb.get(), G, v, o), Q = V3(Qb.get(), G, v, o);
    copy(e, dGai, Gai);
    ctr(e, -1.0, "Gki,ak->Gai", Gij, T1, 1.0, dGai, prec);   // cxx:205
    ctr(e, 1.0, "Gbc,cj->Gbj", Gab, T1, 0.0, Q, prec);       // cxx:208 (Gamma_Gbc T_cj)
    e.record("dGai2b", dGb.get(), G * v * o);
    e.record("Q", Qb.get(), G * v * o);
    ctr(e, 1.0, "Gai,Gbj->abij", dGai, Q, 1.0, R2, prec);    // cxx:208-209
    ctr(e, -1.0, "jika,bk->abij", Vijka, T1, 1.0, R2, prec); // cxx:213
    DBuf Pb(&e.pool, v * o * o * o);
    View P = V4(Pb.get(), v, o, o, o);
    ctr(e, 1.0, "acik,cj->aikj", V, T1, 0.0, P, prec);       // cxx:214-215
    e.record("P", Pb.get(), v * o * o * o);
    ctr(e, -1.0, "bk,aikj->abij", T1, P, 1.0, R2, prec);
  }
  {
    // Xakic (cxx:218-252)
    DBuf dGb(&e.pool, G * v * o);
    View dGai = V3(dGb.get(), G, v, o);
    copy(e, dGai, Gai);
    ctr(e, -1.0, "Gil,al->Gai", Gij, T1, 1.0, dGai, prec);  // cxx:227
    ctr(e, 1.0, "Gad,di->Gai", Gab, T1, 1.0, dGai, prec);   // cxx:230
    DBuf Xb2(&e.pool, v * o * o * v);
    View Xakic = V4(Xb2.get(), v, o, o, v);
    ctr(e, 1.0, "Gai,Gck->akic", dGai, Gai, 0.0, Xakic, prec);  // cxx:234
    DBuf Yb(&e.pool, n2);
    View Y = V4(Yb.get(), v, v, o, o);
    add_t1t1(e, Yb.get(), T2p, T1p, 2.0, int(v), int(o));    // cxx:238-240
    ctr(e, -0.5, "dclk,dail->akic", V, Y, 1.0, Xakic, prec);   // cxx:242
    ctr(e, 1.0, "dclk,adil->akic", V, T2, 1.0, Xakic, prec);   // cxx:243
    if (!e.distinguishable)
      ctr(e, -0.5, "cdlk,adil->akic", V, T2, 1.0, Xakic, prec);  // cxx:244-246
    e.record("Xakic", Xb2.get(), v * o * o * v);
    // Y = 2 T[cbkj] - T[bckj] (cxx:248-249)
    axpby(e, n2, 2.0, T2p, 0.0, Yb.get());
    sum(e, -1.0, T2, "bckj", 1.0, Y, "cbkj");
    ctr(e, 1.0, "akic,cbkj->abij", Xakic, Y, 1.0, R2, prec);  // cxx:251
  }
  {
    // Xakci (cxx:254-290)
    DBuf dGabb(&e.pool, G * v * v), dGijb(&e.pool, G * o * o);
    View dGab = V3(dGabb.get(), G, v, v), dGij = V3(dGijb.get(), G, o, o);
    copy(e, dGab, Gab);
    copy(e, dGij, Gij);
    ctr(e, -1.0, "Gcl,al->Gac", Gai, T1, 1.0, dGab, prec);  // cxx:270
    ctr(e, 1.0, "Gdk,di->Gki", Gai, T1, 1.0, dGij, prec);   // cxx:273
    e.record("dGab4", dGabb.get(), G * v * v);
    e.record("dGij4", dGijb.get(), G * o * o);
    DBuf Xb3(&e.pool, v * o * v * o);
    View Xakci = V4(Xb3.get(), v, o, v, o);
    ctr(e, 1.0, "Gac,Gki->akci", dGab, dGij, 0.0, Xakci, prec);  // cxx:277-278
    if (!e.distinguishable)
      ctr(e, -0.5, "cdlk,dail->akci", V, T2, 1.0, Xakci, prec);  // cxx:283-285
    e.frozen.push_back(Xb3.get());
    e.record("Xakci", Xb3.get(), v * o * v * o);
    ctr(e, -1.0, "akci,cbkj->abij", Xakci, T2, 1.0, R2, prec);   // cxx:288
    ctr(e, -1.0, "bkci,ackj->abij", Xakci, T2, 1.0, R2, prec);   // cxx:289
    e.unfreeze(Xb3.get());  // Xb3 returns to the pool at the end of this block
  }
  {
    // R_abij += R_baji (cxx:292)
    DBuf Rt(&e.pool, n2);
    copy(e, V4(Rt.get(), v, v, o, o), R2);
    sum(e, 1.0, V4(Rt.get(), v, v, o, o), "baji", 1.0, R2, "abij");
  }
  axpby(e, n2, 1.0, e.Vabij.get(), 1.0, R2p);  // cxx:301
  {
    // Xklij (cxx:303-334)
    DBuf Xb4(&e.pool, o * o * o * o);
    View Xklij = V4(Xb4.get(), o, o, o, o);
    copy(e, Xklij, Vijkl);
    ctr(e, 1.0, "klic,cj->klij", Vijka, T1, 1.0, Xklij, prec);  // cxx:308
    ctr(e, 1.0, "lkjc,ci->klij", Vijka, T1, 1.0, Xklij, prec);  // cxx:309
    e.record("Xklij1", Xb4.get(), o * o * o * o);
    ctr(e, 1.0, "klij,abkl->abij", Xklij, X, 1.0, R2, prec);    // cxx:313
    if (e.distinguishable) {
      // Xklij = T[ci] V[cdkl] T[dj] (cxx:319): V contracted with X - T2
      DBuf Wb(&e.pool, n2);
      axpby(e, n2, 1.0, Xb.get(), 0.0, Wb.get());
      axpby(e, n2, -1.0, T2p, 1.0, Wb.get());
      ctr(e, 1.0, "cdkl,cdij->klij", V, V4(Wb.get(), v, v, o, o), 0.0, Xklij, prec);
    } else {
      ctr(e, 1.0, "cdkl,cdij->klij", V, X, 0.0, Xklij, prec);   // cxx:317
    }
    e.record("Xklij2", Xb4.get(), o * o * o * o);
    // ppl = 1: the T1 part of this term comes with the ladder (cxx:326-333)
    ctr(e, 1.0, "klij,abkl->abij", Xklij, e.ppl ? T2 : X, 1.0, R2, prec);
  }
  if (e.ppl && e.NR > 0) {
    // factor-based ladder (cxx:340-372)
    factor_ladder(e, T1p, Xb.get(), R2p, e.prec_ladder);
  } else if (e.ppl) {
    // particle-particle ladder (cxx:336-453)
    if (slice_size <= 0) slice_size = int(o);
    DBuf dGabb(&e.pool, G * v * v);
    View dGab = V3(dGabb.get(), G, v, v);
    copy(e, dGab, Gab);
    ctr(e, -1.0, "Gbk,ak->Gab", Gai, T1, 1.0, dGab, prec);  // cxx:397-400
    e.record("dGabP", dGabb.get(), G * v * v);
    const int nslices = int((v + slice_size - 1) / slice_size);
    std::vector<DBuf> sl(nslices);
    std::vector<int64_t> lo(nslices), hi(nslices);
    for (int s = 0; s < nslices; ++s) {  // cxx:402-416
      lo[s] = int64_t(s) * slice_size;
      hi[s] = std::min<int64_t>(lo[s] + slice_size, v);
      sl[s] = DBuf(&e.pool, G * (hi[s] - lo[s]) * v);
      copy(e, V3(sl[s].get(), G, hi[s] - lo[s], v), subview(dGab, {0, lo[s], 0}, {G, hi[s], v}));
    }
    const int64_t smax = std::min<int64_t>(slice_size, v);
    DBuf Vxy(&e.pool, smax * smax * v * v), Rxy(&e.pool, smax * smax * o * o);
    // an anchor keeps the bf16 planes of its V~ slices for the refinement
    // residuals when they fit in half the free device memory
    bool keep_planes = false;
    if (e.recording) {
      e.anchor.ladder.clear();
      e.anchor.ladder_stack = DBuf();
      e.anchor.ladder_rows = 0;
      e.anchor.ladder_off.clear();
      e.anchor.ladder_slice = slice_size;
      const int64_t Kp = ((v * v + 31) / 32) * 32;
      double bytes = 0.0;
      for (int m = 0; m < nslices; ++m)
        for (int n = m; n < nslices; ++n)
          bytes += 4.0 * double(((hi[n] - lo[n]) * (hi[m] - lo[m]) + 255) / 256 * 256) * double(Kp);
      const int64_t Mps = (o * v + 255) / 256 * 256;
      const double s_bytes = 4.0 * double(Mps) * double(Kp);
      // device memory free or cached in the pool's free lists (a run's
      // earlier phases leave buffers there; an allocation that does not fit
      // gives them back first)
      size_t free_b = 0, total_b = 0;
      const bool info = hipMemGetInfo(&free_b, &total_b) == hipSuccess;
      const size_t cached = e.pool.cached_bytes();
      if (getenv("SISI_VERBOSE"))
        fprintf(stderr, "[sisi] anchor ladder planes: %.1f GB (+ %.1f GB S) against half of "
                "%.1f GB free + %.1f GB cached\n", bytes / 1e9, s_bytes / 1e9,
                double(free_b) / 1e9, double(cached) / 1e9);
      free_b += cached;
      keep_planes = info && bytes <= 0.5 * double(free_b);
      e.anchor.ladder_s = B3FrozenA();
      if (keep_planes && info && bytes + s_bytes <= 0.5 * double(free_b) &&
          get_b3_backend() == 0 && 2.0 * double(o * v) * double(v * v) * double(G) >= kB3MinFlops) {
        // S[c,k,b,d] = sum_G Gai[G,c,k] dGab[G,b,d]: product rows m = c + v k,
        // columns n = b + v d, written as split-BF16x3 planes at row k + o b,
        // column c + v d
        B3FrozenA &S = e.anchor.ladder_s;
        S.store = DBuf(&e.pool, size_t((2 * Mps * Kp * 2 + 7) / 8));
        if (Kp != v * v) zero(e, S.store.get(), int64_t(S.store.size()));
        GemmArgs gs;
        gs.A = e.Gai.get();
        gs.sam = G;
        gs.sak = 1;
        gs.B = dGabb.get();
        gs.sbk = 1;
        gs.sbn = G;
        gs.M = int(v * o);
        gs.N = int(v * v);
        gs.K = int(G);
        gs.dm0 = int(v);
        gs.dm1 = int(o);
        gs.cm0 = 1;
        gs.cm1 = Kp;
        gs.dn0 = int(v);
        gs.dn1 = int(v);
        gs.cn0 = o * Kp;
        gs.cn1 = v;
        gs.C = S.store.get();
        gs.out_planes = 1;
        gs.plane_lo = Mps * Kp;
        gemm_b3(e, gs, 3);
        S.M = int(o * v);
        S.K = int(v * v);
        S.Mp = int(Mps);
        S.Kp = int(Kp);
      }
    }
    // reduced-precision ladder with plane output: the slices are re-laid as
    // [G, c, x] so that the V~ product's rows m = c + v x run along c, the
    // unit-stride index of the planes' rows (coalesced bf16 stores)
    // (only when every pair's two products are large enough for the bf16
    // kernels, so the precision of each product is the one the stored-pair
    // ladder would use)
    const int pl = e.prec_ladder;
    const int64_t smin = hi[nslices - 1] - lo[nslices - 1];
    const bool planes = e.ladder_planes && !e.tape && !keep_planes &&
                        (pl == SISI_BF16X3 || pl == SISI_BF16) && get_b3_backend() == 0 &&
                        2.0 * double(smin * smin) * double(v * v) * double(std::min(G, o * o)) >=
                            kB3MinFlops;
    if (planes) {
      const int64_t Kp = (v * v + 31) / 32 * 32;
      const int64_t Mmax = (smax * smax + 255) / 256 * 256;
      std::vector<DBuf> slt(nslices);
      for (int s2 = 0; s2 < nslices; ++s2) {
        const int64_t Ns = hi[s2] - lo[s2];
        slt[s2] = DBuf(&e.pool, G * Ns * v);
        // element (G, x, c) at G + NG c + NG v x
        copy(e, dense_strided(slt[s2].get(), {G, Ns, v}, {1, G * v, G}),
             subview(dGab, {0, lo[s2], 0}, {G, hi[s2], v}));
        e.frozen.push_back(slt[s2].get());  // split into planes once, used by every pair
      }
      B3FrozenA fa;
      fa.store = DBuf(&e.pool, size_t((2 * Mmax * Kp * 2 + 7) / 8));
      // the plane columns past v^2 (k padding) are never written: zero once
      if (Kp != v * v) zero(e, fa.store.get(), int64_t(fa.store.size()));
      for (int m = 0; m < nslices; ++m)
        for (int n = m; n < nslices; ++n) {  // cxx:418-451
          const int64_t Nx = hi[n] - lo[n], Ny = hi[m] - lo[m];
          const int64_t Mp = (Nx * Ny + 255) / 256 * 256;
          View Rxyij = V4(Rxy.get(), Nx, Ny, o, o);
          // V~[x,y,c,d] = sum_G slice_n[G,x,c] slice_m[G,y,d] (cxx:431-434) as
          // plane row r = x + Nx y, column c + v d
          GemmArgs g1;
          g1.A = slt[n].get();
          g1.sam = G;
          g1.sak = 1;
          g1.B = slt[m].get();
          g1.sbk = 1;
          g1.sbn = G;
          g1.M = int(v * Nx);
          g1.N = int(v * Ny);
          g1.K = int(G);
          g1.dm0 = int(v);
          g1.dm1 = int(Nx);
          g1.cm0 = 1;
          g1.cm1 = Kp;
          g1.dn0 = int(v);
          g1.dn1 = int(Ny);
          g1.cn0 = v;
          g1.cn1 = Nx * Kp;
          g1.C = fa.store.get();
          g1.out_planes = pl == SISI_BF16 ? 2 : 1;
          g1.plane_lo = Mp * Kp;
          gemm_b3(e, g1, pl == SISI_BF16 ? 1 : 3);
          fa.M = int(Nx * Ny);
          fa.K = int(v * v);
          fa.Mp = int(Mp);
          fa.Kp = int(Kp);
          // R~[x,y,i,j] = sum_cd V~[x,y,c,d] X[c,d,i,j] (cxx:443)
          GemmArgs g2;
          g2.B = Xb.get();
          g2.sbk = 1;
          g2.sbn = v * v;
          g2.M = int(Nx * Ny);
          g2.N = int(o * o);
          g2.K = int(v * v);
          g2.C = Rxy.get();
          g2.dm0 = g2.M;
          g2.dn0 = g2.N;
          g2.cm0 = 1;
          g2.cn0 = Nx * Ny;
          gemm_b3_frozen_a(e, g2, fa, pl == SISI_BF16 ? 1 : 3);
          const int64_t a = lo[n], b = lo[m];
          copy(e, subview(R2, {a, b, 0, 0}, {a + Nx, b + Ny, o, o}), Rxyij, 1.0, 1.0);
          if (a > b) sum(e, 1.0, Rxyij, "xyij", 1.0, subview(R2, {b, a, 0, 0}, {b + Ny, a + Nx, o, o}), "yxji");
        }
      for (auto &b : slt) e.unfreeze(b.get());
    }
    // kept planes: stacked in one store when every pair is large enough for
    // the bf16 kernels (rows of pair p from ladder_off[p], each pair padded
    // to 256 rows), else one store per pair
    const int64_t smin_k = hi[nslices - 1] - lo[nslices - 1];
    const int64_t Kp_k = (v * v + 31) / 32 * 32;
    uint16_t *stack_hi = nullptr;
    if (keep_planes && 2.0 * double(smin_k * smin_k) * double(o * o) * double(v * v) >= kB3MinFlops) {
      Anchor &an = e.anchor;
      int64_t rows = 0;
      for (int m = 0; m < nslices; ++m)
        for (int n = m; n < nslices; ++n) {
          an.ladder_off.push_back(rows);
          rows += ((hi[n] - lo[n]) * (hi[m] - lo[m]) + 255) / 256 * 256;
        }
      an.ladder_rows = rows;
      an.ladder_stack = DBuf(&e.pool, size_t((2 * rows * Kp_k * 2 + 7) / 8));
      stack_hi = reinterpret_cast<uint16_t *>(an.ladder_stack.get());
    }
    for (int m = 0, pair = 0; m < nslices && !planes; ++m)
      for (int n = m; n < nslices; ++n, ++pair) {  // cxx:418-451
        const int64_t Nx = hi[n] - lo[n], Ny = hi[m] - lo[m];
        View Vxycd = V4(Vxy.get(), Nx, Ny, v, v);
        View Rxyij = V4(Rxy.get(), Nx, Ny, o, o);
        ctr(e, 1.0, "Gxc,Gyd->xycd", V3(sl[n].get(), G, Nx, v), V3(sl[m].get(), G, Ny, v), 0.0,
            Vxycd, e.prec_ladder);  // cxx:431-434
        if (keep_planes) {
          e.anchor.ladder.emplace_back();
          B3FrozenA &fz = e.anchor.ladder.back();
          if (stack_hi) {
            const int64_t off = e.anchor.ladder_off[size_t(pair)] * Kp_k;
            const int Mp = int((Nx * Ny + 255) / 256 * 256);
            uint16_t *h = stack_hi + off, *l = stack_hi + e.anchor.ladder_rows * Kp_k + off;
            b3_split_into(e, Vxy.get(), 1, Nx * Ny, int(Nx * Ny), int(v * v), Mp, int(Kp_k), h, l);
            fz.M = int(Nx * Ny);
            fz.K = int(v * v);
            fz.Mp = Mp;
            fz.Kp = int(Kp_k);
            fz.ext = h;
            fz.lo_off = e.anchor.ladder_rows * Kp_k;
          } else if (2.0 * double(Nx * Ny) * double(o * o) * double(v * v) >= kB3MinFlops) {
            b3_freeze_a(e, Vxy.get(), 1, Nx * Ny, int(Nx * Ny), int(v * v), fz);
          }
        }
        ctr(e, 1.0, "xycd,cdij->xyij", Vxycd, X, 0.0, Rxyij, e.prec_ladder);  // cxx:443
        // sliceIntoResiduum (ClusterSinglesDoublesAlgorithm.cxx:887-915)
        const int64_t a = lo[n], b = lo[m];
        copy(e, subview(R2, {a, b, 0, 0}, {a + Nx, b + Ny, o, o}), Rxyij, 1.0, 1.0);
        if (a > b) sum(e, 1.0, Rxyij, "xyij", 1.0, subview(R2, {b, a, 0, 0}, {b + Ny, a + Nx, o, o}), "yxji");
      }
  }
  {
    // T1 equations (cxx:457-492)
    ctr(e, 1.0, "ac,ci->ai", Kac, T1, 1.0, R1, prec);    // cxx:466
    ctr(e, -1.0, "ki,ak->ai", Kki, T1, 1.0, R1, prec);   // cxx:467
    DBuf Kckb(&e.pool, v * o), Skib(&e.pool, o * o);
    View Kck = V2(Kckb.get(), v, o);
    ctr(e, 1.0, "cdkl,dl->ck", U, T1, 0.0, Kck, prec);   // cxx:470-471
    e.record("Kck", Kckb.get(), v * o);
    ctr(e, 2.0, "ck,caki->ai", Kck, T2, 1.0, R1, prec);  // cxx:474
    ctr(e, -1.0, "ck,caik->ai", Kck, T2, 1.0, R1, prec); // cxx:475
    ctr(e, 1.0, "ck,ci->ki", Kck, T1, 0.0, V2(Skib.get(), o, o), prec);  // cxx:476
    e.record("Ski", Skib.get(), o * o);
    ctr(e, 1.0, "ak,ki->ai", T1, V2(Skib.get(), o, o), 1.0, R1, prec);
    ctr(e, 2.0, "acik,ck->ai", V, T1, 1.0, R1, prec);      // cxx:477
    ctr(e, -1.0, "ciak,ck->ai", Vaibj, T1, 1.0, R1, prec); // cxx:478
    DBuf Ygb(&e.pool, G * v * o);
    View Yg = V3(Ygb.get(), G, v, o);
    ctr(e, 1.0, "Gdk,cdik->Gci", Gai, X, 0.0, Yg, prec);  // cxx:480-481
    ctr(e, 2.0, "Gca,Gci->ai", Gab, Yg, 1.0, R1, prec);
    ctr(e, 1.0, "Gck,cdik->Gdi", Gai, X, 0.0, Yg, prec);  // cxx:484-485
    ctr(e, -1.0, "Gda,Gdi->ai", Gab, Yg, 1.0, R1, prec);
    ctr(e, -2.0, "klic,ackl->ai", Vijka, X, 1.0, R1, prec);  // cxx:489
    ctr(e, 1.0, "lkic,ackl->ai", Vijka, X, 1.0, R1, prec);   // cxx:490
  }
  e.thaw();
}

// Residual at T = T_A + d by refinement around the recorded FP64 anchor
// T_A: R(T) = R(T_A) + dR, where dR is evaluated from the reference's term
// list with every product of T-dependent factors expanded as
//   P(T) Q(T) - P(T_A) Q(T_A) = dP Q(T) + P(T_A) dQ
// and every product with a constant as c dP.  Each contraction of dR carries
// a d factor, so its rounding error is relative to |d|, not to |R|: dR can
// run on the reduced-precision MFMA paths while R keeps FP64 accuracy.
// Elementwise intermediates (X, Y, dressed vertices) are formed in FP64 at T
// and differenced.  The anchor must have been recorded (e.recording) by an
// FP64 residual at T_A with the same slice size.
void ccsd_residual_delta(Engine &e, int slice_size, const double *T1c, const double *T2c,
                         double *R1p, double *R2p) {
  SISI_REQUIRE(!e.cplx, SISI_ERR_INVALID, "refinement residuals are not built for complex amplitudes");
  SISI_REQUIRE(e.NR == 0, SISI_ERR_INVALID, "refinement residuals are not built with CoulombFactors");
  SISI_REQUIRE(e.formulation == SISI_FORMULATION_VERTEX, SISI_ERR_INVALID,
               "refinement residuals are built for the vertex formulation only");
  Anchor &an = e.anchor;
  SISI_REQUIRE(an.valid, SISI_ERR_STATE, "refinement residual without an anchor");
  SISI_REQUIRE(an.ppl == e.ppl && an.distinguishable == e.distinguishable, SISI_ERR_STATE,
               "refinement residual with another residual variant than its anchor");
  const int64_t o = e.No, v = e.Nv, G = e.NG;
  const int64_t n2 = v * v * o * o, n1 = v * o;
  const int prec = e.precision;
  e.thaw();
  double *T1p = const_cast<double *>(T1c), *T2p = const_cast<double *>(T2c);
  View T1 = V2(T1p, v, o), T2 = V4(T2p, v, v, o, o);
  View T1A = V2(an.T1.get(), v, o);
  View V = V4(e.Vabij.get(), v, v, o, o), U = V4(e.Uabij.get(), v, v, o, o);
  View Vaibj = V4(e.Vaibj.get(), v, o, v, o), Vijka = V4(e.Vijka.get(), o, o, o, v);
  View Gij = V3(e.Gij.get(), G, o, o), Gai = V3(e.Gai.get(), G, v, o),
       Gab = V3(e.Gab.get(), G, v, v);
  auto A = [&](const char *name) -> double * {
    auto it = an.v.find(name);
    SISI_REQUIRE(it != an.v.end(), SISI_ERR_STATE, std::string("anchor lacks ") + name);
    return it->second.get();
  };
  // amplitude changes d1 = T1 - T1_A, d2 = T2 - T2_A (FP64)
  DBuf d1b(&e.pool, n1), d2b(&e.pool, n2);
  axpby(e, n1, 1.0, T1p, 0.0, d1b.get());
  axpby(e, n1, -1.0, an.T1.get(), 1.0, d1b.get());
  axpby(e, n2, 1.0, T2p, 0.0, d2b.get());
  axpby(e, n2, -1.0, an.T2.get(), 1.0, d2b.get());
  View d1 = V2(d1b.get(), v, o), d2 = V4(d2b.get(), v, v, o, o);
  e.frozen = {T1p, T2p, d1b.get(), d2b.get(), an.T1.get(), an.T2.get(), e.Vabij.get(),
              e.Uabij.get(), e.Vaibj.get(), e.Vijka.get(), e.Gij.get(), e.Gai.get(), e.Gab.get()};
  // X = T2 + T1 T1 at T and its change (FP64 elementwise)
  DBuf Xnb(&e.pool, n2), dXb(&e.pool, n2);
  add_t1t1(e, Xnb.get(), T2p, T1p, 1.0, int(v), int(o));
  add_t1t1(e, dXb.get(), an.T2.get(), an.T1.get(), 1.0, int(v), int(o));
  axpby(e, n2, 1.0, Xnb.get(), -1.0, dXb.get());
  View Xn = V4(Xnb.get(), v, v, o, o), dX = V4(dXb.get(), v, v, o, o);
  e.frozen.push_back(Xnb.get());
  e.frozen.push_back(dXb.get());
  DBuf dR2b(&e.pool, n2), dR1b(&e.pool, n1);
  zero(e, dR2b.get(), n2);
  zero(e, dR1b.get(), n1);
  View dR2 = V4(dR2b.get(), v, v, o, o), dR1 = V2(dR1b.get(), v, o);

  DBuf dKacb(&e.pool, v * v), dLacb(&e.pool, v * v), dKkib(&e.pool, o * o), dLkib(&e.pool, o * o);
  View dKac = V2(dKacb.get(), v, v), dLac = V2(dLacb.get(), v, v), dKki = V2(dKkib.get(), o, o),
       dLki = V2(dLkib.get(), o, o);
  {
    ctr(e, -1.0, "cdkl,adkl->ac", U, dX, 0.0, dKac, prec);
    DBuf dYdb;
    if (e.distinguishable) {  // change of Y = T2 + 2 T1 T1
      dYdb = DBuf(&e.pool, n2);
      DBuf YAb(&e.pool, n2);
      add_t1t1(e, dYdb.get(), T2p, T1p, 2.0, int(v), int(o));
      add_t1t1(e, YAb.get(), an.T2.get(), an.T1.get(), 2.0, int(v), int(o));
      axpby(e, n2, -1.0, YAb.get(), 1.0, dYdb.get());
      ctr(e, -0.5, "cdkl,adkl->ac", U, V4(dYdb.get(), v, v, o, o), 0.0, dLac, prec);
    } else {
      copy(e, dLac, dKac);
    }
    DBuf Zb(&e.pool, G), Yb(&e.pool, G * v * o);
    ctr(e, 1.0, "Gdk,dk->G", Gai, d1, 0.0, V1(Zb.get(), G), prec);
    ctr(e, 2.0, "Gca,G->ac", Gab, V1(Zb.get(), G), 1.0, dLac, prec);
    ctr(e, 1.0, "Gda,dk->Gak", Gab, d1, 0.0, V3(Yb.get(), G, v, o), prec);
    ctr(e, -1.0, "Gck,Gak->ac", Gai, V3(Yb.get(), G, v, o), 1.0, dLac, prec);
    ctr(e, 1.0, "cdkl,cdil->ki", U, dX, 0.0, dKki, prec);
    if (e.distinguishable)
      ctr(e, 0.5, "cdkl,cdil->ki", U, V4(dYdb.get(), v, v, o, o), 0.0, dLki, prec);
    else
      copy(e, dLki, dKki);
    ctr(e, 2.0, "klic,cl->ki", Vijka, d1, 1.0, dLki, prec);
    ctr(e, -1.0, "lkic,cl->ki", Vijka, d1, 1.0, dLki, prec);
    // Lac T2, Lki T2
    ctr(e, 1.0, "ac,cbij->abij", dLac, T2, 1.0, dR2, prec);
    ctr(e, 1.0, "ac,cbij->abij", V2(A("Lac"), v, v), d2, 1.0, dR2, prec);
    ctr(e, -1.0, "ki,abkj->abij", dLki, T2, 1.0, dR2, prec);
    ctr(e, -1.0, "ki,abkj->abij", V2(A("Lki"), o, o), d2, 1.0, dR2, prec);
  }
  {
    // dressed Gamma_ai (Gamma_ai - Gamma_ki T_ak) times Q = Gamma_bc T_cj
    DBuf ddGb(&e.pool, G * v * o), dQb(&e.pool, G * v * o), Qnb(&e.pool, G * v * o);
    View ddG = V3(ddGb.get(), G, v, o), dQ = V3(dQb.get(), G, v, o), Qn = V3(Qnb.get(), G, v, o);
    ctr(e, -1.0, "Gki,ak->Gai", Gij, d1, 0.0, ddG, prec);
    ctr(e, 1.0, "Gbc,cj->Gbj", Gab, d1, 0.0, dQ, prec);
    axpby(e, G * v * o, 1.0, A("Q"), 0.0, Qnb.get());
    axpby(e, G * v * o, 1.0, dQb.get(), 1.0, Qnb.get());
    ctr(e, 1.0, "Gai,Gbj->abij", ddG, Qn, 1.0, dR2, prec);
    ctr(e, 1.0, "Gai,Gbj->abij", V3(A("dGai2b"), G, v, o), dQ, 1.0, dR2, prec);
    ctr(e, -1.0, "jika,bk->abij", Vijka, d1, 1.0, dR2, prec);
    // T1 (V T1): d1 P(T) + T1_A dP
    DBuf dPb(&e.pool, v * o * o * o), Pnb(&e.pool, v * o * o * o);
    View dP = V4(dPb.get(), v, o, o, o), Pn = V4(Pnb.get(), v, o, o, o);
    ctr(e, 1.0, "acik,cj->aikj", V, d1, 0.0, dP, prec);
    axpby(e, v * o * o * o, 1.0, A("P"), 0.0, Pnb.get());
    axpby(e, v * o * o * o, 1.0, dPb.get(), 1.0, Pnb.get());
    ctr(e, -1.0, "bk,aikj->abij", d1, Pn, 1.0, dR2, prec);
    ctr(e, -1.0, "bk,aikj->abij", T1A, dP, 1.0, dR2, prec);
  }
  {
    // Xakic: (dressed Gamma_ai) Gamma_ck + V (Y, T2) terms; then Xakic Y2
    DBuf ddGb(&e.pool, G * v * o);
    View ddG = V3(ddGb.get(), G, v, o);
    ctr(e, -1.0, "Gil,al->Gai", Gij, d1, 0.0, ddG, prec);
    ctr(e, 1.0, "Gad,di->Gai", Gab, d1, 1.0, ddG, prec);
    DBuf dXkb(&e.pool, v * o * o * v);
    View dXakic = V4(dXkb.get(), v, o, o, v);
    ctr(e, 1.0, "Gai,Gck->akic", ddG, Gai, 0.0, dXakic, prec);
    DBuf dYb(&e.pool, n2), YAb(&e.pool, n2);
    add_t1t1(e, dYb.get(), T2p, T1p, 2.0, int(v), int(o));
    add_t1t1(e, YAb.get(), an.T2.get(), an.T1.get(), 2.0, int(v), int(o));
    axpby(e, n2, -1.0, YAb.get(), 1.0, dYb.get());
    YAb.reset();
    ctr(e, -0.5, "dclk,dail->akic", V, V4(dYb.get(), v, v, o, o), 1.0, dXakic, prec);
    ctr(e, 1.0, "dclk,adil->akic", V, d2, 1.0, dXakic, prec);
    if (!e.distinguishable) ctr(e, -0.5, "cdlk,adil->akic", V, d2, 1.0, dXakic, prec);
    // Y2 = 2 T2 - T2^(bc) at T and its change
    DBuf Y2b(&e.pool, n2);
    View Y2 = V4(Y2b.get(), v, v, o, o), dY2 = V4(dYb.get(), v, v, o, o);
    axpby(e, n2, 2.0, T2p, 0.0, Y2b.get());
    sum(e, -1.0, T2, "bckj", 1.0, Y2, "cbkj");
    axpby(e, n2, 2.0, d2b.get(), 0.0, dYb.get());
    sum(e, -1.0, d2, "bckj", 1.0, dY2, "cbkj");
    ctr(e, 1.0, "akic,cbkj->abij", dXakic, Y2, 1.0, dR2, prec);
    ctr(e, 1.0, "akic,cbkj->abij", V4(A("Xakic"), v, o, o, v), dY2, 1.0, dR2, prec);
  }
  {
    // Xakci = (dressed Gamma_ab)(dressed Gamma_ij) - V T2 / 2; then - Xakci T2 (two terms)
    DBuf ddGab(&e.pool, G * v * v), ddGij(&e.pool, G * o * o), Gijn(&e.pool, G * o * o);
    ctr(e, -1.0, "Gcl,al->Gac", Gai, d1, 0.0, V3(ddGab.get(), G, v, v), prec);
    ctr(e, 1.0, "Gdk,di->Gki", Gai, d1, 0.0, V3(ddGij.get(), G, o, o), prec);
    axpby(e, G * o * o, 1.0, A("dGij4"), 0.0, Gijn.get());
    axpby(e, G * o * o, 1.0, ddGij.get(), 1.0, Gijn.get());
    DBuf dXcb(&e.pool, v * o * v * o);
    View dXakci = V4(dXcb.get(), v, o, v, o);
    ctr(e, 1.0, "Gac,Gki->akci", V3(ddGab.get(), G, v, v), V3(Gijn.get(), G, o, o), 0.0, dXakci,
        prec);
    ctr(e, 1.0, "Gac,Gki->akci", V3(A("dGab4"), G, v, v), V3(ddGij.get(), G, o, o), 1.0, dXakci,
        prec);
    if (!e.distinguishable) ctr(e, -0.5, "cdlk,dail->akci", V, d2, 1.0, dXakci, prec);
    View XakciA = V4(A("Xakci"), v, o, v, o);
    ctr(e, -1.0, "akci,cbkj->abij", dXakci, T2, 1.0, dR2, prec);
    ctr(e, -1.0, "akci,cbkj->abij", XakciA, d2, 1.0, dR2, prec);
    ctr(e, -1.0, "bkci,ackj->abij", dXakci, T2, 1.0, dR2, prec);
    ctr(e, -1.0, "bkci,ackj->abij", XakciA, d2, 1.0, dR2, prec);
  }
  {
    DBuf Rt(&e.pool, n2);
    copy(e, V4(Rt.get(), v, v, o, o), dR2);
    sum(e, 1.0, V4(Rt.get(), v, v, o, o), "baji", 1.0, dR2, "abij");
  }
  {
    // Xklij (hole ladder): (Vijkl + Vijka T1) X and (V X) T2
    DBuf dXk(&e.pool, o * o * o * o);
    View dXklij = V4(dXk.get(), o, o, o, o);
    ctr(e, 1.0, "klic,cj->klij", Vijka, d1, 0.0, dXklij, prec);
    ctr(e, 1.0, "lkjc,ci->klij", Vijka, d1, 1.0, dXklij, prec);
    ctr(e, 1.0, "klij,abkl->abij", dXklij, Xn, 1.0, dR2, prec);
    ctr(e, 1.0, "klij,abkl->abij", V4(A("Xklij1"), o, o, o, o), dX, 1.0, dR2, prec);
    if (e.distinguishable) {  // change of T1 V T1: V contracted with dX - d2
      DBuf dWb(&e.pool, n2);
      axpby(e, n2, 1.0, dXb.get(), 0.0, dWb.get());
      axpby(e, n2, -1.0, d2b.get(), 1.0, dWb.get());
      ctr(e, 1.0, "cdkl,cdij->klij", V, V4(dWb.get(), v, v, o, o), 0.0, dXklij, prec);
    } else {
      ctr(e, 1.0, "cdkl,cdij->klij", V, dX, 0.0, dXklij, prec);
    }
    ctr(e, 1.0, "klij,abkl->abij", dXklij, e.ppl ? T2 : Xn, 1.0, dR2, prec);
    ctr(e, 1.0, "klij,abkl->abij", V4(A("Xklij2"), o, o, o, o), e.ppl ? d2 : dX, 1.0, dR2, prec);
  }
  if (e.ppl) {
    // ladder: V~(T) X(T) - V~(T_A) X(T_A) = dV~ X(T) + V~(T_A) dX per slice pair,
    // dV~ = dG[n]^T G(T)[m] + G_A[n]^T dG[m] for the dressed Gamma_ab
    if (slice_size <= 0) slice_size = int(o);
    const int nslices = int((v + slice_size - 1) / slice_size);
    std::vector<int64_t> lo(nslices), hi(nslices);
    for (int s = 0; s < nslices; ++s) {
      lo[s] = int64_t(s) * slice_size;
      hi[s] = std::min<int64_t>(lo[s] + slice_size, v);
    }
    const int64_t smax = std::min<int64_t>(slice_size, v);
    const int pl = e.prec_ladder;
    const int npairs = nslices * (nslices + 1) / 2;
    // V~(T_A) of each pair from the anchor's bf16 planes when the ladder runs
    // in bf16 or split-BF16x3 (no rebuild, no split), else rebuilt here
    const bool planes = (pl == SISI_BF16 || pl == SISI_BF16X3) && an.ladder_slice == slice_size &&
                        an.ladder.size() == size_t(npairs);
    // low-rank form, taken when the anchor kept S and every pair's planes and
    // every product is large enough for the bf16 kernels (the precision of
    // each product is then the one of the slice-pair form): with
    // D = dGab(T) - dGab(T_A) = -sum_k d1[a,k] Gai[G,b,k],
    //   dV~ = D^T G_A + G_A^T D + D^T D, and contracted with X(T):
    //   (D^T G_A) X -> -sum_k d1[a,k] QA[k,b,i,j], QA = S X (one product),
    //   (G_A^T D) X -> the same at (b,a,j,i) (X[c,d,i,j] = X[d,c,j,i]),
    //   (D^T D) X   -> sum_kl d1[a,k] d1[b,l] QW[k,l,i,j], QW = V X (V = Vabij);
    // every term carries a d1 factor, as the per-pair products do
    const int64_t smin = hi[nslices - 1] - lo[nslices - 1];
    const B3FrozenA &Sf = an.ladder_s;
    bool lowrank = planes && e.ladder_planes && !e.tape && get_b3_backend() == 0 &&
                   2.0 * double(smin * smin) * double(v * v) * double(std::min(G, o * o)) >=
                       kB3MinFlops &&
                   Sf.store.get() && Sf.M == int(o * v) && Sf.K == int(v * v);
    for (int p = 0; lowrank && p < npairs; ++p) lowrank = an.ladder[size_t(p)].valid();
    // dressed Gamma_ab change and value at T (the slice-pair forms)
    DBuf ddGb, Gnb;
    if (!lowrank) {
      ddGb = DBuf(&e.pool, G * v * v);
      Gnb = DBuf(&e.pool, G * v * v);
      ctr(e, -1.0, "Gbk,ak->Gab", Gai, d1, 0.0, V3(ddGb.get(), G, v, v), prec);
      axpby(e, G * v * v, 1.0, A("dGabP"), 0.0, Gnb.get());
      axpby(e, G * v * v, 1.0, ddGb.get(), 1.0, Gnb.get());
    }
    View ddG = V3(ddGb.get(), G, v, v);
    View GA = V3(A("dGabP"), G, v, v), Gn = V3(Gnb.get(), G, v, v);
    if (lowrank) {
      const int passes = pl == SISI_BF16 ? 1 : 3;
      DBuf QAb(&e.pool, o * v * o * o), QWb(&e.pool, o * o * o * o), Ycb(&e.pool, v * o * o * o),
          tAb(&e.pool, n2);
      GemmArgs gq;
      gq.B = Xnb.get();
      gq.sbk = 1;
      gq.sbn = v * v;
      gq.M = int(o * v);
      gq.N = int(o * o);
      gq.K = int(v * v);
      gq.C = QAb.get();
      gq.cm0 = 1;
      gq.cn0 = o * v;
      gq.dm0 = gq.M;
      gq.dn0 = gq.N;
      gq.alpha = 1.0;
      gq.beta = 0.0;
      gemm_b3_frozen_a(e, gq, Sf, passes);  // QA[k,b,i,j], rows k + o b
      // L collects the half of the ladder change that dR2 receives twice,
      // L + L^(baji): the first-order term, half the D^T D term, and per
      // slice pair V~(T_A) dX (halved on diagonal pairs, whose block holds
      // both orders of a and b), written by the products in place
      View QA = V4(QAb.get(), o, v, o, o), QW = V4(QWb.get(), o, o, o, o),
           Yc = V4(Ycb.get(), v, o, o, o), L = V4(tAb.get(), v, v, o, o);
      ctr(e, -1.0, "ak,kbij->abij", d1, QA, 0.0, L, prec);
      ctr(e, 0.5, "cdkl,cdij->klij", V, Xn, 0.0, QW, prec);
      ctr(e, 1.0, "ak,klij->alij", d1, QW, 0.0, Yc, prec);
      ctr(e, 1.0, "bl,alij->abij", d1, Yc, 1.0, L, prec);
      if (an.ladder_rows > 0) {
        // every pair in one product over the stacked planes, then each pair's
        // rows (x + Nx y from its offset) added into its block of L
        const int64_t rows = an.ladder_rows;
        DBuf Cst(&e.pool, rows * o * o);
        B3FrozenA st;
        st.M = int(rows);
        st.K = int(v * v);
        st.Mp = int(rows);
        st.Kp = an.ladder[0].Kp;
        st.ext = an.ladder[0].ext;
        st.lo_off = an.ladder[0].lo_off;
        GemmArgs gs;
        gs.B = dXb.get();
        gs.sbk = 1;
        gs.sbn = v * v;
        gs.M = int(rows);
        gs.N = int(o * o);
        gs.K = int(v * v);
        gs.C = Cst.get();
        gs.cm0 = 1;
        gs.cn0 = rows;
        gs.dm0 = gs.M;
        gs.dn0 = gs.N;
        gs.alpha = 1.0;
        gs.beta = 0.0;
        gemm_b3_frozen_a(e, gs, st, passes);
        int pair = 0;
        for (int m = 0; m < nslices; ++m)
          for (int n = m; n < nslices; ++n, ++pair) {
            const int64_t Nx = hi[n] - lo[n], Ny = hi[m] - lo[m];
            const int64_t a0 = lo[n], b0 = lo[m];
            View Cp = dense_strided(Cst.get() + an.ladder_off[size_t(pair)], {Nx, Ny, o, o},
                                    {1, Nx, rows, rows * o});
            copy(e, subview(L, {a0, b0, 0, 0}, {a0 + Nx, b0 + Ny, o, o}), Cp,
                 n == m ? 0.5 : 1.0, 1.0);
          }
      }
      int pair = 0;
      for (int m = 0; m < nslices && an.ladder_rows == 0; ++m)
        for (int n = m; n < nslices; ++n, ++pair) {
          // L[a0 + x, b0 + y, i, j] += V~_A(x + Nx y, c + v d) dX(c + v d, i + o j)
          const int64_t Nx = hi[n] - lo[n], Ny = hi[m] - lo[m];
          const int64_t a0 = lo[n], b0 = lo[m];
          GemmArgs g2;
          g2.B = dXb.get();
          g2.sbk = 1;
          g2.sbn = v * v;
          g2.M = int(Nx * Ny);
          g2.N = int(o * o);
          g2.K = int(v * v);
          g2.C = tAb.get() + a0 + v * b0;
          g2.dm0 = int(Nx);
          g2.dm1 = int(Ny);
          g2.cm0 = 1;
          g2.cm1 = v;
          g2.dn0 = g2.N;
          g2.cn0 = v * v;
          g2.alpha = n == m ? 0.5 : 1.0;
          g2.beta = 1.0;
          gemm_b3_frozen_a(e, g2, an.ladder[size_t(pair)], passes);
        }
      sum(e, 1.0, L, "abij", 1.0, dR2, "abij");
      sum(e, 1.0, L, "baji", 1.0, dR2, "abij");
    }
    std::vector<DBuf> sA(lowrank ? 0 : nslices), sN(lowrank ? 0 : nslices), sD(lowrank ? 0 : nslices);
    for (int s = 0; s < nslices && !lowrank; ++s) {
      const int64_t n = G * (hi[s] - lo[s]) * v;
      sA[s] = DBuf(&e.pool, n);
      sN[s] = DBuf(&e.pool, n);
      sD[s] = DBuf(&e.pool, n);
      copy(e, V3(sA[s].get(), G, hi[s] - lo[s], v), subview(GA, {0, lo[s], 0}, {G, hi[s], v}));
      copy(e, V3(sN[s].get(), G, hi[s] - lo[s], v), subview(Gn, {0, lo[s], 0}, {G, hi[s], v}));
      copy(e, V3(sD[s].get(), G, hi[s] - lo[s], v), subview(ddG, {0, lo[s], 0}, {G, hi[s], v}));
      e.frozen.push_back(sA[s].get());
      e.frozen.push_back(sN[s].get());
      e.frozen.push_back(sD[s].get());
    }
    DBuf VAb, dVb, Rxy;
    if (!lowrank) {
      VAb = DBuf(&e.pool, smax * smax * v * v);
      dVb = DBuf(&e.pool, smax * smax * v * v);
      Rxy = DBuf(&e.pool, smax * smax * o * o);
    }
    int pair = 0;
    for (int m = 0; m < nslices && !lowrank; ++m)
      for (int n = m; n < nslices; ++n, ++pair) {
        const int64_t Nx = hi[n] - lo[n], Ny = hi[m] - lo[m];
        View VA = V4(VAb.get(), Nx, Ny, v, v), dV = V4(dVb.get(), Nx, Ny, v, v);
        View Rxyij = V4(Rxy.get(), Nx, Ny, o, o);
        const B3FrozenA *fa =
            planes && an.ladder[size_t(pair)].valid() ? &an.ladder[size_t(pair)] : nullptr;
        if (!fa)
          ctr(e, 1.0, "Gxc,Gyd->xycd", V3(sA[n].get(), G, Nx, v), V3(sA[m].get(), G, Ny, v), 0.0,
              VA, pl);
        ctr(e, 1.0, "Gxc,Gyd->xycd", V3(sD[n].get(), G, Nx, v), V3(sN[m].get(), G, Ny, v), 0.0,
            dV, pl);
        ctr(e, 1.0, "Gxc,Gyd->xycd", V3(sA[n].get(), G, Nx, v), V3(sD[m].get(), G, Ny, v), 1.0,
            dV, pl);
        ctr(e, 1.0, "xycd,cdij->xyij", dV, Xn, 0.0, Rxyij, pl);
        if (fa) {
          // Rxyij(m = x + Nx y, n = i + o j) += sum_k V~_A(m, k = c + v d) dX(k, n)
          GemmArgs g;
          g.B = dXb.get();
          g.sbk = 1;
          g.sbn = v * v;
          g.C = Rxy.get();
          g.M = int(Nx * Ny);
          g.N = int(o * o);
          g.K = int(v * v);
          g.cm0 = 1;
          g.cn0 = Nx * Ny;
          g.dm0 = g.M;
          g.dn0 = g.N;
          g.alpha = 1.0;
          g.beta = 1.0;
          gemm_b3_frozen_a(e, g, *fa, pl == SISI_BF16 ? 1 : 3);
        } else {
          ctr(e, 1.0, "xycd,cdij->xyij", VA, dX, 1.0, Rxyij, pl);
        }
        const int64_t a = lo[n], b = lo[m];
        copy(e, subview(dR2, {a, b, 0, 0}, {a + Nx, b + Ny, o, o}), Rxyij, 1.0, 1.0);
        if (a > b)
          sum(e, 1.0, Rxyij, "xyij", 1.0, subview(dR2, {b, a, 0, 0}, {b + Ny, a + Nx, o, o}),
              "yxji");
      }
    for (int s2 = 0; s2 < nslices && !lowrank; ++s2) {  // slices return to the pool at block end
      e.unfreeze(sA[s2].get());
      e.unfreeze(sN[s2].get());
      e.unfreeze(sD[s2].get());
    }
  }
  {
    // T1 equations
    ctr(e, 1.0, "ac,ci->ai", dKac, T1, 1.0, dR1, prec);
    ctr(e, 1.0, "ac,ci->ai", V2(A("Kac"), v, v), d1, 1.0, dR1, prec);
    ctr(e, -1.0, "ki,ak->ai", dKki, T1, 1.0, dR1, prec);
    ctr(e, -1.0, "ki,ak->ai", V2(A("Kki"), o, o), d1, 1.0, dR1, prec);
    DBuf dKckb(&e.pool, v * o), Kcknb(&e.pool, v * o), dSb(&e.pool, o * o), Snb(&e.pool, o * o);
    View dKck = V2(dKckb.get(), v, o), Kckn = V2(Kcknb.get(), v, o), KckA = V2(A("Kck"), v, o);
    ctr(e, 1.0, "cdkl,dl->ck", U, d1, 0.0, dKck, prec);
    axpby(e, v * o, 1.0, A("Kck"), 0.0, Kcknb.get());
    axpby(e, v * o, 1.0, dKckb.get(), 1.0, Kcknb.get());
    ctr(e, 2.0, "ck,caki->ai", dKck, T2, 1.0, dR1, prec);
    ctr(e, 2.0, "ck,caki->ai", KckA, d2, 1.0, dR1, prec);
    ctr(e, -1.0, "ck,caik->ai", dKck, T2, 1.0, dR1, prec);
    ctr(e, -1.0, "ck,caik->ai", KckA, d2, 1.0, dR1, prec);
    View dS = V2(dSb.get(), o, o), Sn = V2(Snb.get(), o, o);
    ctr(e, 1.0, "ck,ci->ki", dKck, T1, 0.0, dS, prec);
    ctr(e, 1.0, "ck,ci->ki", KckA, d1, 1.0, dS, prec);
    axpby(e, o * o, 1.0, A("Ski"), 0.0, Snb.get());
    axpby(e, o * o, 1.0, dSb.get(), 1.0, Snb.get());
    ctr(e, 1.0, "ak,ki->ai", d1, Sn, 1.0, dR1, prec);
    ctr(e, 1.0, "ak,ki->ai", T1A, dS, 1.0, dR1, prec);
    ctr(e, 2.0, "acik,ck->ai", V, d1, 1.0, dR1, prec);
    ctr(e, -1.0, "ciak,ck->ai", Vaibj, d1, 1.0, dR1, prec);
    DBuf Ygb(&e.pool, G * v * o);
    View Yg = V3(Ygb.get(), G, v, o);
    ctr(e, 1.0, "Gdk,cdik->Gci", Gai, dX, 0.0, Yg, prec);
    ctr(e, 2.0, "Gca,Gci->ai", Gab, Yg, 1.0, dR1, prec);
    ctr(e, 1.0, "Gck,cdik->Gdi", Gai, dX, 0.0, Yg, prec);
    ctr(e, -1.0, "Gda,Gdi->ai", Gab, Yg, 1.0, dR1, prec);
    ctr(e, -2.0, "klic,ackl->ai", Vijka, dX, 1.0, dR1, prec);
    ctr(e, 1.0, "lkic,ackl->ai", Vijka, dX, 1.0, dR1, prec);
  }
  e.thaw();
  // R(T) = R(T_A) + dR (FP64)
  axpby(e, n2, 1.0, an.R2.get(), 0.0, R2p);
  axpby(e, n2, 1.0, dR2b.get(), 1.0, R2p);
  axpby(e, n1, 1.0, an.R1.get(), 0.0, R1p);
  axpby(e, n1, 1.0, dR1b.get(), 1.0, R1p);
}

// FP64 residual at T that also becomes the new refinement anchor.
void ccsd_residual_anchor(Engine &e, int slice_size, const double *T1, const double *T2,
                          double *R1, double *R2, int precision) {
  SISI_REQUIRE(!e.cplx, SISI_ERR_INVALID, "refinement residuals are not built for complex amplitudes");
  SISI_REQUIRE(e.NR == 0, SISI_ERR_INVALID, "refinement residuals are not built with CoulombFactors");
  SISI_REQUIRE(is_precision(precision), SISI_ERR_INVALID, "unknown anchor precision");
  const int64_t n1 = int64_t(e.n1()), n2 = int64_t(e.n2());
  const int p = e.precision, pl = e.prec_ladder;
  e.precision = precision;
  e.prec_ladder = precision;
  e.anchor.valid = false;
  e.anchor.ladder.clear();
  e.anchor.ladder_s = B3FrozenA();
  e.recording = true;
  try {
    ccsd_residual(e, 1, true, slice_size, T1, T2, R1, R2);
  } catch (...) {
    e.recording = false;
    e.precision = p;
    e.prec_ladder = pl;
    throw;
  }
  e.recording = false;
  e.precision = p;
  e.prec_ladder = pl;
  Anchor &an = e.anchor;
  if (an.T1.size() != size_t(n1)) an.T1 = DBuf(&e.pool, n1);
  if (an.T2.size() != size_t(n2)) an.T2 = DBuf(&e.pool, n2);
  if (an.R1.size() != size_t(n1)) an.R1 = DBuf(&e.pool, n1);
  if (an.R2.size() != size_t(n2)) an.R2 = DBuf(&e.pool, n2);
  axpby(e, n1, 1.0, T1, 0.0, an.T1.get());
  axpby(e, n2, 1.0, T2, 0.0, an.T2.get());
  axpby(e, n1, 1.0, R1, 0.0, an.R1.get());
  axpby(e, n2, 1.0, R2, 0.0, an.R2.get());
  an.valid = true;
  an.age = 0;
  an.precision = precision;
  an.ppl = e.ppl;
  an.distinguishable = e.distinguishable;
}

// The HPFockMatrix term spins * sum_ai T_ai f_ia (cxx:199-211; uploaded for
// a real vertex only).
static double noncanonical_term(Engine &e, const double *T1) {
  if (!e.fai.get()) return 0.0;
  SISI_REQUIRE(e.fai_cplx == e.cplx, SISI_ERR_STATE,
               "HPFockMatrix was set for the other (real / complex) mode of the vertex");
  const int64_t n1 = int64_t(e.Nv) * e.No;
  return e.spins * dot(e, n1, T1, e.fai.get());
}

void ccsd_energy(Engine &e, const double *T1, const double *T2, double &E, double &direct,
                 double &exchange, double *E_imag) {
  SISI_REQUIRE(integrals_ready(e), SISI_ERR_STATE, "energy: integrals not built");
  if (e.cplx) {  // F = complex: the real parts are reported (cxx:30-33, 185-189)
    SISI_REQUIRE(!e.anti_energy, SISI_ERR_INVALID,
                 "energy: the antisymmetrized energy is built for real amplitudes");
    double Ez[2], dz[2], xz[2];
    zccsd_energy(e, T1, T2, Ez, dz, xz);
    // spins (unrestricted) as in the real branch
    for (int k = 0; k < 2; ++k) {
      dz[k] *= 0.25 * e.spins * e.spins;
      xz[k] *= 0.5 * e.spins;
      Ez[k] = dz[k] + xz[k];
    }
    E = Ez[0];
    direct = dz[0];
    exchange = xz[0];
    if (E_imag) *E_imag = Ez[1];
    e.noncanonical = noncanonical_term(e, T1);
    E += e.noncanonical;
    return;
  }
  if (e.formulation == SISI_FORMULATION_UCCSD) {  // antisymmetrized (cxx:163-167)
    E = direct = uccsd_energy(e, T1, T2);
    exchange = 0.0;
    if (E_imag) *E_imag = 0.0;
    e.noncanonical = 0.0;
    return;
  }
  if (e.anti_energy) {
    // 0.25 T_abkl V_klab + 0.5 T_aj T_cl V_jlac (cxx:163-167) with V_ijab =
    // V_abij: 0.25 sum_abij (T2 + 2 T1 T1)_abij V_abij
    const int64_t n2 = e.n2();
    DBuf X(&e.pool, n2);
    add_t1t1(e, X.get(), T2, T1, 2.0, e.Nv, e.No);
    E = direct = 0.25 * dot(e, n2, X.get(), e.Vabij.get());
    exchange = 0.0;
  } else {
    energy_terms(e, T1, T2, direct, exchange);
    // direct 1/2 spins^2 tau V_ijab, exchange -1/2 spins tau V_ijba
    // (ClusterSinglesDoublesAlgorithm.cxx:163-170); energy_terms uses spins = 2
    direct *= 0.25 * e.spins * e.spins;
    exchange *= 0.5 * e.spins;
    E = direct + exchange;
  }
  if (E_imag) *E_imag = 0.0;
  e.noncanonical = noncanonical_term(e, T1);
  E += e.noncanonical;
}

void ccsd_pair_energies(Engine &e, const double *T1, const double *T2, double *pair) {
  SISI_REQUIRE(integrals_ready(e), SISI_ERR_STATE, "pair energies: integrals not built");
  SISI_REQUIRE(e.formulation != SISI_FORMULATION_UCCSD, SISI_ERR_INVALID,
               "pair energies are built for the vertex and full-integral CCSD energies");
  if (e.cplx) {
    SISI_REQUIRE(!e.anti_energy, SISI_ERR_INVALID,
                 "pair energies: the antisymmetrized form is built for real amplitudes");
    return zccsd_pair_energies(e, T1, T2, pair);
  }
  const int64_t o = e.No, v = e.Nv;
  DBuf Xb(&e.pool, v * v * o * o);
  if (e.anti_energy) {
    // 0.25 T_abij V_ijab + 0.5 T_ai T_bj V_ijab (cxx:222-226)
    SISI_REQUIRE(e.Vabij.get(), SISI_ERR_STATE, "pair energies: integrals not built");
    add_t1t1(e, Xb.get(), T2, T1, 2.0, int(v), int(o));
    ctr(e, 0.25, "abij,abij->ij", V4(Xb.get(), v, v, o, o), V4(e.Vabij.get(), v, v, o, o), 0.0,
        V2(pair, o, o), SISI_FP64);
    return;
  }
  SISI_REQUIRE(e.Uabij.get(), SISI_ERR_STATE, "pair energies: integrals not built");
  add_t1t1(e, Xb.get(), T2, T1, 1.0, int(v), int(o));  // tau = T2 + T1 T1
  // the four terms of cxx:226-231 are tau_abij (2 V_abij - V_baij) = tau U
  // (fixed factors: the reference's pair energies do not read `unrestricted`)
  ctr(e, 1.0, "abij,abij->ij", V4(Xb.get(), v, v, o, o), V4(e.Uabij.get(), v, v, o, o), 0.0,
      V2(pair, o, o), SISI_FP64);
}

// ---------------------------------------------------------------------------
// Driver
// ---------------------------------------------------------------------------

// Solve the bordered DIIS system (dsysv_ in DiisMixer.cxx:16-40) in FP64 by
// Gaussian elimination with partial pivoting.
std::vector<double> diis_solve(std::vector<std::vector<double>> A, std::vector<double> b) {
  const int n = int(b.size());
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
    SISI_REQUIRE(std::isfinite(A[p][c]), SISI_ERR_NUMERIC,
                 "DIIS: non-finite overlap (the iteration diverged)");
    SISI_REQUIRE(A[p][c] != 0.0, SISI_ERR_NUMERIC, "DIIS: singular overlap matrix");
    std::swap(A[p], A[c]);
    std::swap(b[p], b[c]);
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r][c] / A[c][c];
      for (int k = c; k < n; ++k) A[r][k] -= f * A[c][k];
      b[r] -= f * b[c];
    }
  }
  std::vector<double> x(n);
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
    for (int k = r + 1; k < n; ++k) s -= A[r][k] * x[k];
    x[r] = s / A[r][r];
  }
  return x;
}

namespace {

double fock_dot(Engine &e, const double *x1, const double *x2, const double *y1,
                const double *y2) {
  return dot(e, int64_t(e.n1()), x1, y1) + dot(e, int64_t(e.n2()), x2, y2);
}

// Residual replayed from a hipGraph.  The plain residual issues only kernels
// and device copies on e.stream (no host synchronisation, no memsets), so
// one capture per configuration serves every later iteration: T is copied
// into the graph's fixed input buffers, the graph is launched and R copied
// out.  For small systems the residual is bound by launch overhead (a few
// hundred launches of a few microseconds each), which the replay removes.
// Every pool buffer the captured launches touch stays reserved until the
// graph is dropped.  Returns false (and the caller runs the residual
// eagerly) when the stream cannot be captured.
bool graph_residual(Engine &e, int slice, const double *T1, const double *T2, double *R1,
                    double *R2) {
  Engine::ResidualGraph &g = e.rgraph;
  const std::vector<int> key = {e.precision,       slice,   e.formulation,  int(e.ppl),
                                int(e.distinguishable), e.NR, int(e.cplx), int(e.run_initial),
                                get_b3_backend()};
  if (g.exec && g.key != key) e.drop_residual_graph();
  const int64_t n1 = int64_t(e.n1()), n2 = int64_t(e.n2());
  if (!g.exec) {
    e.drop_residual_graph();
    g.T1 = DBuf(&e.pool, size_t(n1));
    g.T2 = DBuf(&e.pool, size_t(n2));
    g.R1 = DBuf(&e.pool, size_t(n1));
    g.R2 = DBuf(&e.pool, size_t(n2));
    copy(e, V1(g.T1.get(), n1), V1(const_cast<double *>(T1), n1));
    copy(e, V1(g.T2.get(), n2), V1(const_cast<double *>(T2), n2));
    // eager pass first: library handles, algorithm choices and lazily
    // created state are set up outside the capture
    ccsd_residual(e, 1, e.run_initial, slice, g.T1.get(), g.T2.get(), g.R1.get(), g.R2.get());
    e.sync();
    hipGraph_t graph = nullptr;
    // no device-wide operation of another thread may run during the capture
    std::lock_guard<std::recursive_mutex> lock(device_wide_mutex());
    e.pool.begin_capture();
    hipError_t err = hipStreamBeginCapture(e.stream, hipStreamCaptureModeRelaxed);
    if (err == hipSuccess) {
      try {
        ccsd_residual(e, 1, e.run_initial, slice, g.T1.get(), g.T2.get(), g.R1.get(), g.R2.get());
      } catch (...) {
        (void)hipStreamEndCapture(e.stream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        e.pool.release(e.pool.end_capture());
        e.drop_residual_graph();
        throw;
      }
      err = hipStreamEndCapture(e.stream, &graph);
    }
    g.held = e.pool.end_capture();
    if (err == hipSuccess) err = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (err != hipSuccess) {
      (void)hipGetLastError();
      g.exec = nullptr;
      e.drop_residual_graph();
      return false;
    }
    g.key = key;
  } else {
    copy(e, V1(g.T1.get(), n1), V1(const_cast<double *>(T1), n1));
    copy(e, V1(g.T2.get(), n2), V1(const_cast<double *>(T2), n2));
  }
  SISI_HIP(hipGraphLaunch(g.exec, e.stream));
  copy(e, V1(R1, n1), V1(g.R1.get(), n1));
  copy(e, V1(R2, n2), V1(g.R2.get(), n2));
  return true;
}

}  // namespace

void ccsd_begin(Engine &e, const sisi_ccsd_opts &opts, const double *T1h, const double *T2h) {
  SISI_REQUIRE(integrals_ready(e), SISI_ERR_STATE, "ccsd: integrals not built");
  SISI_REQUIRE(opts.mixer == SISI_MIXER_LINEAR || opts.mixer == SISI_MIXER_DIIS,
               SISI_ERR_INVALID, "Mixer not implemented");
  SISI_REQUIRE(opts.mixer != SISI_MIXER_DIIS || opts.max_residua >= 1, SISI_ERR_INVALID,
               "maxResidua must be >= 1");
  SISI_REQUIRE(opts.refine_interval >= 0, SISI_ERR_INVALID, "refine_interval must be >= 0");
  SISI_REQUIRE(opts.bulk_refine_interval >= 0, SISI_ERR_INVALID,
               "bulk_refine_interval must be >= 0");
  SISI_REQUIRE(opts.bulk_refine_interval == 0 || opts.precision == SISI_FP64 ||
                   (!e.cplx && e.NR == 0 && e.formulation == SISI_FORMULATION_VERTEX),
               SISI_ERR_INVALID,
               "bulk refinement residuals are built for the real vertex formulation only");
  SISI_REQUIRE(!e.cplx || opts.precision == SISI_FP64 || opts.refine_interval == 0,
               SISI_ERR_INVALID, "refinement residuals are not built for complex amplitudes");
  SISI_REQUIRE(e.NR == 0 || opts.precision == SISI_FP64 || opts.refine_interval == 0,
               SISI_ERR_INVALID, "refinement residuals are not built with CoulombFactors");
  SISI_REQUIRE(e.formulation == SISI_FORMULATION_VERTEX || opts.precision == SISI_FP64 ||
                   opts.refine_interval == 0,
               SISI_ERR_INVALID, "refinement residuals are built for the vertex formulation only");
  SISI_REQUIRE(opts.refine_precision == SISI_FP32 || opts.refine_precision == SISI_BF16X3 ||
                   opts.refine_precision == SISI_BF16,
               SISI_ERR_INVALID, "refine_precision must be SISI_FP32, SISI_BF16X3 or SISI_BF16");
  e.drop_residual_graph();
  e.opts = opts;
  e.ppl = opts.ppl != 0;
  e.distinguishable = opts.distinguishable != 0;
  e.T1 = DBuf(&e.pool, e.n1());
  e.T2 = DBuf(&e.pool, e.n2());
  e.run_initial = (T2h != nullptr);
  if (T1h) e.upload_t(e.T1, T1h, e.n1());
  else zero(e, e.T1.get(), int64_t(e.n1()));
  if (T2h) e.upload_t(e.T2, T2h, e.n2());
  else zero(e, e.T2.get(), int64_t(e.n2()));
  e.iteration = 0;
  e.fp64_iterations = 0;
  e.prev_energy = e.prev_energy_imag = 0.0;
  e.last_energy = e.last_direct = e.last_exchange = e.last_energy_imag = 0.0;
  e.last_delta = 1.0;
  e.converged = false;
  e.confirm_pending = false;
  e.confirm_hold = 0;
  e.energies.clear();
  e.diis.clear();
  e.B.clear();
  e.diis_next = e.diis_count = 0;
  e.lin_have = false;
  e.lin1.reset();
  e.lin2.reset();
  e.anchor = Anchor();
  if (opts.mixer == SISI_MIXER_DIIS) {
    const int N = opts.max_residua;
    e.diis.resize(N);
    e.B.assign(N + 1, std::vector<double>(N + 1, 0.0));
    for (int k = 1; k <= N; ++k) e.B[0][k] = e.B[k][0] = -1.0;  // DiisMixer.cxx:83-96
  }
  e.cur_precision = opts.precision;
  SISI_REQUIRE(opts.coarse_threshold <= 0.0 || is_precision(opts.coarse_precision), SISI_ERR_INVALID,
               "coarse_precision must be a precision");
  e.coarse_active = opts.precision != SISI_FP64 && opts.coarse_threshold > 0.0;
  e.run_active = true;
}

void ccsd_iterate(Engine &e, int n) {
  SISI_REQUIRE(e.run_active, SISI_ERR_STATE, "ccsd: no run (call begin first)");
  const sisi_ccsd_opts &o = e.opts;
  for (int step = 0; step < n && !e.converged && e.iteration < o.max_iterations; ++step) {
    // precision schedule: coarse precision until |dT|/|T| <= coarse_threshold
    // (if set), bulk precision until |dT|/|T| <= switch_threshold, FP64
    // afterwards (and FP64 throughout for the plain path)
    const bool bulk = o.precision != SISI_FP64 && e.cur_precision != SISI_FP64;
    const bool refine = !bulk && o.precision != SISI_FP64 && o.refine_interval > 0;
    // bulk refinement: anchors in the bulk precision (not in the coarse
    // stage) every bulk_refine_interval iterations, refinement residuals
    // between; the anchor error (the bulk precision's) is what the bulk
    // phase tolerates anyway, the refinement error scales with the change
    const bool bulk_refine = bulk && !e.coarse_active && o.bulk_refine_interval > 0;
    int prec = bulk ? (e.coarse_active ? o.coarse_precision : o.precision) : SISI_FP64;
    DBuf R1(&e.pool, e.n1()), R2(&e.pool, e.n2());
    if (bulk_refine) {
      if (!e.anchor.valid || e.anchor.precision != prec ||
          e.anchor.age >= o.bulk_refine_interval) {
        set_precision(e, prec);
        ccsd_residual_anchor(e, o.slice_size, e.T1.get(), e.T2.get(), R1.get(), R2.get(), prec);
      } else {
        set_precision(e, o.refine_precision);
        ccsd_residual_delta(e, o.slice_size, e.T1.get(), e.T2.get(), R1.get(), R2.get());
        e.anchor.age++;
        prec = o.refine_precision;
      }
    } else if (refine) {
      // refinement phase: FP64 anchor every refine_interval iterations,
      // reduced-precision refinement residuals in between
      if (!e.anchor.valid || e.anchor.precision != SISI_FP64 ||
          e.anchor.age >= o.refine_interval) {
        set_precision(e, SISI_FP64);
        ccsd_residual_anchor(e, o.slice_size, e.T1.get(), e.T2.get(), R1.get(), R2.get());
        prec = SISI_FP64;
      } else {
        set_precision(e, o.refine_precision);
        ccsd_residual_delta(e, o.slice_size, e.T1.get(), e.T2.get(), R1.get(), R2.get());
        e.anchor.age++;
        prec = o.refine_precision;
      }
    } else {
      set_precision(e, prec);
      // graphed residual after the MP2 start (which is a copy) and with
      // launch timing off
      const bool graphed = o.graphs != 0 && (e.iteration > 0 || e.run_initial) &&
                           !e.timers->enabled &&
                           graph_residual(e, o.slice_size, e.T1.get(), e.T2.get(), R1.get(),
                                          R2.get());
      if (!graphed)
        ccsd_residual(e, e.iteration, e.run_initial, o.slice_size, e.T1.get(), e.T2.get(),
                      R1.get(), R2.get());
    }
    DBuf d1(&e.pool, e.n1()), d2(&e.pool, e.n2());
    estimate_amplitudes(e, 2, R1.get(), e.T1.get(), d1.get(), o.level_shift);
    estimate_amplitudes(e, 4, R2.get(), e.T2.get(), d2.get(), o.level_shift);
    if (o.mixer == SISI_MIXER_DIIS) {
      // DiisMixer::append (DiisMixer.cxx:103-180)
      const int N = o.max_residua;
      DiisSlot &s = e.diis[e.diis_next];
      s.t1 = std::move(R1);
      s.t2 = std::move(R2);
      s.d1 = std::move(d1);
      s.d2 = std::move(d2);
      for (int i = 0; i < N; ++i) {
        if (!e.diis[i].d2.get()) continue;
        const double ov = 2.0 * fock_dot(e, e.diis[i].d1.get(), e.diis[i].d2.get(), s.d1.get(),
                                         s.d2.get());
        e.B[e.diis_next + 1][i + 1] = ov;
        e.B[i + 1][e.diis_next + 1] = ov;
      }
      if (e.diis_count < N) ++e.diis_count;
      const int dim = e.diis_count + 1;
      std::vector<std::vector<double>> Bm(dim, std::vector<double>(dim));
      for (int r = 0; r < dim; ++r)
        for (int c = 0; c < dim; ++c) Bm[r][c] = e.B[r][c];
      std::vector<double> rhs(dim, 0.0);
      rhs[0] = -1.0;
      std::vector<double> col = diis_solve(Bm, rhs);
      DBuf n1(&e.pool, e.n1()), n2(&e.pool, e.n2());
      zero(e, n1.get(), int64_t(e.n1()));
      zero(e, n2.get(), int64_t(e.n2()));
      for (int j = 0; j < e.diis_count; ++j) {
        const int i = (e.diis_next + N - j) % N;
        axpby(e, int64_t(e.n1()), col[i + 1], e.diis[i].t1.get(), 1.0, n1.get());
        axpby(e, int64_t(e.n2()), col[i + 1], e.diis[i].t2.get(), 1.0, n2.get());
      }
      e.diis_next = (e.diis_next + 1) % N;
      e.T1 = std::move(n1);
      e.T2 = std::move(n2);
      // amplitude change of this iteration, read back for the test below
      const DiisSlot &cur = e.diis[(e.diis_next + N - 1) % N];
      const double dd = fock_dot(e, cur.d1.get(), cur.d2.get(), cur.d1.get(), cur.d2.get());
      const double tt = fock_dot(e, e.T1.get(), e.T2.get(), e.T1.get(), e.T2.get());
      e.last_delta = std::sqrt(std::fabs(dd / tt));
    } else {
      // LinearMixer::append (LinearMixer.cxx:30-44)
      if (e.lin_have) {
        axpby(e, int64_t(e.n1()), 1.0 - o.mixing_ratio, e.lin1.get(), o.mixing_ratio, R1.get());
        axpby(e, int64_t(e.n2()), 1.0 - o.mixing_ratio, e.lin2.get(), o.mixing_ratio, R2.get());
      }
      e.lin1 = std::move(R1);
      e.lin2 = std::move(R2);
      e.lin_have = true;
      DBuf n1(&e.pool, e.n1()), n2(&e.pool, e.n2());
      copy(e, V1(n1.get(), int64_t(e.n1())), V1(e.lin1.get(), int64_t(e.n1())));
      copy(e, V1(n2.get(), int64_t(e.n2())), V1(e.lin2.get(), int64_t(e.n2())));
      e.T1 = std::move(n1);
      e.T2 = std::move(n2);
      const double dd = fock_dot(e, d1.get(), d2.get(), d1.get(), d2.get());
      const double tt = fock_dot(e, e.T1.get(), e.T2.get(), e.T1.get(), e.T2.get());
      e.last_delta = std::sqrt(std::fabs(dd / tt));
    }
    // mixed runs keep T2 exactly symmetric (T_abij = T_baji): the reduced-
    // precision and refinement residuals treat an antisymmetric part of T2
    // differently from the FP64 one, and DIIS lets that part grow when it is
    // not projected out
    if (o.precision != SISI_FP64 && !e.cplx && e.formulation == SISI_FORMULATION_VERTEX)
      symmetrize_doubles(e, e.T2.get(), e.Nv, e.No);
    double E, dir, exc, Ei;
    ccsd_energy(e, e.T1.get(), e.T2.get(), E, dir, exc, &Ei);
    SISI_REQUIRE(std::isfinite(E) && std::isfinite(e.last_delta), SISI_ERR_NUMERIC,
                 "CCSD iteration " + std::to_string(e.iteration + 1) +
                     " diverged (non-finite energy or amplitudes)");
    e.last_energy = E;
    e.last_energy_imag = Ei;
    e.last_direct = dir;
    e.last_exchange = exc;
    e.energies.push_back(E);
    if (prec == SISI_FP64) e.fp64_iterations++;
    e.iteration++;
    // hand over from the coarse stage, and to FP64 once the bulk iterations
    // reached their threshold
    if (e.coarse_active && e.last_delta <= o.coarse_threshold && e.iteration > 1)
      e.coarse_active = false;
    if (prec != SISI_FP64 && !e.coarse_active && e.last_delta <= o.switch_threshold &&
        e.iteration > 1)
      e.cur_precision = SISI_FP64;
    // convergence test (ClusterSinglesDoublesAlgorithm.cxx:103-106), declared on
    // FP64 residuals only; a refinement iteration that meets it asks for an
    // FP64 anchor next, which then confirms it
    // |(e - previousE) / e| of the complex energies in complex mode
    const double de = std::hypot(E - e.prev_energy, Ei - e.prev_energy_imag) / std::hypot(E, Ei);
    const bool met = de < std::fabs(o.energy_convergence) &&
                     e.last_delta * e.last_delta <
                         std::fabs(o.amplitudes_convergence * o.amplitudes_convergence);
    // A confirmation that fails (DIIS keeps the energy oscillating near the
    // threshold) holds further requests for a few refinement iterations, so
    // that the run does not spend an FP64 anchor on every other iteration.
    constexpr int kConfirmHold = 4;
    if (met && prec == SISI_FP64) e.converged = true;
    else if (refine && prec == SISI_FP64 && e.confirm_pending) e.confirm_hold = kConfirmHold;
    if (prec == SISI_FP64) e.confirm_pending = false;
    if (refine && prec != SISI_FP64) {
      if (met && e.confirm_hold == 0) {
        e.anchor.age = o.refine_interval;
        e.confirm_pending = true;
      } else if (e.confirm_hold > 0) {
        --e.confirm_hold;
      }
    }
    e.prev_energy = E;
    e.prev_energy_imag = Ei;
  }
}

void fill_report(const Engine &e, sisi_ccsd_report *r) {
  r->energy = e.last_energy;
  r->direct = e.last_direct;
  r->exchange = e.last_exchange;
  r->delta = e.last_delta;
  r->iterations = e.iteration;
  r->converged = e.converged ? 1 : 0;
  r->fp64_iterations = e.fp64_iterations;
  r->precision = e.coarse_active ? e.opts.coarse_precision : e.cur_precision;
  r->status = SISI_OK;
  r->energy_imag = e.last_energy_imag;
}

}  // namespace sisi
