// Full-integral formulation of the CCSD residual:
// CcsdEnergyFromCoulombIntegralsReference::getResiduum (cxx:29-295), the
// Hirata form with stored PPPP V_abcd and PPPH V_abci and no vertex.  Every
// three-operand expression of the reference is evaluated as two binary MFMA
// contractions; pairs that act on the same operand are merged through
// X = T2 + T1 T1, Y = T2 + 2 T1 T1 and U = 2 V_abij - V_baij.
#include <cmath>

#include "ccsd.hpp"
#include "contract.hpp"
#include "tensor_ops.hpp"

namespace sisi {

namespace {
View D2(double *p, int64_t a, int64_t b) { return dense(p, {a, b}); }
View D4(double *p, int64_t a, int64_t b, int64_t c, int64_t d) { return dense(p, {a, b, c, d}); }
}  // namespace

void set_integrals(Engine &e, int which, const double *V) {
  SISI_REQUIRE(V != nullptr, SISI_ERR_INVALID, "null integrals");
  SISI_REQUIRE(!e.cplx, SISI_ERR_INVALID, "integrals are uploaded for the real path only");
  const int64_t o = e.No, v = e.Nv;
  DBuf *dst = nullptr;
  int64_t n = 0;
  switch (which) {
    case SISI_PPHH: dst = &e.Vabij; n = v * v * o * o; break;
    case SISI_PHPH: dst = &e.Vaibj; n = v * o * v * o; break;
    case SISI_HHHH: dst = &e.Vijkl; n = o * o * o * o; break;
    case SISI_HHHP: dst = &e.Vijka; n = o * o * o * v; break;
    case SISI_PPPP: dst = &e.Vabcd; n = v * v * v * v; break;
    case SISI_PPPH: dst = &e.Vabci; n = v * v * v * o; break;
    case SISI_PHHP: dst = &e.ring_Vaijb; n = v * o * o * v; break;
    case SISI_HHPP: dst = &e.ring_Vijab; n = o * o * v * v; break;
    default: throw Error(SISI_ERR_INVALID, "unknown integral class");
  }
  for (int64_t k = 0; k < n; ++k)
    SISI_REQUIRE(std::isfinite(V[k]), SISI_ERR_NUMERIC, "non-finite integrals");
  *dst = DBuf(&e.pool, size_t(n));
  e.upload(*dst, V, size_t(n));
  e.given_integrals |= 1u << which;
  if (which == SISI_PPHH) {
    // constants of the energy and of the K/L intermediates
    View Vv = D4(e.Vabij.get(), v, v, o, o);
    e.Vbaij = DBuf(&e.pool, size_t(n));
    e.Uabij = DBuf(&e.pool, size_t(n));
    sum(e, 1.0, Vv, "baij", 0.0, D4(e.Vbaij.get(), v, v, o, o), "abij");
    axpby(e, n, 2.0, e.Vabij.get(), 0.0, e.Uabij.get());
    axpby(e, n, -1.0, e.Vbaij.get(), 1.0, e.Uabij.get());
  }
  const unsigned core = (1u << SISI_PPHH) | (1u << SISI_PHPH) | (1u << SISI_HHHH) | (1u << SISI_HHHP);
  if ((e.given_integrals & core) == core) e.have_integrals = true;
  e.drop_triples_cache();
  e.drop_residual_graph();
  e.sync();
}

void build_vabcd(Engine &e) {
  SISI_REQUIRE(e.Gab.get() && !e.cplx, SISI_ERR_STATE, "PPPP/PPPH need the real vertex");
  const int64_t o = e.No, v = e.Nv, G = e.NG;
  // CoulombIntegralsFromVertex.cxx:423, 430
  if (!e.Vabcd.get()) {
    e.Vabcd = DBuf(&e.pool, size_t(v * v * v * v));
    ctr(e, 1.0, "Gac,Gbd->abcd", dense(e.Gab.get(), {G, v, v}), dense(e.Gab.get(), {G, v, v}), 0.0,
        D4(e.Vabcd.get(), v, v, v, v), SISI_FP64);
  }
  if (!e.Vabci.get()) {
    e.Vabci = DBuf(&e.pool, size_t(v * v * v * o));
    ctr(e, 1.0, "Gac,Gbi->abci", dense(e.Gab.get(), {G, v, v}), dense(e.Gai.get(), {G, v, o}), 0.0,
        D4(e.Vabci.get(), v, v, v, o), SISI_FP64);
  }
}

void ccsd_residual_full(Engine &e, int it, bool initial_doubles, const double *T1c,
                        const double *T2c, double *R1p, double *R2p) {
  SISI_REQUIRE(e.have_integrals && e.Vabcd.get(), SISI_ERR_STATE,
               "full-integral residual: PPHH, PHPH, HHHH, HHHP and PPPP integrals needed");
  const int64_t o = e.No, v = e.Nv, n2 = v * v * o * o;
  double *T1p = const_cast<double *>(T1c), *T2p = const_cast<double *>(T2c);
  View T1 = D2(T1p, v, o), T2 = D4(T2p, v, v, o, o), R1 = D2(R1p, v, o),
       R2 = D4(R2p, v, v, o, o);
  View V = D4(e.Vabij.get(), v, v, o, o), U = D4(e.Uabij.get(), v, v, o, o),
       Vaibj = D4(e.Vaibj.get(), v, o, v, o), Vijkl = D4(e.Vijkl.get(), o, o, o, o),
       Vijka = D4(e.Vijka.get(), o, o, o, v), Vabcd = D4(e.Vabcd.get(), v, v, v, v);
  const int prec = e.precision;
  zero(e, R1p, v * o);
  zero(e, R2p, n2);
  DBuf Xb(&e.pool, n2);
  View X = D4(Xb.get(), v, v, o, o);
  add_t1t1(e, Xb.get(), T2p, T1p, 1.0, int(v), int(o));
  if (e.formulation == SISI_FORMULATION_ONLY_PPL) {  // cxx:56-73
    ctr(e, 1.0, "abcd,cdij->abij", Vabcd, X, 0.0, R2, e.prec_ladder);
    return;
  }
  if (it == 0 && !initial_doubles) {  // MP2 start (cxx:50-54)
    copy(e, R2, V);
    return;
  }
  SISI_REQUIRE(e.Vabci.get(), SISI_ERR_STATE, "full-integral residual: PPPH integrals needed");
  View Vabci = D4(e.Vabci.get(), v, v, v, o);
  e.thaw();
  e.frozen = {T1p, T2p, Xb.get(), e.Vabij.get(), e.Uabij.get(), e.Vaibj.get(), e.Vijka.get(),
              e.Vabcd.get(), e.Vabci.get()};
  DBuf Yb(&e.pool, n2);
  View Y = D4(Yb.get(), v, v, o, o);
  add_t1t1(e, Yb.get(), T2p, T1p, 2.0, int(v), int(o));

  DBuf Kacb(&e.pool, v * v), Lacb(&e.pool, v * v), Kkib(&e.pool, o * o), Lkib(&e.pool, o * o);
  View Kac = D2(Kacb.get(), v, v), Lac = D2(Lacb.get(), v, v), Kki = D2(Kkib.get(), o, o),
       Lki = D2(Lkib.get(), o, o);
  // Kac, Lac, Kki, Lki (cxx:171-190)
  ctr(e, -1.0, "cdkl,adkl->ac", U, X, 0.0, Kac, prec);
  copy(e, Lac, Kac);
  ctr(e, 2.0, "cdak,dk->ac", Vabci, T1, 1.0, Lac, prec);
  ctr(e, -1.0, "dcak,dk->ac", Vabci, T1, 1.0, Lac, prec);
  ctr(e, 1.0, "cdkl,cdil->ki", U, X, 0.0, Kki, prec);
  copy(e, Lki, Kki);
  ctr(e, 2.0, "klic,cl->ki", Vijka, T1, 1.0, Lki, prec);
  ctr(e, -1.0, "lkic,cl->ki", Vijka, T1, 1.0, Lki, prec);
  ctr(e, 1.0, "ac,cbij->abij", Lac, T2, 1.0, R2, prec);   // cxx:193
  ctr(e, -1.0, "ki,abkj->abij", Lki, T2, 1.0, R2, prec);  // cxx:196
  {
    // singles terms of cxx:199-202
    ctr(e, 1.0, "baci,cj->abij", Vabci, T1, 1.0, R2, prec);
    DBuf Pb(&e.pool, v * o * o * o);
    View P = D4(Pb.get(), v, o, o, o);
    ctr(e, 1.0, "bkci,cj->bkij", Vaibj, T1, 0.0, P, prec);
    ctr(e, -1.0, "ak,bkij->abij", T1, P, 1.0, R2, prec);
    ctr(e, -1.0, "jika,bk->abij", Vijka, T1, 1.0, R2, prec);
    ctr(e, 1.0, "acik,cj->aikj", V, T1, 0.0, P, prec);
    ctr(e, -1.0, "bk,aikj->abij", T1, P, 1.0, R2, prec);
  }
  {
    // Xakic (cxx:205-211) and its contraction with 2 T[cbkj] - T[bckj] (cxx:221-222)
    DBuf Xkb(&e.pool, v * o * o * v);
    View Xakic = D4(Xkb.get(), v, o, o, v);
    sum(e, 1.0, V, "acik", 0.0, Xakic, "akic");
    ctr(e, -1.0, "lkic,al->akic", Vijka, T1, 1.0, Xakic, prec);
    ctr(e, 1.0, "acdk,di->akic", Vabci, T1, 1.0, Xakic, prec);
    ctr(e, -0.5, "dclk,dail->akic", V, Y, 1.0, Xakic, prec);
    ctr(e, 1.0, "dclk,adil->akic", V, T2, 1.0, Xakic, prec);
    ctr(e, -0.5, "cdlk,adil->akic", V, T2, 1.0, Xakic, prec);
    DBuf Y2b(&e.pool, n2);
    View Y2 = D4(Y2b.get(), v, v, o, o);
    axpby(e, n2, 2.0, T2p, 0.0, Y2b.get());
    sum(e, -1.0, T2, "bckj", 1.0, Y2, "cbkj");
    ctr(e, 1.0, "akic,cbkj->abij", Xakic, Y2, 1.0, R2, prec);
  }
  {
    // Xakci (cxx:214-218, 224-225)
    DBuf Xcb(&e.pool, v * o * v * o);
    View Xakci = D4(Xcb.get(), v, o, v, o);
    copy(e, Xakci, Vaibj);
    ctr(e, -1.0, "klic,al->akci", Vijka, T1, 1.0, Xakci, prec);
    ctr(e, 1.0, "adck,di->akci", Vabci, T1, 1.0, Xakci, prec);
    ctr(e, -0.5, "cdlk,dail->akci", V, Y, 1.0, Xakci, prec);
    ctr(e, -1.0, "akci,cbkj->abij", Xakci, T2, 1.0, R2, prec);
    ctr(e, -1.0, "bkci,ackj->abij", Xakci, T2, 1.0, R2, prec);
  }
  {
    DBuf Rt(&e.pool, n2);  // R_abij += R_baji (cxx:229-230)
    copy(e, D4(Rt.get(), v, v, o, o), R2);
    sum(e, 1.0, D4(Rt.get(), v, v, o, o), "baji", 1.0, R2, "abij");
  }
  axpby(e, n2, 1.0, e.Vabij.get(), 1.0, R2p);  // cxx:239
  {
    // Xklij (cxx:242-246), contracted with T2 + T1 T1 (cxx:249, 252)
    DBuf Xb4(&e.pool, o * o * o * o);
    View Xklij = D4(Xb4.get(), o, o, o, o);
    copy(e, Xklij, Vijkl);
    ctr(e, 1.0, "klic,cj->klij", Vijka, T1, 1.0, Xklij, prec);
    ctr(e, 1.0, "lkjc,ci->klij", Vijka, T1, 1.0, Xklij, prec);
    ctr(e, 1.0, "cdkl,cdij->klij", V, X, 1.0, Xklij, prec);
    ctr(e, 1.0, "klij,abkl->abij", Xklij, X, 1.0, R2, prec);
  }
  {
    // Xabcd (cxx:255-257), contracted with T2 + T1 T1 (cxx:260-261)
    DBuf Xdb(&e.pool, v * v * v * v);
    View Xabcd = D4(Xdb.get(), v, v, v, v);
    copy(e, Xabcd, Vabcd);
    ctr(e, -1.0, "cdak,bk->abcd", Vabci, T1, 1.0, Xabcd, prec);
    ctr(e, -1.0, "dcbk,ak->abcd", Vabci, T1, 1.0, Xabcd, prec);
    ctr(e, 1.0, "abcd,cdij->abij", Xabcd, X, 1.0, R2, e.prec_ladder);
  }
  {
    // T1 equations (cxx:271-292)
    ctr(e, 1.0, "ac,ci->ai", Kac, T1, 1.0, R1, prec);
    ctr(e, -1.0, "ki,ak->ai", Kki, T1, 1.0, R1, prec);
    DBuf Kckb(&e.pool, v * o), Skib(&e.pool, o * o);
    View Kck = D2(Kckb.get(), v, o), Ski = D2(Skib.get(), o, o);
    ctr(e, 2.0, "cdkl,dl->ck", V, T1, 0.0, Kck, prec);   // cxx:275
    ctr(e, -1.0, "cdlk,dl->ck", V, T1, 1.0, Kck, prec);  // cxx:276
    ctr(e, 2.0, "ck,caki->ai", Kck, T2, 1.0, R1, prec);
    ctr(e, -1.0, "ck,caik->ai", Kck, T2, 1.0, R1, prec);
    ctr(e, 1.0, "ck,ci->ki", Kck, T1, 0.0, Ski, prec);
    ctr(e, 1.0, "ak,ki->ai", T1, Ski, 1.0, R1, prec);
    ctr(e, 2.0, "acik,ck->ai", V, T1, 1.0, R1, prec);
    ctr(e, -1.0, "akci,ck->ai", Vaibj, T1, 1.0, R1, prec);
    ctr(e, 2.0, "cdak,cdik->ai", Vabci, X, 1.0, R1, prec);
    ctr(e, -1.0, "dcak,cdik->ai", Vabci, X, 1.0, R1, prec);
    ctr(e, -2.0, "klic,ackl->ai", Vijka, X, 1.0, R1, prec);
    ctr(e, 1.0, "lkic,ackl->ai", Vijka, X, 1.0, R1, prec);
  }
  e.thaw();
}

}  // namespace sisi
