// Complex CCSD on the device: integrals from a complex vertex and the
// complex residual of CcsdEnergyFromCoulombIntegrals (cxx:496-891).
//
// Complex tensors are planar: the real part, then the imaginary part of the
// same logical tensor, each dense and first index fastest.  A complex
// product is four real MFMA contractions of the planes,
//   Re C = Re A Re B - sA sB Im A Im B,   Im C = sB Re A Im B + sA Im A Re B,
// with sA = -1 when A enters conjugated (likewise sB), so the reference's
// conjugate-transposed vertices (conjTransposeGammaGai[G,a,i] =
// conj(Gamma_ia) ...) are never materialised: the conjugation becomes a sign
// and the transposition an index order of the planner.  All coefficients of
// the path are real.
#include <cmath>

#include "ccsd.hpp"
#include "contract.hpp"
#include "tensor_ops.hpp"
#include "ztensor.hpp"

namespace sisi {

namespace {

// X = T2 + f T1 T1 (complex outer product; X may alias T2).
void zadd_t1t1(Engine &e, double *X, const double *T2, const double *T1, double f) {
  const int v = e.Nv, o = e.No;
  const int64_t n1 = int64_t(e.n1e()), n2 = int64_t(e.n2e());
  const double *T1r = T1, *T1i = T1 + n1;
  add_t1t1(e, X, T2, T1r, f, v, o);
  add_t1t1(e, X, X, T1i, -f, v, o);
  add_t1t1(e, X + n2, T2 + n2, T1r, f, v, o, T1i);
  add_t1t1(e, X + n2, X + n2, T1i, f, v, o, T1r);
}

// Sum of x*y over n complex elements of two planar arrays (no conjugation):
// out[0] = Re, out[1] = Im.
void zsum_product(Engine &e, int64_t n, const double *x, const double *y, double out[2]) {
  out[0] = dot(e, n, x, y) - dot(e, n, x + n, y + n);
  out[1] = dot(e, n, x, y + n) + dot(e, n, x + n, y);
}

}  // namespace

void zintegrals_from_vertex(Engine &e) {
  e.require_inputs();
  SISI_REQUIRE(e.cplx, SISI_ERR_STATE, "complex integrals need a complex vertex");
  const int64_t o = e.No, v = e.Nv, G = e.NG, P = e.Np;
  // slices (CoulombIntegralsFromVertex.cxx:120-136)
  ZV gam = Z3(e.gamma.get(), G, P, P);
  e.Gij = DBuf(&e.pool, 2 * G * o * o);
  e.Gia = DBuf(&e.pool, 2 * G * o * v);
  e.Gai = DBuf(&e.pool, 2 * G * v * o);
  e.Gab = DBuf(&e.pool, 2 * G * v * v);
  ZV Gij = Z3(e.Gij.get(), G, o, o), Gia = Z3(e.Gia.get(), G, o, v),
     Gai = Z3(e.Gai.get(), G, v, o), Gab = Z3(e.Gab.get(), G, v, v);
  zcopy(e, Gij, zsub(gam, {0, 0, 0}, {G, o, o}));
  zcopy(e, Gia, zsub(gam, {0, 0, o}, {G, o, P}));
  zcopy(e, Gai, zsub(gam, {0, o, 0}, {G, P, o}));
  zcopy(e, Gab, zsub(gam, {0, o, o}, {G, P, P}));
  e.Vabij = DBuf(&e.pool, 2 * v * v * o * o);
  e.Vaijb = DBuf(&e.pool, 2 * v * o * o * v);
  e.Vijab = DBuf(&e.pool, 2 * o * o * v * v);
  e.Vaibj = DBuf(&e.pool, 2 * v * o * v * o);
  e.Vijkl = DBuf(&e.pool, 2 * o * o * o * o);
  e.Vijka = DBuf(&e.pool, 2 * o * o * o * v);
  e.Vaijk = DBuf(&e.pool, 2 * v * o * o * o);
  const int p = SISI_FP64;
  // V = sum_G conj(Gamma_G,rp) Gamma_G,sq with the conjugate transposes of
  // cxx:744-760 written as conjugated operands
  zctr(e, 1.0, "Gia,Gbj->abij", Gia, true, Gai, false, 0.0,
       Z4(e.Vabij.get(), v, v, o, o), p);  // :765
  zctr(e, 1.0, "Gja,Gib->aijb", Gia, true, Gia, false, 0.0,
       Z4(e.Vaijb.get(), v, o, o, v), p);  // :772
  zctr(e, 1.0, "Gai,Gjb->ijab", Gai, true, Gia, false, 0.0,
       Z4(e.Vijab.get(), o, o, v, v), p);  // :779
  zctr(e, 1.0, "Gba,Gij->aibj", Gab, true, Gij, false, 0.0,
       Z4(e.Vaibj.get(), v, o, v, o), p);  // :786
  zctr(e, 1.0, "Gki,Gjl->ijkl", Gij, true, Gij, false, 0.0,
       Z4(e.Vijkl.get(), o, o, o, o), p);  // :793
  zctr(e, 1.0, "Gki,Gja->ijka", Gij, true, Gia, false, 0.0,
       Z4(e.Vijka.get(), o, o, o, v), p);  // :800
  zctr(e, 1.0, "Gja,Gik->aijk", Gia, true, Gij, false, 0.0,
       Z4(e.Vaijk.get(), v, o, o, o), p);  // :807
  // U_klcd = 2 V_klcd - V_kldc of HHPP: the pairs (2 Vijab[klcd] - Vijab[kldc])
  // of CcsdEnergyFromCoulombIntegrals.cxx:602-603, 624-625, 871-872
  ZV Vijab = Z4(e.Vijab.get(), o, o, v, v);
  e.Uklcd = DBuf(&e.pool, 2 * o * o * v * v);
  ZV U = Z4(e.Uklcd.get(), o, o, v, v);
  zcopy(e, U, Vijab, false, 2.0);
  zsum(e, -1.0, Vijab, "kldc", false, 1.0, U, "klcd");
  // getEnergy's V_ijab and V_ijba in [a,b,i,j] order
  // (ClusterSinglesDoublesAlgorithm.cxx:170-178)
  e.VijabT = DBuf(&e.pool, 2 * v * v * o * o);
  e.VjiabT = DBuf(&e.pool, 2 * v * v * o * o);
  zsum(e, 1.0, Vijab, "ijab", false, 0.0, Z4(e.VijabT.get(), v, v, o, o), "abij");
  zsum(e, 1.0, Vijab, "ijba", false, 0.0, Z4(e.VjiabT.get(), v, v, o, o), "abij");
  e.Vbaij.reset();
  e.Uabij.reset();
  e.drop_triples_cache();
  e.drop_residual_graph();
  e.have_integrals = true;
  e.sync();
}

void zccsd_residual(Engine &e, int it, bool initial_doubles, int slice_size, const double *T1c,
                    const double *T2c, double *R1p, double *R2p) {
  SISI_REQUIRE(e.have_integrals && e.cplx, SISI_ERR_STATE, "residual: integrals not built");
  const int64_t o = e.No, v = e.Nv, G = e.NG;
  const int64_t n2 = v * v * o * o;
  double *T1p = const_cast<double *>(T1c), *T2p = const_cast<double *>(T2c);
  ZV T1 = Z2(T1p, v, o), T2 = Z4(T2p, v, v, o, o);
  ZV R1 = Z2(R1p, v, o), R2 = Z4(R2p, v, v, o, o);
  ZV Vabij = Z4(e.Vabij.get(), v, v, o, o), Vaijb = Z4(e.Vaijb.get(), v, o, o, v),
     Vijab = Z4(e.Vijab.get(), o, o, v, v), Vaibj = Z4(e.Vaibj.get(), v, o, v, o),
     Vijkl = Z4(e.Vijkl.get(), o, o, o, o), Vijka = Z4(e.Vijka.get(), o, o, o, v),
     Vaijk = Z4(e.Vaijk.get(), v, o, o, o), U = Z4(e.Uklcd.get(), o, o, v, v);
  ZV Gij = Z3(e.Gij.get(), G, o, o), Gia = Z3(e.Gia.get(), G, o, v),
     Gai = Z3(e.Gai.get(), G, v, o), Gab = Z3(e.Gab.get(), G, v, v);
  zero(e, R1p, int64_t(e.n1()));
  if (it == 0 && !initial_doubles) {  // MP2 start (cxx:516-521)
    zcopy(e, R2, Vabij);
    return;
  }
  zero(e, R2p, 2 * n2);
  const int prec = e.precision;
  e.thaw();
  for (const DBuf *b : {&e.Vabij, &e.Vaijb, &e.Vijab, &e.Vaibj, &e.Vijka, &e.Vaijk, &e.Uklcd,
                        &e.Gij, &e.Gia, &e.Gai, &e.Gab}) {
    e.frozen.push_back(b->get());
    e.frozen.push_back(b->get() + b->size() / 2);
  }
  e.frozen.push_back(T1p);
  e.frozen.push_back(T1p + v * o);
  e.frozen.push_back(T2p);
  e.frozen.push_back(T2p + n2);

  DBuf Xb(&e.pool, 2 * n2);
  ZV X = Z4(Xb.get(), v, v, o, o);
  zadd_t1t1(e, Xb.get(), T2p, T1p, 1.0);  // cxx:584-586
  e.frozen.push_back(Xb.get());
  e.frozen.push_back(Xb.get() + n2);
  DBuf Yb;  // Y = T2 + 2 T1 T1 (cxx:611-614, 674-677)
  auto make_Y = [&]() {
    if (!Yb.get()) {
      Yb = DBuf(&e.pool, 2 * n2);
      zadd_t1t1(e, Yb.get(), T2p, T1p, 2.0);
    }
    return Z4(Yb.get(), v, v, o, o);
  };

  DBuf Kacb(&e.pool, 2 * v * v), Lacb(&e.pool, 2 * v * v), Kkib(&e.pool, 2 * o * o),
      Lkib(&e.pool, 2 * o * o);
  ZV Kac = Z2(Kacb.get(), v, v), Lac = Z2(Lacb.get(), v, v), Kki = Z2(Kkib.get(), o, o),
     Lki = Z2(Lkib.get(), o, o);
  {
    // Kac = -U[klcd] X[adkl] (cxx:602-603); Lac (cxx:606-619)
    zctr(e, -1.0, "klcd,adkl->ac", U, false, X, false, 0.0, Kac, prec);
    if (e.distinguishable) zctr(e, -0.5, "klcd,adkl->ac", U, false, make_Y(), false, 0.0, Lac, prec);
    else zcopy(e, Lac, Kac);
    // + 2 conj(Gamma_Gca) Gamma_Gkd T_dk (cxx:620-621)
    DBuf Zb(&e.pool, 2 * G);
    zctr(e, 1.0, "Gkd,dk->G", Gia, false, T1, false, 0.0, Z1(Zb.get(), G), prec);
    zctr(e, 2.0, "Gca,G->ac", Gab, true, Z1(Zb.get(), G), false, 1.0, Lac, prec);
    // - conj(Gamma_Gda) Gamma_Gkc T_dk (cxx:622-623)
    DBuf Ygb(&e.pool, 2 * G * v * o);
    zctr(e, 1.0, "Gda,dk->Gak", Gab, true, T1, false, 0.0, Z3(Ygb.get(), G, v, o), prec);
    zctr(e, -1.0, "Gak,Gkc->ac", Z3(Ygb.get(), G, v, o), false, Gia, false, 1.0, Lac, prec);
    // Kki = U[klcd] X[cdil] (cxx:624-625); Lki (cxx:628-641)
    zctr(e, 1.0, "klcd,cdil->ki", U, false, X, false, 0.0, Kki, prec);
    if (e.distinguishable) zctr(e, 0.5, "klcd,cdil->ki", U, false, make_Y(), false, 0.0, Lki, prec);
    else zcopy(e, Lki, Kki);
    zctr(e, 2.0, "klic,cl->ki", Vijka, false, T1, false, 1.0, Lki, prec);   // :640
    zctr(e, -1.0, "lkic,cl->ki", Vijka, false, T1, false, 1.0, Lki, prec);  // :641
    zctr(e, 1.0, "ac,cbij->abij", Lac, false, T2, false, 1.0, R2, prec);   // :644
    zctr(e, -1.0, "ki,abkj->abij", Lki, false, T2, false, 1.0, R2, prec);  // :647
  }
  {
    // dressed conj-transposed Gamma_ai (cxx:650-660)
    DBuf dGb(&e.pool, 2 * G * v * o), Qb(&e.pool, 2 * G * v * o), Pb(&e.pool, 2 * v * o * o * o);
    ZV dG = Z3(dGb.get(), G, v, o), Q = Z3(Qb.get(), G, v, o), P = Z4(Pb.get(), v, o, o, o);
    zsum(e, 1.0, Gia, "Gia", true, 0.0, dG, "Gai");
    zctr(e, -1.0, "Gik,ak->Gai", Gij, true, T1, false, 1.0, dG, prec);   // :653
    zctr(e, 1.0, "Gbc,cj->Gbj", Gab, false, T1, false, 0.0, Q, prec);
    zctr(e, 1.0, "Gai,Gbj->abij", dG, false, Q, false, 1.0, R2, prec);   // :654-655
    zctr(e, -1.0, "akij,bk->abij", Vaijk, false, T1, false, 1.0, R2, prec);  // :657
    zctr(e, 1.0, "akic,cj->akij", Vaijb, false, T1, false, 0.0, P, prec);    // :658-659
    zctr(e, -1.0, "bk,akij->abij", T1, false, P, false, 1.0, R2, prec);
  }
  {
    // Xakic (cxx:662-694)
    DBuf dGb(&e.pool, 2 * G * v * o), Xkb(&e.pool, 2 * v * o * o * v);
    ZV dG = Z3(dGb.get(), G, v, o), Xakic = Z4(Xkb.get(), v, o, o, v);
    zsum(e, 1.0, Gia, "Gia", true, 0.0, dG, "Gai");
    zctr(e, -1.0, "Gil,al->Gai", Gij, true, T1, false, 1.0, dG, prec);           // :670-671
    zctr(e, 1.0, "Gda,di->Gai", Gab, true, T1, false, 1.0, dG, prec);            // :673-674
    zctr(e, -0.5, "Gdl,dail->Gai", Gai, true, make_Y(), false, 1.0, dG, prec);   // :681-682
    zctr(e, 1.0, "Gai,Gkc->akic", dG, false, Gia, false, 0.0, Xakic, prec);      // :684-685
    zctr(e, 1.0, "lkdc,adil->akic", Vijab, false, T2, false, 1.0, Xakic, prec);  // :687, 689
    if (!e.distinguishable)
      zctr(e, -0.5, "lkcd,adil->akic", Vijab, false, T2, false, 1.0, Xakic, prec);  // :688
    // 2 T[cbkj] - T[bckj] (cxx:692-693)
    DBuf Y2b(&e.pool, 2 * n2);
    ZV Y2 = Z4(Y2b.get(), v, v, o, o);
    axpby(e, 2 * n2, 2.0, T2p, 0.0, Y2b.get());
    zsum(e, -1.0, T2, "bckj", false, 1.0, Y2, "cbkj");
    zctr(e, 1.0, "akic,cbkj->abij", Xakic, false, Y2, false, 1.0, R2, prec);  // :695
  }
  {
    // Xakci (cxx:698-728)
    DBuf dGabb(&e.pool, 2 * G * v * v), dGijb(&e.pool, 2 * G * o * o),
        Xcb(&e.pool, 2 * v * o * v * o);
    ZV dGab = Z3(dGabb.get(), G, v, v), dGij = Z3(dGijb.get(), G, o, o),
       Xakci = Z4(Xcb.get(), v, o, v, o);
    zsum(e, 1.0, Gab, "Gca", true, 0.0, dGab, "Gac");
    zctr(e, -1.0, "Gcl,al->Gac", Gai, true, T1, false, 1.0, dGab, prec);  // :708-709
    zcopy(e, dGij, Gij);
    zctr(e, 1.0, "Gkd,di->Gki", Gia, false, T1, false, 1.0, dGij, prec);  // :711
    zctr(e, 1.0, "Gac,Gki->akci", dGab, false, dGij, false, 0.0, Xakci, prec);  // :714-715
    if (!e.distinguishable)
      zctr(e, -0.5, "lkcd,dail->akci", Vijab, false, T2, false, 1.0, Xakci, prec);  // :719
    zctr(e, -1.0, "akci,cbkj->abij", Xakci, false, T2, false, 1.0, R2, prec);  // :723
    zctr(e, -1.0, "bkci,ackj->abij", Xakci, false, T2, false, 1.0, R2, prec);  // :724
  }
  {
    // R_abij += R_baji (cxx:727)
    DBuf Rt(&e.pool, 2 * n2);
    ZV Rtv = Z4(Rt.get(), v, v, o, o);
    zcopy(e, Rtv, R2);
    zsum(e, 1.0, Rtv, "baji", false, 1.0, R2, "abij");
  }
  axpby(e, 2 * n2, 1.0, e.Vabij.get(), 1.0, R2p);  // cxx:736
  {
    // Xklij (cxx:738-768)
    DBuf Xb4(&e.pool, 2 * o * o * o * o);
    ZV Xklij = Z4(Xb4.get(), o, o, o, o);
    zcopy(e, Xklij, Vijkl);
    zctr(e, 1.0, "klic,cj->klij", Vijka, false, T1, false, 1.0, Xklij, prec);  // :744
    zctr(e, 1.0, "lkjc,ci->klij", Vijka, false, T1, false, 1.0, Xklij, prec);  // :745
    zctr(e, 1.0, "klij,abkl->abij", Xklij, false, X, false, 1.0, R2, prec);    // :748
    if (e.distinguishable) {
      // T[ci] V[klcd] T[dj] (cxx:753): HHPP contracted with X - T2
      DBuf Wb(&e.pool, 2 * n2);
      axpby(e, 2 * n2, 1.0, Xb.get(), 0.0, Wb.get());
      axpby(e, 2 * n2, -1.0, T2p, 1.0, Wb.get());
      zctr(e, 1.0, "klcd,cdij->klij", Vijab, false, Z4(Wb.get(), v, v, o, o), false, 0.0, Xklij,
           prec);
    } else {
      zctr(e, 1.0, "klcd,cdij->klij", Vijab, false, X, false, 0.0, Xklij, prec);  // :751
    }
    zctr(e, 1.0, "klij,abkl->abij", Xklij, false, e.ppl ? T2 : X, false, 1.0, R2, prec);  // :763-767
  }
  if (e.ppl && e.NR > 0) {
    // factor-based ladder for complex amplitudes (cxx:786-813)
    zfactor_ladder(e, T1p, Xb.get(), R2p, e.prec_ladder);
  } else if (e.ppl) {
    // particle-particle ladder (cxx:814-858) with the dressed vertices of
    // sliceCoupledCoulombIntegrals (ClusterSinglesDoublesAlgorithm.cxx:531-557):
    // right = Gamma_Gab - Gamma_Gkb T_ak, left = conj(Gamma_Gba) - conj(Gamma_Gbk) T_ak
    if (slice_size <= 0) slice_size = int(o);
    DBuf Lb(&e.pool, 2 * G * v * v), Rb(&e.pool, 2 * G * v * v);
    ZV left = Z3(Lb.get(), G, v, v), right = Z3(Rb.get(), G, v, v);
    zcopy(e, right, Gab);
    zctr(e, -1.0, "Gkb,ak->Gab", Gia, false, T1, false, 1.0, right, prec);
    zsum(e, 1.0, Gab, "Gba", true, 0.0, left, "Gab");
    zctr(e, -1.0, "Gbk,ak->Gab", Gai, true, T1, false, 1.0, left, prec);
    const int64_t s = std::min<int64_t>(slice_size, v);
    DBuf Ls(&e.pool, 2 * G * s * v), Rs(&e.pool, 2 * G * s * v), Vxy(&e.pool, 2 * s * s * v * v),
        Rxy(&e.pool, 2 * s * s * o * o);
    for (int64_t b = 0; b < v; b += slice_size)
      for (int64_t a = b; a < v; a += slice_size) {  // cxx:836-858
        const int64_t Nx = std::min<int64_t>(a + slice_size, v) - a;
        const int64_t Ny = std::min<int64_t>(b + slice_size, v) - b;
        ZV Lx = Z3(Ls.get(), G, Nx, v), Ry = Z3(Rs.get(), G, Ny, v);
        zcopy(e, Lx, zsub(left, {0, a, 0}, {G, a + Nx, v}));
        zcopy(e, Ry, zsub(right, {0, b, 0}, {G, b + Ny, v}));
        ZV Vxycd = Z4(Vxy.get(), Nx, Ny, v, v), Rxyij = Z4(Rxy.get(), Nx, Ny, o, o);
        zctr(e, 1.0, "Gxc,Gyd->xycd", Lx, false, Ry, false, 0.0, Vxycd, e.prec_ladder);
        zctr(e, 1.0, "xycd,cdij->xyij", Vxycd, false, X, false, 0.0, Rxyij, e.prec_ladder);
        // sliceIntoResiduum (ClusterSinglesDoublesAlgorithm.cxx:887-915)
        zcopy(e, zsub(R2, {a, b, 0, 0}, {a + Nx, b + Ny, o, o}), Rxyij, false, 1.0, 1.0);
        if (a > b) zsum(e, 1.0, Rxyij, "xyij", false, 1.0,
                        zsub(R2, {b, a, 0, 0}, {b + Ny, a + Nx, o, o}), "yxji");
      }
  }
  {
    // T1 equations (cxx:861-891)
    zctr(e, 1.0, "ac,ci->ai", Kac, false, T1, false, 1.0, R1, prec);    // :869
    zctr(e, -1.0, "ki,ak->ai", Kki, false, T1, false, 1.0, R1, prec);   // :870
    DBuf Kckb(&e.pool, 2 * v * o), Sb(&e.pool, 2 * o * o), Wb(&e.pool, 2 * G * v * o);
    ZV Kck = Z2(Kckb.get(), v, o), S = Z2(Sb.get(), o, o), W = Z3(Wb.get(), G, v, o);
    zctr(e, 1.0, "klcd,dl->ck", U, false, T1, false, 0.0, Kck, prec);  // :873-874
    zctr(e, 2.0, "ck,caki->ai", Kck, false, T2, false, 1.0, R1, prec);   // :877
    zctr(e, -1.0, "ck,caik->ai", Kck, false, T2, false, 1.0, R1, prec);  // :878
    zctr(e, 1.0, "ck,ci->ki", Kck, false, T1, false, 0.0, S, prec);      // :879
    zctr(e, 1.0, "ak,ki->ai", T1, false, S, false, 1.0, R1, prec);
    zctr(e, 2.0, "akic,ck->ai", Vaijb, false, T1, false, 1.0, R1, prec);   // :880
    zctr(e, -1.0, "akci,ck->ai", Vaibj, false, T1, false, 1.0, R1, prec);  // :881
    zctr(e, 1.0, "Gkd,cdik->Gci", Gia, false, X, false, 0.0, W, prec);     // :883-884
    zctr(e, 2.0, "Gca,Gci->ai", Gab, true, W, false, 1.0, R1, prec);
    zctr(e, 1.0, "Gkc,cdik->Gdi", Gia, false, X, false, 0.0, W, prec);     // :885-886
    zctr(e, -1.0, "Gda,Gdi->ai", Gab, true, W, false, 1.0, R1, prec);
    zctr(e, -2.0, "klic,ackl->ai", Vijka, false, X, false, 1.0, R1, prec);  // :888
    zctr(e, 1.0, "lkic,ackl->ai", Vijka, false, X, false, 1.0, R1, prec);   // :889
  }
  e.thaw();
}

void zccsd_energy(Engine &e, const double *T1, const double *T2, double E[2], double direct[2],
                  double exchange[2]) {
  const int64_t n2 = int64_t(e.n2e());
  DBuf tau(&e.pool, 2 * n2);
  zadd_t1t1(e, tau.get(), T2, T1, 1.0);
  // direct = 2 sum tau_abij V_ijab, exchange = -sum tau_abij V_ijba
  // (ClusterSinglesDoublesAlgorithm.cxx:170-178 with spins = 2)
  double d[2], x[2];
  zsum_product(e, n2, tau.get(), e.VijabT.get(), d);
  zsum_product(e, n2, tau.get(), e.VjiabT.get(), x);
  for (int k = 0; k < 2; ++k) {
    direct[k] = 2.0 * d[k];
    exchange[k] = -x[k];
    E[k] = direct[k] + exchange[k];
  }
}

void zccsd_pair_energies(Engine &e, const double *T1c, const double *T2c, double *pair) {
  const int64_t o = e.No, v = e.Nv, n2 = int64_t(e.n2e());
  DBuf tau(&e.pool, 2 * n2), W(&e.pool, 2 * n2);
  zadd_t1t1(e, tau.get(), T2c, T1c, 1.0);
  // e_ij = sum_ab tau_abij (2 V_ijab - V_ijba) (cxx:227-232)
  axpby(e, 2 * n2, 2.0, e.VijabT.get(), 0.0, W.get());
  axpby(e, 2 * n2, -1.0, e.VjiabT.get(), 1.0, W.get());
  zctr(e, 1.0, "abij,abij->ij", Z4(tau.get(), v, v, o, o), false, Z4(W.get(), v, v, o, o), false,
       0.0, Z2(pair, o, o), SISI_FP64);
}

}  // namespace sisi
