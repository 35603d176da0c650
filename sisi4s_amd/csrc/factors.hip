// Factor-based particle-particle ladder (CoulombFactors / FactorOrbitals):
// sliceAmplitudesFromCoupledCoulombFactors for real amplitudes
// (ClusterSinglesDoublesAlgorithm.cxx:567-672) over all pairs of factor
// slices (CcsdEnergyFromCoulombIntegrals.cxx:340-372).
//
// With Lambda[G,R] and Pi[p,R] (complex, planar) and X = T2 + T1 T1:
//   X1[R,d,i,j]  = sum_c X[c,d,i,j] conj(Pi_cR)            (R in the left slice)
//   XRS[R,S,i,j] = sum_d X1[R,d,i,j] Pi_dS * V_RS,  V_RS = sum_G conj(Lambda_GR) Lambda_GS
//   X2[R,b,i,j]  = sum_S XRS[R,S,i,j] dPi_bS,       dPi_aR = Pi_aR - sum_k Pi_kR T_ak
//   R_abij      += Re sum_R X2[R,b,i,j] conj(dPi_aR)
// (X2 is summed over the right slices before the last product, which is
// linear in it).  It equals the vertex ladder when
// Gamma[G,p,q] = sum_R Lambda_GR Pi_pR Pi_qR, at a cost that scales with NR^2
// instead of Nv^4.  Real factors skip every imaginary-plane product.
#include <cmath>
#include <vector>

#include "ccsd.hpp"
#include "contract.hpp"
#include "tensor_ops.hpp"
#include "ztensor.hpp"

namespace sisi {

void set_coulomb_factors(Engine &e, int NR, int slice, const double *lambda, const double *pi,
                         const double *pi_out) {
  e.drop_residual_graph();
  if (NR <= 0) {
    e.NR = 0;
    e.Lam.reset();
    e.PiF.reset();
    e.PiO.reset();
    e.VRS.reset();
    return;
  }
  SISI_REQUIRE(lambda && pi, SISI_ERR_INVALID, "null CoulombFactors / FactorOrbitals");
  const int64_t G = e.NG, P = e.Np, R = NR;
  bool real = true;
  for (int64_t k = 0; k < G * R && real; ++k) real = lambda[2 * k + 1] == 0.0;
  for (int64_t k = 0; k < P * R && real; ++k) real = pi[2 * k + 1] == 0.0;
  for (int64_t k = 0; k < 2 * G * R; ++k)
    SISI_REQUIRE(std::isfinite(lambda[k]), SISI_ERR_NUMERIC, "non-finite CoulombFactors");
  for (int64_t k = 0; k < 2 * P * R; ++k)
    SISI_REQUIRE(std::isfinite(pi[k]) && (!pi_out || std::isfinite(pi_out[k])), SISI_ERR_NUMERIC,
                 "non-finite FactorOrbitals");
  // interleaved (re, im) -> planar
  auto planar = [](const double *z, int64_t n) {
    std::vector<double> out(size_t(2 * n));
    for (int64_t k = 0; k < n; ++k) {
      out[size_t(k)] = z[2 * k];
      out[size_t(n + k)] = z[2 * k + 1];
    }
    return out;
  };
  e.upload(e.Lam, planar(lambda, G * R).data(), size_t(2 * G * R));
  e.upload(e.PiF, planar(pi, P * R).data(), size_t(2 * P * R));
  e.upload(e.PiO, planar(pi_out ? pi_out : pi, P * R).data(), size_t(2 * P * R));
  e.NR = NR;
  e.factors_slice = slice > 0 ? slice : e.Nv;  // factorsSliceSize default Nv (cxx:345-356)
  e.factors_real = real;
  e.VRS = DBuf(&e.pool, size_t(2 * R * R));
  ZV L = zdense(e.Lam.get(), {G, R});
  zctr(e, 1.0, "GR,GS->RS", L, true, L, false, 0.0, zdense(e.VRS.get(), {R, R}), SISI_FP64);
  e.sync();
}

void factor_ladder(Engine &e, const double *T1p, const double *Xp, double *R2p, int prec) {
  SISI_REQUIRE(e.NR > 0, SISI_ERR_STATE, "no Coulomb factors");
  const int64_t o = e.No, v = e.Nv, P = e.Np, NR = e.NR, fs = std::min<int64_t>(e.factors_slice, NR);
  const bool re_only = e.factors_real;
  View T1 = dense(const_cast<double *>(T1p), {v, o});
  View X = dense(const_cast<double *>(Xp), {v, v, o, o});
  View R2 = dense(R2p, {v, v, o, o});
  ZV Pi = zdense(e.PiF.get(), {P, NR});
  ZV PiA = zsub(Pi, {o, 0}, {P, NR}), PiI = zsub(Pi, {0, 0}, {o, NR});
  // complex product, or the real planes alone for real factors
  auto zc = [&](double alpha, const char *spec, const ZV &A, bool ca, const ZV &B, bool cb,
                double beta, const ZV &C) {
    if (re_only) ctr(e, alpha, spec, A.re, B.re, beta, C.re, prec);
    else zctr(e, alpha, spec, A, ca, B, cb, beta, C, prec);
  };
  // dPi[a,R] = Pi_aR - sum_k Pi_kR T_ak (cxx:639-650)
  DBuf dPib(&e.pool, size_t(2 * v * NR));
  ZV dPi = zdense(dPib.get(), {v, NR});
  zcopy(e, dPi, PiA);
  ctr(e, -1.0, "kR,ak->aR", PiI.re, T1, 1.0, dPi.re, prec);
  if (!re_only) ctr(e, -1.0, "kR,ak->aR", PiI.im, T1, 1.0, dPi.im, prec);
  DBuf X1b(&e.pool, size_t(2 * fs * v * o * o)), X2b(&e.pool, size_t(2 * fs * v * o * o)),
      XSb(&e.pool, size_t(2 * fs * fs * o * o));
  for (int64_t R0 = 0; R0 < NR; R0 += fs) {
    const int64_t Rx = std::min(fs, NR - R0);
    ZV X1 = zdense(X1b.get(), {Rx, v, o, o}), X2 = zdense(X2b.get(), {Rx, v, o, o});
    ZV PL = zsub(PiA, {0, R0}, {v, R0 + Rx}), dPL = zsub(dPi, {0, R0}, {v, R0 + Rx});
    // X1 = X conj(Pi_L): X real (cxx:607-610)
    ctr(e, 1.0, "cdij,cR->Rdij", X, PL.re, 0.0, X1.re, prec);
    if (!re_only) ctr(e, -1.0, "cdij,cR->Rdij", X, PL.im, 0.0, X1.im, prec);
    for (int64_t S0 = 0; S0 < NR; S0 += fs) {
      const int64_t Ry = std::min(fs, NR - S0);
      ZV XS = zdense(XSb.get(), {Rx, Ry, o, o});
      ZV PS = zsub(PiA, {0, S0}, {v, S0 + Ry}), dPS = zsub(dPi, {0, S0}, {v, S0 + Ry});
      zc(1.0, "Rdij,dS->RSij", X1, false, PS, false, 0.0, XS);                  // cxx:613-614
      scale_pairs(e, XSb.get(), Rx, Ry, o * o, e.VRS.get(), NR, NR * NR, R0, S0, re_only);  // :623
      zc(1.0, "RSij,bS->Rbij", XS, false, dPS, false, S0 == 0 ? 0.0 : 1.0, X2);  // cxx:660
    }
    // R += Re sum_R X2[R,b,i,j] conj(dPi_aR) (cxx:666-668)
    ctr(e, 1.0, "Rbij,aR->abij", X2.re, dPL.re, 1.0, R2, prec);
    if (!re_only) ctr(e, 1.0, "Rbij,aR->abij", X2.im, dPL.im, 1.0, R2, prec);
  }
}

// Complex amplitudes (ClusterSinglesDoublesAlgorithm.cxx:745-880), T1, X and
// R2 planar complex, with Pi = FactorOrbitals and Po = OutgoingFactorOrbitals:
//   X1[R,d,i,j]  = sum_c X[c,d,i,j] conj(Po_cR)
//   XRS[R,S,i,j] = sum_d X1[R,d,i,j] Pi_dS * V_RS
//   X2[R,b,i,j]  = sum_S XRS[R,S,i,j] D_bS,   D_bS = Po_bS - sum_k Po_kS T_bk
//   R_abij      += sum_R X2[R,b,i,j] C_aR,    C_aR = conj(Pi_aR) - sum_k conj(Pi_kR) T_ak
// It equals the complex vertex ladder when Gamma[G,q,r] = sum_R Lambda_GR Po_qR Pi_rR.
void zfactor_ladder(Engine &e, const double *T1p, const double *Xp, double *R2p, int prec) {
  SISI_REQUIRE(e.NR > 0, SISI_ERR_STATE, "no Coulomb factors");
  const int64_t o = e.No, v = e.Nv, P = e.Np, NR = e.NR, fs = std::min<int64_t>(e.factors_slice, NR);
  ZV T1 = zdense(const_cast<double *>(T1p), {v, o});
  ZV X = zdense(const_cast<double *>(Xp), {v, v, o, o});
  ZV R2 = zdense(R2p, {v, v, o, o});
  ZV Pi = zdense(e.PiF.get(), {P, NR}), Po = zdense(e.PiO.get(), {P, NR});
  ZV PiA = zsub(Pi, {o, 0}, {P, NR}), PiI = zsub(Pi, {0, 0}, {o, NR});
  ZV PoA = zsub(Po, {o, 0}, {P, NR}), PoI = zsub(Po, {0, 0}, {o, NR});
  DBuf Db(&e.pool, size_t(2 * v * NR)), Cb(&e.pool, size_t(2 * v * NR));
  ZV D = zdense(Db.get(), {v, NR}), C = zdense(Cb.get(), {v, NR});
  zcopy(e, D, PoA);  // cxx:851-853
  zctr(e, -1.0, "kS,bk->bS", PoI, false, T1, false, 1.0, D, prec);
  zcopy(e, C, PiA, true);  // cxx:855-857
  zctr(e, -1.0, "kR,ak->aR", PiI, true, T1, false, 1.0, C, prec);
  DBuf X1b(&e.pool, size_t(2 * fs * v * o * o)), X2b(&e.pool, size_t(2 * fs * v * o * o)),
      XSb(&e.pool, size_t(2 * fs * fs * o * o));
  for (int64_t R0 = 0; R0 < NR; R0 += fs) {
    const int64_t Rx = std::min(fs, NR - R0);
    ZV X1 = zdense(X1b.get(), {Rx, v, o, o}), X2 = zdense(X2b.get(), {Rx, v, o, o});
    zctr(e, 1.0, "cdij,cR->Rdij", X, false, zsub(PoA, {0, R0}, {v, R0 + Rx}), true, 0.0, X1,
         prec);  // cxx:827
    for (int64_t S0 = 0; S0 < NR; S0 += fs) {
      const int64_t Ry = std::min(fs, NR - S0);
      ZV XS = zdense(XSb.get(), {Rx, Ry, o, o});
      zctr(e, 1.0, "Rdij,dS->RSij", X1, false, zsub(PiA, {0, S0}, {v, S0 + Ry}), false, 0.0, XS,
           prec);  // cxx:830
      scale_pairs(e, XSb.get(), Rx, Ry, o * o, e.VRS.get(), NR, NR * NR, R0, S0, false);  // :834
      zctr(e, 1.0, "RSij,bS->Rbij", XS, false, zsub(D, {0, S0}, {v, S0 + Ry}), false,
           S0 == 0 ? 0.0 : 1.0, X2, prec);  // cxx:867
    }
    zctr(e, 1.0, "Rbij,aR->abij", X2, false, zsub(C, {0, R0}, {v, R0 + Rx}), false, 1.0, R2,
         prec);  // cxx:874
  }
}

}  // namespace sisi
