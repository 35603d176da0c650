// Direct-ring CCD: DrccdEnergyFromCoulombIntegrals::getResiduum
// (DrccdEnergyFromCoulombIntegrals.cxx:31-80) on the engine's CCSD driver
// (ClusterSinglesDoublesAlgorithm::run), real amplitudes.  Doubles only:
//   R_abij = V_abij + s V_akic T_cbkj + s V_bkjc T_acik
//            + C_alid T_dblj,   C_alid = s W_klcd T_acik,
//   W_klcd = s V_klcd (- V_lkcd with adjacentPairsExchange),
// s = 2 (1 when unrestricted), V_akic = PHHP, V_klcd = HHPP.  Every product
// is a ring contraction (M = N = Nv No, K = Nv No) on the MFMA GEMM family.
#include "ccsd.hpp"
#include "contract.hpp"
#include "tensor_ops.hpp"
#include "ztensor.hpp"

namespace sisi {

void set_drccd(Engine &e, int unrestricted, int linearized, int adjacent_pairs_exchange) {
  e.spins = unrestricted ? 1.0 : 2.0;  // ClusterSinglesDoublesAlgorithm.cxx:133, cxx:43
  e.ring_linearized = linearized != 0;
  e.ring_adjacent = adjacent_pairs_exchange != 0;
  e.drop_residual_graph();
}

bool integrals_ready(const Engine &e) {
  if (e.formulation == SISI_FORMULATION_UCCSD) return uccsd_ready(e);
  if (e.formulation == SISI_FORMULATION_DRCCD) {
    if (e.cplx) return e.have_integrals && e.Vaijb.get() != nullptr && e.Vijab.get() != nullptr;
    return e.Vabij.get() != nullptr && e.Uabij.get() != nullptr;
  }
  return e.have_integrals;
}

namespace {

// F = complex (the reference's template with complex PPHH, PHHP and HHPP of
// CoulombIntegralsFromVertex's complex path): the same products, no
// conjugations, on planar complex amplitudes.
void zdrccd_residual(Engine &e, int it, bool initial_doubles, const double *T2c, double *R1p,
                     double *R2p) {
  const int64_t o = e.No, v = e.Nv;
  const int64_t n2 = v * v * o * o;
  double *T2p = const_cast<double *>(T2c);
  ZV T2 = zdense(T2p, {v, v, o, o}), R2 = zdense(R2p, {v, v, o, o});
  ZV V = zdense(e.Vabij.get(), {v, v, o, o}), Vaijb = zdense(e.Vaijb.get(), {v, o, o, v}),
     Vijab = zdense(e.Vijab.get(), {o, o, v, v});
  zero(e, R1p, int64_t(e.n1()));
  zcopy(e, R2, V);  // cxx:57, 77
  if (it == 0 && !initial_doubles) return;
  const int prec = e.precision;
  const double s = e.spins;
  e.thaw();
  for (const double *p : {T2p, e.Vaijb.get(), e.Vijab.get()}) {
    e.frozen.push_back(p);
    e.frozen.push_back(p + n2);
  }
  zctr(e, s, "akic,cbkj->abij", Vaijb, false, T2, false, 1.0, R2, prec);  // cxx:58
  zctr(e, s, "bkjc,acik->abij", Vaijb, false, T2, false, 1.0, R2, prec);  // cxx:59
  if (!e.ring_linearized) {
    DBuf Wb(&e.pool, size_t(2 * n2)), Cb(&e.pool, size_t(2 * n2));
    ZV W = zdense(Wb.get(), {o, o, v, v}), C = zdense(Cb.get(), {v, o, o, v});
    zcopy(e, W, Vijab, false, s);                                            // cxx:61-62
    if (e.ring_adjacent) zsum(e, -1.0, Vijab, "jiab", false, 1.0, W, "ijab");  // cxx:63-65
    zctr(e, s, "klcd,acik->alid", W, false, T2, false, 0.0, C, prec);          // cxx:68
    zctr(e, 1.0, "alid,dblj->abij", C, false, T2, false, 1.0, R2, prec);       // cxx:69
  }
  e.thaw();
}

}  // namespace

void drccd_residual(Engine &e, int it, bool initial_doubles, const double *T2c, double *R1p,
                    double *R2p) {
  SISI_REQUIRE(integrals_ready(e), SISI_ERR_STATE, "direct-ring CCD needs the PPHH integrals");
  if (e.cplx) return zdrccd_residual(e, it, initial_doubles, T2c, R1p, R2p);
  const int64_t o = e.No, v = e.Nv;
  const int64_t n2 = v * v * o * o;
  double *T2p = const_cast<double *>(T2c);
  View V = dense(e.Vabij.get(), {v, v, o, o});
  View T2 = dense(T2p, {v, v, o, o}), R2 = dense(R2p, {v, v, o, o});
  zero(e, R1p, v * o);
  copy(e, R2, V);  // cxx:57 and the MP2 start, cxx:77
  if (it == 0 && !initial_doubles) return;
  // PHHP and HHPP: uploaded, or V_aijb = V_abji and V_ijab = V_abij as
  // CoulombIntegralsFromVertex.cxx:491, 515 derive them
  DBuf Aloc, Bloc;
  const double *Vaijb_p = e.ring_Vaijb.get(), *Vijab_p = e.ring_Vijab.get();
  if (!Vaijb_p) {
    Aloc = DBuf(&e.pool, size_t(n2));
    sum(e, 1.0, V, "abji", 0.0, dense(Aloc.get(), {v, o, o, v}), "aijb");
    Vaijb_p = Aloc.get();
  }
  if (!Vijab_p) {
    Bloc = DBuf(&e.pool, size_t(n2));
    sum(e, 1.0, V, "abij", 0.0, dense(Bloc.get(), {o, o, v, v}), "ijab");
    Vijab_p = Bloc.get();
  }
  View Vaijb = dense(const_cast<double *>(Vaijb_p), {v, o, o, v});
  View Vijab = dense(const_cast<double *>(Vijab_p), {o, o, v, v});
  const int prec = e.precision;
  const double s = e.spins;
  e.thaw();
  e.frozen = {T2p, Vaijb_p, Vijab_p};
  ctr(e, s, "akic,cbkj->abij", Vaijb, T2, 1.0, R2, prec);  // cxx:58
  ctr(e, s, "bkjc,acik->abij", Vaijb, T2, 1.0, R2, prec);  // cxx:59
  if (!e.ring_linearized) {
    DBuf Wb(&e.pool, size_t(n2)), Cb(&e.pool, size_t(n2));
    View W = dense(Wb.get(), {o, o, v, v}), C = dense(Cb.get(), {v, o, o, v});
    axpby(e, n2, s, Vijab_p, 0.0, Wb.get());                      // cxx:61-62
    if (e.ring_adjacent) sum(e, -1.0, Vijab, "jiab", 1.0, W, "ijab");  // cxx:63-65
    ctr(e, s, "klcd,acik->alid", W, T2, 0.0, C, prec);            // cxx:67-68
    ctr(e, 1.0, "alid,dblj->abij", C, T2, 1.0, R2, prec);         // cxx:69
  }
  e.thaw();
}

}  // namespace sisi
