// Spin-orbital CCSD: UccsdAmplitudesFromCoulombIntegrals::getResiduumTemplate
// (UccsdAmplitudesFromCoulombIntegrals.cxx:107-241), canonical orbitals.
// The residual is <Phi_i^a|Hbar|Phi>, <Phi_ij^ab|Hbar|Phi> minus the
// diagonal Fock terms, evaluated in the intermediate form of Stanton and
// Gauss (J. Chem. Phys. 94, 4334 (1991)) from six uploaded antisymmetrized
// classes <pq||rs> (real): PPHH V_abij, HHHH V_ijkl, PPPP V_abcd, HHHP V_ijka,
// PPPH V_abci, PHPH V_aibj.  Every other block is one of these with its
// indices reordered, which the contraction specs below express directly
// (the planner permutes operands as needed):
//   <mn||ef> = V_efmn   <ma||fe> = -V_fe,am (PPPH)   <mb||ej> = -V_bm,ej (PHPH)
//   <mn||ie> = V_mnie (HHHP)   <mb||ij> = V_ij,mb (HHHP)   <ab||ej> = V_abej (PPPH)
// The particle ladder (1/2) tau_ijef W_abef is expanded so that no Nv^4
// intermediate is formed besides the stored <ab||ef>.  Amplitudes T1[a,i],
// T2[a,b,i,j]; the energy is the antisymmetrized branch of getEnergy
// (ClusterSinglesDoublesAlgorithm.cxx:163-167).
#include "ccsd.hpp"
#include "contract.hpp"
#include "tensor_ops.hpp"

namespace sisi {

namespace {

constexpr unsigned kUccsdClasses = (1u << SISI_PPHH) | (1u << SISI_HHHH) | (1u << SISI_PPPP) |
                                   (1u << SISI_HHHP) | (1u << SISI_PPPH) | (1u << SISI_PHPH);

// R += x - x^(swap of the two named index letters), x in "abij" layout
void add_antisym(Engine &e, const View &x, const char *swapped, const View &R) {
  sum(e, 1.0, x, "abij", 1.0, R, "abij");
  sum(e, -1.0, x, swapped, 1.0, R, "abij");
}

}  // namespace

bool uccsd_ready(const Engine &e) {
  return !e.cplx && (e.given_integrals & kUccsdClasses) == kUccsdClasses;
}

bool ip_eom_ready(const Engine &e) {
  constexpr unsigned kIp = kUccsdClasses & ~(1u << SISI_PPPP);
  return !e.cplx && (e.given_integrals & kIp) == kIp;
}

bool ea_eom_ready(const Engine &e) {
  constexpr unsigned kEa = kUccsdClasses & ~(1u << SISI_HHHH);
  return !e.cplx && (e.given_integrals & kEa) == kEa;
}

void uccsd_residual(Engine &e, int it, bool initial_doubles, const double *T1c, const double *T2c,
                    double *R1p, double *R2p) {
  SISI_REQUIRE(!e.cplx, SISI_ERR_INVALID, "the spin-orbital CCSD residual is built for real amplitudes");
  SISI_REQUIRE(uccsd_ready(e), SISI_ERR_STATE,
               "spin-orbital CCSD needs the antisymmetrized PPHH, HHHH, PPPP, HHHP, PPPH and PHPH "
               "integrals uploaded");
  const int64_t o = e.No, v = e.Nv;
  const int64_t n2 = v * v * o * o;
  double *T1p = const_cast<double *>(T1c), *T2p = const_cast<double *>(T2c);
  View T1 = dense(T1p, {v, o}), T2 = dense(T2p, {v, v, o, o});
  View R1 = dense(R1p, {v, o}), R2 = dense(R2p, {v, v, o, o});
  View V = dense(e.Vabij.get(), {v, v, o, o}), Hoooo = dense(e.Vijkl.get(), {o, o, o, o}),
       Pvvvv = dense(e.Vabcd.get(), {v, v, v, v}), Q = dense(e.Vijka.get(), {o, o, o, v}),
       P3 = dense(e.Vabci.get(), {v, v, v, o}), X = dense(e.Vaibj.get(), {v, o, v, o});
  zero(e, R1p, v * o);
  copy(e, R2, V);  // <ij||ab>; the MP2 start (cxx:158-161) is this alone
  if (it == 0 && !initial_doubles) return;
  const int prec = e.precision;
  e.thaw();
  e.frozen = {T1p, T2p, e.Vabij.get(), e.Vijkl.get(), e.Vabcd.get(), e.Vijka.get(),
              e.Vabci.get(), e.Vaibj.get()};

  // tau~ = T2 + 1/2 (T_ai T_bj - T_bi T_aj), tau = T2 + T_ai T_bj - T_bi T_aj
  DBuf O1b(&e.pool, n2), taub(&e.pool, n2), tautb(&e.pool, n2);
  View O1 = dense(O1b.get(), {v, v, o, o}), tau = dense(taub.get(), {v, v, o, o}),
       taut = dense(tautb.get(), {v, v, o, o});
  zero(e, O1b.get(), n2);
  add_t1t1(e, O1b.get(), O1b.get(), T1p, 1.0, int(v), int(o));
  copy(e, tau, T2);
  add_antisym(e, O1, "baij", tau);
  copy(e, taut, T2);
  sum(e, 0.5, O1, "abij", 1.0, taut, "abij");
  sum(e, -0.5, O1, "baij", 1.0, taut, "abij");
  e.frozen.push_back(taub.get());
  e.frozen.push_back(tautb.get());

  // one-particle intermediates
  DBuf Faeb(&e.pool, v * v), Fmib(&e.pool, o * o), Fmeb(&e.pool, o * v);
  View Fae = dense(Faeb.get(), {v, v}), Fmi = dense(Fmib.get(), {o, o}),
       Fme = dense(Fmeb.get(), {o, v});
  ctr(e, -1.0, "fm,feam->ae", T1, P3, 0.0, Fae, prec);          // t_mf <ma||fe>
  ctr(e, -0.5, "afmn,efmn->ae", taut, V, 1.0, Fae, prec);       // -1/2 tau~_mnaf <mn||ef>
  ctr(e, 1.0, "en,mnie->mi", T1, Q, 0.0, Fmi, prec);            // t_ne <mn||ie>
  ctr(e, 0.5, "efin,efmn->mi", taut, V, 1.0, Fmi, prec);        // 1/2 tau~_inef <mn||ef>
  ctr(e, 1.0, "fn,efmn->me", T1, V, 0.0, Fme, prec);            // t_nf <mn||ef>

  // T1 equations
  ctr(e, 1.0, "ei,ae->ai", T1, Fae, 0.0, R1, prec);
  ctr(e, -1.0, "am,mi->ai", T1, Fmi, 1.0, R1, prec);
  ctr(e, 1.0, "aeim,me->ai", T2, Fme, 1.0, R1, prec);
  ctr(e, -1.0, "fn,anfi->ai", T1, X, 1.0, R1, prec);            // -t_nf <na||if>
  ctr(e, 0.5, "efim,efam->ai", T2, P3, 1.0, R1, prec);          // -1/2 t_imef <ma||ef>
  ctr(e, 0.5, "aemn,nmie->ai", T2, Q, 1.0, R1, prec);           // -1/2 t_mnae <nm||ei>

  DBuf tmpb(&e.pool, n2);
  View tmp = dense(tmpb.get(), {v, v, o, o});
  // P(ab) t_ijae (F_be - 1/2 t_mb F_me)
  {
    DBuf Gb(&e.pool, v * v);
    View G = dense(Gb.get(), {v, v});
    copy(e, G, Fae);
    ctr(e, -0.5, "bm,me->be", T1, Fme, 1.0, G, prec);
    ctr(e, 1.0, "aeij,be->abij", T2, G, 0.0, tmp, prec);
    add_antisym(e, tmp, "baij", R2);
  }
  // -P(ij) t_imab (F_mj + 1/2 t_je F_me)
  {
    DBuf Hb(&e.pool, o * o);
    View H = dense(Hb.get(), {o, o});
    copy(e, H, Fmi);
    ctr(e, 0.5, "ej,me->mj", T1, Fme, 1.0, H, prec);
    ctr(e, -1.0, "abim,mj->abij", T2, H, 0.0, tmp, prec);
    add_antisym(e, tmp, "abji", R2);
  }
  // hole ladder 1/2 tau_mnab W_mnij and the tau.tau part of the particle
  // ladder, 1/8 tau_mnab Y_mnij with Y_mnij = tau_ijef <mn||ef>
  {
    DBuf Yb(&e.pool, o * o * o * o), Wb(&e.pool, o * o * o * o), Ab(&e.pool, o * o * o * o);
    View Y = dense(Yb.get(), {o, o, o, o}), W = dense(Wb.get(), {o, o, o, o}),
         A = dense(Ab.get(), {o, o, o, o});
    ctr(e, 1.0, "efij,efmn->mnij", tau, V, 0.0, Y, prec);
    ctr(e, 1.0, "ej,mnie->mnij", T1, Q, 0.0, A, prec);        // t_je <mn||ie>
    // 1/2 W_mnij + 1/8 Y = 1/2 <mn||ij> + 1/2 P(ij) A + (1/8 + 1/8) Y
    axpby(e, o * o * o * o, 0.5, e.Vijkl.get(), 0.0, Wb.get());
    sum(e, 0.5, A, "mnij", 1.0, W, "mnij");
    sum(e, -0.5, A, "mnji", 1.0, W, "mnij");
    axpby(e, o * o * o * o, 0.25, Yb.get(), 1.0, Wb.get());
    ctr(e, 1.0, "abmn,mnij->abij", tau, W, 1.0, R2, prec);
  }
  // particle ladder: 1/2 tau_ijef <ab||ef> - 1/2 P(ab) t_mb tau_ijef <am||ef>
  {
    ctr(e, 0.5, "abef,efij->abij", Pvvvv, tau, 1.0, R2, prec);
    DBuf Zb(&e.pool, v * o * o * o);
    View Z = dense(Zb.get(), {v, o, o, o});
    ctr(e, 1.0, "efam,efij->amij", P3, tau, 0.0, Z, prec);    // tau_ijef <am||ef>
    ctr(e, -0.5, "bm,amij->abij", T1, Z, 0.0, tmp, prec);
    add_antisym(e, tmp, "baij", R2);
  }
  // P(ij) P(ab) (t_imae W_mbej - t_ie t_ma <mb||ej>)
  {
    DBuf Wb(&e.pool, o * v * v * o), Sb(&e.pool, n2), Ub(&e.pool, o * v * o * o);
    View W = dense(Wb.get(), {o, v, v, o}), S = dense(Sb.get(), {v, v, o, o}),
         U = dense(Ub.get(), {o, v, o, o});
    sum(e, -1.0, X, "bmej", 0.0, W, "mbej");                  // <mb||ej>
    ctr(e, -1.0, "fj,efbm->mbej", T1, P3, 1.0, W, prec);      // t_jf <mb||ef>
    ctr(e, 1.0, "bn,mnje->mbej", T1, Q, 1.0, W, prec);        // -t_nb <mn||ej>
    // S_fbjn = 1/2 t_jnfb + t_jf t_nb
    axpby(e, n2, 0.5, T2p, 0.0, Sb.get());
    add_t1t1(e, Sb.get(), Sb.get(), T1p, 1.0, int(v), int(o));
    ctr(e, -1.0, "fbjn,efmn->mbej", S, V, 1.0, W, prec);
    ctr(e, 1.0, "aeim,mbej->abij", T2, W, 0.0, tmp, prec);
    ctr(e, -1.0, "ei,bmej->mbij", T1, X, 0.0, U, prec);       // t_ie <mb||ej>
    ctr(e, -1.0, "am,mbij->abij", T1, U, 1.0, tmp, prec);
    sum(e, 1.0, tmp, "abij", 1.0, R2, "abij");
    sum(e, -1.0, tmp, "abji", 1.0, R2, "abij");
    sum(e, -1.0, tmp, "baij", 1.0, R2, "abij");
    sum(e, 1.0, tmp, "baji", 1.0, R2, "abij");
  }
  // P(ij) t_ie <ab||ej> - P(ab) t_ma <mb||ij>
  ctr(e, 1.0, "ei,abej->abij", T1, P3, 0.0, tmp, prec);
  add_antisym(e, tmp, "abji", R2);
  ctr(e, -1.0, "am,ijmb->abij", T1, Q, 0.0, tmp, prec);
  add_antisym(e, tmp, "baij", R2);
  e.thaw();
}

double uccsd_energy(Engine &e, const double *T1, const double *T2) {
  SISI_REQUIRE(uccsd_ready(e), SISI_ERR_STATE, "spin-orbital CCSD integrals not uploaded");
  const int64_t o = e.No, v = e.Nv, n2 = v * v * o * o;
  DBuf Yb(&e.pool, n2);
  add_t1t1(e, Yb.get(), T2, T1, 2.0, int(v), int(o));  // T2 + 2 T1 T1
  return 0.25 * dot(e, n2, Yb.get(), e.Vabij.get());
}

}  // namespace sisi
