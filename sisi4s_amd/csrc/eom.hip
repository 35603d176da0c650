// IP-EOM-CCSD right-apply on the device: SimilarityTransformedHamiltonian::
// right_apply_CCSD_IP (SimilarityTransformedHamiltonian.cxx:64-68), the
// product that UCcsdIPEquationOfMotionDavidson (cxx:43-205) repeats inside
// its Davidson iterations.  Spin orbitals, canonical orbitals, the six
// antisymmetrized classes of the spin-orbital CCSD engine (PPHH V_abij, HHHH
// V_ijkl, HHHP V_ijka, PPPH V_abci, PHPH V_aibj), amplitudes T1[a,i],
// T2[a,b,i,j]; R = (r1[i], r2[i,j,a]) with r2 antisymmetric in i, j.  The
// similarity-transformed blocks follow the intermediates of Stanton and Gauss
// (J. Chem. Phys. 101, 8938 (1994)); every product runs in FP64 on the
// contraction planner.  The blocks depend on the amplitudes only: they are
// built on the first right-apply after the amplitudes are set and kept on
// the engine, so each Davidson right-apply contracts r with stored blocks
// (the reference's intermediates: 1 form, cxx:64-68).
#include "ccsd.hpp"
#include "contract.hpp"
#include "tensor_ops.hpp"

#include <algorithm>

namespace sisi {

namespace {

// diagonal of an n x n matrix += eps
void add_diagonal(Engine &e, double *M, int64_t n, const double *eps) {
  View d;
  d.p = M;
  d.order = 1;
  d.len[0] = n;
  d.stride[0] = n + 1;
  copy(e, d, dense(const_cast<double *>(eps), {n}), 1.0, 1.0);
}

// tau = T2 + T_ai T_bj - T_bi T_aj and tau~ = T2 + (T_ai T_bj - T_bi T_aj) / 2
// of the amplitudes set (temporaries of the block builds)
struct Taus {
  DBuf taub, tautb;
  View tau, taut;
};

Taus make_taus(Engine &e) {
  const int64_t o = e.No, v = e.Nv, n2 = v * v * o * o;
  View T2 = dense(e.eom.T2.get(), {v, v, o, o});
  DBuf O1b(&e.pool, n2);
  Taus t{DBuf(&e.pool, n2), DBuf(&e.pool, n2), View(), View()};
  View O1 = dense(O1b.get(), {v, v, o, o});
  t.tau = dense(t.taub.get(), {v, v, o, o});
  t.taut = dense(t.tautb.get(), {v, v, o, o});
  zero(e, O1b.get(), n2);
  add_t1t1(e, O1b.get(), O1b.get(), e.eom.T1.get(), 1.0, int(v), int(o));
  copy(e, t.tau, T2);
  sum(e, 1.0, O1, "abij", 1.0, t.tau, "abij");
  sum(e, -1.0, O1, "baij", 1.0, t.tau, "abij");
  copy(e, t.taut, T2);
  sum(e, 0.5, O1, "abij", 1.0, t.taut, "abij");
  sum(e, -0.5, O1, "baij", 1.0, t.taut, "abij");
  return t;
}

// F_me, F_mi, F_ae and W_mbej, shared by the IP and EA right-applies.
void build_common_blocks(Engine &e, const Taus &t) {
  if (e.eom.Fme.get()) return;
  const int64_t o = e.No, v = e.Nv, n2 = v * v * o * o;
  const int p = SISI_FP64;
  View T1 = dense(e.eom.T1.get(), {v, o});
  View V = dense(e.Vabij.get(), {v, v, o, o}), Q = dense(e.Vijka.get(), {o, o, o, v}),
       P3 = dense(e.Vabci.get(), {v, v, v, o}), X = dense(e.Vaibj.get(), {v, o, v, o});
  // one-body blocks
  e.eom.Fme = DBuf(&e.pool, o * v);
  e.eom.Foo = DBuf(&e.pool, o * o);
  e.eom.Fvv = DBuf(&e.pool, v * v);
  View Fme = dense(e.eom.Fme.get(), {o, v}), Foo = dense(e.eom.Foo.get(), {o, o}),
       Fvv = dense(e.eom.Fvv.get(), {v, v});
  ctr(e, 1.0, "fn,efmn->me", T1, V, 0.0, Fme, p);        // t_nf <mn||ef>
  ctr(e, 1.0, "en,mnie->mi", T1, Q, 0.0, Foo, p);        // t_ne <mn||ie>
  ctr(e, 0.5, "efin,efmn->mi", t.taut, V, 1.0, Foo, p);  // tau~_inef <mn||ef> / 2
  ctr(e, 0.5, "ei,me->mi", T1, Fme, 1.0, Foo, p);        // t_ie F_me / 2
  add_diagonal(e, e.eom.Foo.get(), o, e.eps_i.get());    // f_mi
  ctr(e, -1.0, "fm,feam->ae", T1, P3, 0.0, Fvv, p);      // t_mf <ma||fe>
  ctr(e, -0.5, "afmn,efmn->ae", t.taut, V, 1.0, Fvv, p); // -tau~_mnaf <mn||ef> / 2
  ctr(e, -0.5, "am,me->ae", T1, Fme, 1.0, Fvv, p);       // -t_ma F_me / 2
  add_diagonal(e, e.eom.Fvv.get(), v, e.eps_a.get());    // f_ae

  // W_mbej = <mb||ej> + t_jf <mb||ef> - t_nb <mn||ej> - (t_jnfb + t_jf t_nb) <mn||ef>
  e.eom.Wovvo = DBuf(&e.pool, o * v * v * o);
  DBuf Sb(&e.pool, n2);
  View Wovvo = dense(e.eom.Wovvo.get(), {o, v, v, o}), S = dense(Sb.get(), {v, v, o, o});
  sum(e, -1.0, X, "bmej", 0.0, Wovvo, "mbej");
  ctr(e, -1.0, "fj,efbm->mbej", T1, P3, 1.0, Wovvo, p);
  ctr(e, 1.0, "bn,mnje->mbej", T1, Q, 1.0, Wovvo, p);
  add_t1t1(e, Sb.get(), e.eom.T2.get(), e.eom.T1.get(), 1.0, int(v), int(o));  // t_jnfb + t_jf t_nb
  ctr(e, -1.0, "fbjn,efmn->mbej", S, V, 1.0, Wovvo, p);
}

// W_mnij, W_mnie and W_mbij of the IP right-apply.
void build_ip_blocks(Engine &e) {
  if (e.eom.ip) return;
  const int64_t o = e.No, v = e.Nv;
  const int p = SISI_FP64;
  Taus t = make_taus(e);
  build_common_blocks(e, t);
  View T1 = dense(e.eom.T1.get(), {v, o}), T2 = dense(e.eom.T2.get(), {v, v, o, o});
  View V = dense(e.Vabij.get(), {v, v, o, o}), W4o = dense(e.Vijkl.get(), {o, o, o, o}),
       Q = dense(e.Vijka.get(), {o, o, o, v}), P3 = dense(e.Vabci.get(), {v, v, v, o}),
       X = dense(e.Vaibj.get(), {v, o, v, o});
  View Fme = dense(e.eom.Fme.get(), {o, v});

  // W_mnij = <mn||ij> + P(ij) t_je <mn||ie> + tau_ijef <mn||ef> / 2
  e.eom.Woooo = DBuf(&e.pool, o * o * o * o);
  DBuf Ab(&e.pool, o * o * o * o);
  View Woooo = dense(e.eom.Woooo.get(), {o, o, o, o}), A = dense(Ab.get(), {o, o, o, o});
  copy(e, Woooo, W4o);
  ctr(e, 1.0, "ej,mnie->mnij", T1, Q, 0.0, A, p);
  sum(e, 1.0, A, "mnij", 1.0, Woooo, "mnij");
  sum(e, -1.0, A, "mnji", 1.0, Woooo, "mnij");
  ctr(e, 0.5, "efij,efmn->mnij", t.tau, V, 1.0, Woooo, p);

  // W_mnie = <mn||ie> + t_if <mn||fe>
  e.eom.Wooov = DBuf(&e.pool, o * o * o * v);
  View Wooov = dense(e.eom.Wooov.get(), {o, o, o, v});
  copy(e, Wooov, Q);
  ctr(e, 1.0, "fi,femn->mnie", T1, V, 1.0, Wooov, p);

  // W_mbij = <mb||ij> - F_me t_ijbe - t_nb W_mnij + <mb||ef> tau_ijef / 2
  //          + P(ij) <mn||ie> t_jnbe + P(ij) t_ie (<mb||ej> - t_njbf <mn||ef>)
  e.eom.Wovoo = DBuf(&e.pool, o * v * o * o);
  DBuf tmpb(&e.pool, o * v * o * o), Yb(&e.pool, o * v * v * o);
  View Wovoo = dense(e.eom.Wovoo.get(), {o, v, o, o}), tmp = dense(tmpb.get(), {o, v, o, o}),
       Y = dense(Yb.get(), {o, v, v, o});
  sum(e, 1.0, Q, "ijmb", 0.0, Wovoo, "mbij");
  ctr(e, -1.0, "me,beij->mbij", Fme, T2, 1.0, Wovoo, p);
  ctr(e, -1.0, "bn,mnij->mbij", T1, Woooo, 1.0, Wovoo, p);
  ctr(e, -0.5, "efbm,efij->mbij", P3, t.tau, 1.0, Wovoo, p);
  ctr(e, 1.0, "mnie,bejn->mbij", Q, T2, 0.0, tmp, p);
  sum(e, 1.0, tmp, "mbij", 1.0, Wovoo, "mbij");
  sum(e, -1.0, tmp, "mbji", 1.0, Wovoo, "mbij");
  ctr(e, -1.0, "ei,bmej->mbij", T1, X, 0.0, tmp, p);
  ctr(e, 1.0, "bfnj,efmn->mbej", T2, V, 0.0, Y, p);     // t_njbf <mn||ef>
  ctr(e, -1.0, "ei,mbej->mbij", T1, Y, 1.0, tmp, p);
  sum(e, 1.0, tmp, "mbij", 1.0, Wovoo, "mbij");
  sum(e, -1.0, tmp, "mbji", 1.0, Wovoo, "mbij");
  e.eom.ip = true;
}

// W_abef, W_amef and W_abei of the EA right-apply.
void build_ea_blocks(Engine &e) {
  if (e.eom.ea) return;
  const int64_t o = e.No, v = e.Nv;
  const int p = SISI_FP64;
  Taus t = make_taus(e);
  build_common_blocks(e, t);
  View T1 = dense(e.eom.T1.get(), {v, o}), T2 = dense(e.eom.T2.get(), {v, v, o, o});
  View V = dense(e.Vabij.get(), {v, v, o, o}), W4v = dense(e.Vabcd.get(), {v, v, v, v}),
       Q = dense(e.Vijka.get(), {o, o, o, v}), P3 = dense(e.Vabci.get(), {v, v, v, o}),
       X = dense(e.Vaibj.get(), {v, o, v, o});
  View Fme = dense(e.eom.Fme.get(), {o, v});

  // W_abef = <ab||ef> - P(ab) t_mb <am||ef> + tau_mnab <mn||ef> / 2, <mb||ef> = -V_efbm
  e.eom.Wvvvv = DBuf(&e.pool, v * v * v * v);
  View Wvvvv = dense(e.eom.Wvvvv.get(), {v, v, v, v});
  copy(e, Wvvvv, W4v);
  ctr(e, 1.0, "am,efbm->abef", T1, P3, 1.0, Wvvvv, p);
  ctr(e, -1.0, "bm,efam->abef", T1, P3, 1.0, Wvvvv, p);
  ctr(e, 0.5, "abmn,efmn->abef", t.tau, V, 1.0, Wvvvv, p);

  // W_amef = <am||ef> - t_na <nm||ef>
  e.eom.Wvovv = DBuf(&e.pool, v * o * v * v);
  View Wvovv = dense(e.eom.Wvovv.get(), {v, o, v, v});
  sum(e, 1.0, P3, "efam", 0.0, Wvovv, "amef");
  ctr(e, -1.0, "an,efnm->amef", T1, V, 1.0, Wvovv, p);

  // W_abei = <ab||ei> - F_me t_miab + t_if W_abef + <mn||ei> tau_mnab / 2
  //          - P(ab) <mb||ef> t_miaf - P(ab) t_ma (<mb||ei> - t_nibf <mn||ef>)
  e.eom.Wvvvo = DBuf(&e.pool, v * v * v * o);
  DBuf tmpb(&e.pool, v * v * v * o), Yb(&e.pool, o * v * v * o);
  View Wvvvo = dense(e.eom.Wvvvo.get(), {v, v, v, o}), tmp = dense(tmpb.get(), {v, v, v, o}),
       Y = dense(Yb.get(), {o, v, v, o});
  copy(e, Wvvvo, P3);
  ctr(e, -1.0, "me,abmi->abei", Fme, T2, 1.0, Wvvvo, p);
  ctr(e, 1.0, "fi,abef->abei", T1, Wvvvv, 1.0, Wvvvo, p);
  ctr(e, -0.5, "mnie,abmn->abei", Q, t.tau, 1.0, Wvvvo, p);
  ctr(e, -1.0, "efbm,afmi->abei", P3, T2, 0.0, tmp, p);  // <mb||ef> t_miaf
  sum(e, -1.0, tmp, "abei", 1.0, Wvvvo, "abei");
  sum(e, 1.0, tmp, "baei", 1.0, Wvvvo, "abei");
  sum(e, -1.0, X, "bmei", 0.0, Y, "mbei");                 // <mb||ei> = -V_bmei
  ctr(e, -1.0, "bfni,efmn->mbei", T2, V, 1.0, Y, p);       // - t_nibf <mn||ef>
  ctr(e, 1.0, "am,mbei->abei", T1, Y, 0.0, tmp, p);
  sum(e, -1.0, tmp, "abei", 1.0, Wvvvo, "abei");
  sum(e, 1.0, tmp, "baei", 1.0, Wvvvo, "abei");
  e.eom.ea = true;
}

}  // namespace

void eom_set_amplitudes(Engine &e, const double *T1, const double *T2) {
  const size_t n1 = size_t(e.Nv) * e.No, n2 = n1 * n1;
  if (e.eom.have_t && std::equal(T1, T1 + n1, e.eom.hT1.begin()) &&
      std::equal(T2, T2 + n2, e.eom.hT2.begin()))
    return;  // same amplitudes: the blocks built from them stay
  e.drop_eom();
  e.eom.hT1.assign(T1, T1 + n1);
  e.eom.hT2.assign(T2, T2 + n2);
  e.upload(e.eom.T1, T1, n1);
  e.upload(e.eom.T2, T2, n2);
  e.eom.have_t = true;
}

void eom_ip_sigma(Engine &e, const double *r1c, const double *r2c, double *s1p, double *s2p) {
  SISI_REQUIRE(e.formulation == SISI_FORMULATION_UCCSD && ip_eom_ready(e), SISI_ERR_STATE,
               "the IP right-apply needs the antisymmetrized PPHH, HHHH, HHHP, PPPH and PHPH "
               "integrals (SISI_FORMULATION_UCCSD)");
  SISI_REQUIRE(e.eps_i.get() && e.eps_a.get(), SISI_ERR_STATE, "eigenenergies not set");
  SISI_REQUIRE(e.eom.have_t, SISI_ERR_STATE, "EOM amplitudes not set");
  const int64_t o = e.No, v = e.Nv;
  const int p = SISI_FP64;
  e.thaw();
  build_ip_blocks(e);
  View T2 = dense(e.eom.T2.get(), {v, v, o, o}), V = dense(e.Vabij.get(), {v, v, o, o});
  View r1 = dense(const_cast<double *>(r1c), {o}), r2 = dense(const_cast<double *>(r2c), {o, o, v});
  View s1 = dense(s1p, {o}), s2 = dense(s2p, {o, o, v});
  View Fme = dense(e.eom.Fme.get(), {o, v}), Foo = dense(e.eom.Foo.get(), {o, o}),
       Fvv = dense(e.eom.Fvv.get(), {v, v}), Woooo = dense(e.eom.Woooo.get(), {o, o, o, o}),
       Wovvo = dense(e.eom.Wovvo.get(), {o, v, v, o}),
       Wooov = dense(e.eom.Wooov.get(), {o, o, o, v}),
       Wovoo = dense(e.eom.Wovoo.get(), {o, v, o, o});

  // sigma_1 = -F_mi r_m + F_me r_mie - W_nmie r_mne / 2
  ctr(e, -1.0, "mi,m->i", Foo, r1, 0.0, s1, p);
  ctr(e, 1.0, "me,mie->i", Fme, r2, 1.0, s1, p);
  ctr(e, -0.5, "nmie,mne->i", Wooov, r2, 1.0, s1, p);
  // sigma_2 = -W_maji r_m + F_ae r_ije - P(ij) F_mi r_mja + W_mnij r_mna / 2
  //           + P(ij) W_maei r_mje + (<mn||ef> r_mnf / 2) t_ijae
  DBuf t3b(&e.pool, o * o * v), gb(&e.pool, v);
  View t3 = dense(t3b.get(), {o, o, v}), gv = dense(gb.get(), {v});
  ctr(e, -1.0, "maji,m->ija", Wovoo, r1, 0.0, s2, p);
  ctr(e, 1.0, "ae,ije->ija", Fvv, r2, 1.0, s2, p);
  ctr(e, 1.0, "mi,mja->ija", Foo, r2, 0.0, t3, p);
  sum(e, -1.0, t3, "ija", 1.0, s2, "ija");
  sum(e, 1.0, t3, "jia", 1.0, s2, "ija");
  ctr(e, 0.5, "mnij,mna->ija", Woooo, r2, 1.0, s2, p);
  ctr(e, 1.0, "maei,mje->ija", Wovvo, r2, 0.0, t3, p);
  sum(e, 1.0, t3, "ija", 1.0, s2, "ija");
  sum(e, -1.0, t3, "jia", 1.0, s2, "ija");
  ctr(e, 0.5, "efmn,mnf->e", V, r2, 0.0, gv, p);
  ctr(e, 1.0, "e,aeij->ija", gv, T2, 1.0, s2, p);
  e.thaw();
}

// EA-EOM-CCSD right-apply: SimilarityTransformedHamiltonian::right_apply_CCSD_EA
// (SimilarityTransformedHamiltonian.cxx:289-294), repeated by
// UCcsdEAEquationOfMotionDavidson (UCcsdEAEquationOfMotionDavidson.cxx:43-351).
// R = (r1[a], r2[j,a,b]) with r2 antisymmetric in a, b.  The particle-hole
// counterpart of the IP blocks (Stanton and Gauss, J. Chem. Phys. 103, 1064
// (1995)); needs PPPP for W_abef (Nv^4 doubles, as the reference's stored
// Vabcd), HHHH is not read.
void eom_ea_sigma(Engine &e, const double *r1c, const double *r2c, double *s1p, double *s2p) {
  SISI_REQUIRE(e.formulation == SISI_FORMULATION_UCCSD && ea_eom_ready(e), SISI_ERR_STATE,
               "the EA right-apply needs the antisymmetrized PPHH, PPPP, HHHP, PPPH and PHPH "
               "integrals (SISI_FORMULATION_UCCSD)");
  SISI_REQUIRE(e.eps_i.get() && e.eps_a.get(), SISI_ERR_STATE, "eigenenergies not set");
  SISI_REQUIRE(e.eom.have_t, SISI_ERR_STATE, "EOM amplitudes not set");
  const int64_t o = e.No, v = e.Nv;
  const int p = SISI_FP64;
  e.thaw();
  build_ea_blocks(e);
  View T2 = dense(e.eom.T2.get(), {v, v, o, o}), V = dense(e.Vabij.get(), {v, v, o, o});
  View r1 = dense(const_cast<double *>(r1c), {v}), r2 = dense(const_cast<double *>(r2c), {o, v, v});
  View s1 = dense(s1p, {v}), s2 = dense(s2p, {o, v, v});
  View Fme = dense(e.eom.Fme.get(), {o, v}), Foo = dense(e.eom.Foo.get(), {o, o}),
       Fvv = dense(e.eom.Fvv.get(), {v, v}), Wovvo = dense(e.eom.Wovvo.get(), {o, v, v, o}),
       Wvvvv = dense(e.eom.Wvvvv.get(), {v, v, v, v}),
       Wvovv = dense(e.eom.Wvovv.get(), {v, o, v, v}),
       Wvvvo = dense(e.eom.Wvvvo.get(), {v, v, v, o});

  // sigma_1 = F_ac r_c + F_ld r_lad + W_alcd r_lcd / 2
  ctr(e, 1.0, "ac,c->a", Fvv, r1, 0.0, s1, p);
  ctr(e, 1.0, "ld,lad->a", Fme, r2, 1.0, s1, p);
  ctr(e, 0.5, "alcd,lcd->a", Wvovv, r2, 1.0, s1, p);
  // sigma_2 = W_abcj r_c + P(ab) F_ac r_jcb - F_lj r_lab + P(ab) W_lbdj r_lad
  //           + W_abcd r_jcd / 2 - (<kl||cd> r_lcd / 2) t_kjab
  DBuf t3b(&e.pool, o * v * v), gb(&e.pool, o);
  View t3 = dense(t3b.get(), {o, v, v}), gk = dense(gb.get(), {o});
  ctr(e, 1.0, "abcj,c->jab", Wvvvo, r1, 0.0, s2, p);
  ctr(e, 1.0, "ac,jcb->jab", Fvv, r2, 0.0, t3, p);
  sum(e, 1.0, t3, "jab", 1.0, s2, "jab");
  sum(e, -1.0, t3, "jba", 1.0, s2, "jab");
  ctr(e, -1.0, "lj,lab->jab", Foo, r2, 1.0, s2, p);
  ctr(e, 1.0, "lbdj,lad->jab", Wovvo, r2, 0.0, t3, p);
  sum(e, 1.0, t3, "jab", 1.0, s2, "jab");
  sum(e, -1.0, t3, "jba", 1.0, s2, "jab");
  ctr(e, 0.5, "abcd,jcd->jab", Wvvvv, r2, 1.0, s2, p);
  ctr(e, 0.5, "cdkl,lcd->k", V, r2, 0.0, gk, p);
  ctr(e, -1.0, "k,abkj->jab", gk, T2, 1.0, s2, p);
  e.thaw();
}

}  // namespace sisi
