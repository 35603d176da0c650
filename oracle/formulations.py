"""Independent cross-formulations from the reference tree (test infrastructure).

These pin the restatement in ``ccsd.py`` / ``triples.py``: a mistake there (a
dropped term, a wrong sign, a transposed operand) makes them disagree.

* ``residual_full_integrals`` -- CcsdEnergyFromCoulombIntegralsReference.cxx:29-295
  (stored V_abcd / V_abci, no vertex, no slicing); ``residual_only_ppl`` its
  onlyPPL branch (cxx:56-73).
* ``triples_unsliced``       -- PerturbativeTriples.cxx:19-97 (runInMemory:
  6-index T_abcijk, Z_ai / Z_abij intermediates).
* ``mp2_closed_form``        -- ParticleHoleCoulombVertexDecomposition.cxx:279-323,
  Mp2NaturalOrbitals.cxx:57-74: E = sum V(2V - V')/D.
"""

import numpy as np

from .ccsd import ein, integrals_from_vertex, vabcd_from_vertex


def residual_full_integrals(it, T1, T2, gamma, No, initial_doubles=False):
    """CcsdEnergyFromCoulombIntegralsReference::getResiduum (cxx:29-295)."""
    ints = integrals_from_vertex(gamma, No)
    Vabij, Vaibj, Vijkl, Vijka = ints["PPHH"], ints["PHPH"], ints["HHHH"], ints["HHHP"]
    Vabcd, Vabci = vabcd_from_vertex(gamma, No)
    R1 = np.zeros_like(T1)
    R2 = np.zeros_like(T2)
    if it == 0 and not initial_doubles:
        R2[...] = Vabij
        return R1, R2
    tt = ein("ai,bj->abij", T1, T1)
    Kac = (-2.0 * ein("cdkl,adkl->ac", Vabij, T2) + ein("dckl,adkl->ac", Vabij, T2)
           - 2.0 * ein("cdkl,ak,dl->ac", Vabij, T1, T1) + ein("dckl,ak,dl->ac", Vabij, T1, T1))
    Lac = Kac + 2.0 * ein("cdak,dk->ac", Vabci, T1) - ein("dcak,dk->ac", Vabci, T1)
    Kki = (2.0 * ein("cdkl,cdil->ki", Vabij, T2) - ein("dckl,cdil->ki", Vabij, T2)
           + 2.0 * ein("cdkl,ci,dl->ki", Vabij, T1, T1) - ein("dckl,ci,dl->ki", Vabij, T1, T1))
    Lki = Kki + 2.0 * ein("klic,cl->ki", Vijka, T1) - ein("lkic,cl->ki", Vijka, T1)
    R2 += ein("ac,cbij->abij", Lac, T2)
    R2 -= ein("ki,abkj->abij", Lki, T2)
    R2 += ein("baci,cj->abij", Vabci, T1)
    R2 -= ein("bkci,ak,cj->abij", Vaibj, T1, T1)
    R2 -= ein("jika,bk->abij", Vijka, T1)
    R2 -= ein("acik,cj,bk->abij", Vabij, T1, T1)
    Xakic = (ein("acik->akic", Vabij) - ein("lkic,al->akic", Vijka, T1)
             + ein("acdk,di->akic", Vabci, T1) - 0.5 * ein("dclk,dail->akic", Vabij, T2)
             - ein("dclk,di,al->akic", Vabij, T1, T1) + ein("dclk,adil->akic", Vabij, T2)
             - 0.5 * ein("cdlk,adil->akic", Vabij, T2))
    Xakci = (ein("akci->akci", Vaibj) - ein("klic,al->akci", Vijka, T1)
             + ein("adck,di->akci", Vabci, T1) - 0.5 * ein("cdlk,dail->akci", Vabij, T2)
             - ein("cdlk,di,al->akci", Vabij, T1, T1))
    R2 += 2.0 * ein("akic,cbkj->abij", Xakic, T2) - ein("akic,bckj->abij", Xakic, T2)
    R2 -= ein("akci,cbkj->abij", Xakci, T2)
    R2 -= ein("bkci,ackj->abij", Xakci, T2)
    R2 += ein("baji->abij", R2)
    R2 += Vabij
    Xklij = (Vijkl + ein("klic,cj->klij", Vijka, T1) + ein("lkjc,ci->klij", Vijka, T1)
             + ein("cdkl,cdij->klij", Vabij, T2) + ein("cdkl,ci,dj->klij", Vabij, T1, T1))
    R2 += ein("klij,abkl->abij", Xklij, T2)
    R2 += ein("klij,ak,bl->abij", Xklij, T1, T1)
    Xabcd = Vabcd - ein("cdak,bk->abcd", Vabci, T1) - ein("dcbk,ak->abcd", Vabci, T1)
    R2 += ein("abcd,cdij->abij", Xabcd, T2)
    R2 += ein("abcd,ci,dj->abij", Xabcd, T1, T1)
    Kck = 2.0 * ein("cdkl,dl->ck", Vabij, T1) - ein("cdlk,dl->ck", Vabij, T1)
    R1 += ein("ac,ci->ai", Kac, T1) - ein("ki,ak->ai", Kki, T1)
    R1 += 2.0 * ein("ck,caki->ai", Kck, T2) - ein("ck,caik->ai", Kck, T2)
    R1 += ein("ck,ci,ak->ai", Kck, T1, T1)
    R1 += 2.0 * ein("acik,ck->ai", Vabij, T1) - ein("akci,ck->ai", Vaibj, T1)
    R1 += 2.0 * ein("cdak,cdik->ai", Vabci, T2) - ein("dcak,cdik->ai", Vabci, T2)
    R1 += 2.0 * ein("cdak,ci,dk->ai", Vabci, T1, T1) - ein("dcak,ci,dk->ai", Vabci, T1, T1)
    R1 += -2.0 * ein("klic,ackl->ai", Vijka, T2) + ein("lkic,ackl->ai", Vijka, T2)
    R1 += -2.0 * ein("klic,ak,cl->ai", Vijka, T1, T1) + ein("lkic,ak,cl->ai", Vijka, T1, T1)
    return R1, R2


def residual_only_ppl(T1, T2, gamma, No):
    """onlyPPL branch of CcsdEnergyFromCoulombIntegralsReference::getResiduum
    (cxx:56-73): R2 = V_abcd (T2 + T1 T1), R1 = 0, at every iteration."""
    Vabcd, _ = vabcd_from_vertex(gamma, No)
    return np.zeros_like(T1), ein("abcd,cdij->abij", Vabcd, T2 + ein("ai,bj->abij", T1, T1))


def triples_unsliced(T1, T2, gamma, eps_i, eps_a):
    """PerturbativeTriples::runInMemory (PerturbativeTriples.cxx:19-97).
    Returns the (T) energy (without the CCSD energy)."""
    No = eps_i.shape[0]
    ints = integrals_from_vertex(gamma, No)
    Vabij, Vijka = ints["PPHH"], ints["HHHP"]
    _, Vabci = vabcd_from_vertex(gamma, No)
    Z = ein("bcdk,adij->abcijk", Vabci, T2) - ein("jklc,abil->abcijk", Vijka, T2)
    D = (eps_i[None, None, None, :, None, None] + eps_i[None, None, None, None, :, None]
         + eps_i[None, None, None, None, None, :] - eps_a[:, None, None, None, None, None]
         - eps_a[None, :, None, None, None, None] - eps_a[None, None, :, None, None, None])
    T = Z / D
    Zc = T.copy()
    for perm in ("bacjik", "acbikj", "cbakji", "cabkij", "bcajki"):
        T += ein(perm + "->abcijk", Zc)
    Zai = (2.0 * ein("acdikl,cdkl->ai", T, Vabij) - ein("acdikl,dckl->ai", T, Vabij)
           - 2.0 * ein("acdlki,cdkl->ai", T, Vabij) + ein("acdlki,dckl->ai", T, Vabij))
    Zabij = (2.0 * ein("acdijk,cdbk->abij", T, Vabci) - ein("acdijk,dcbk->abij", T, Vabci)
             - ein("acdkji,cdbk->abij", T, Vabci)
             - 2.0 * ein("abcikl,kljc->abij", T, Vijka) + ein("abcikl,lkjc->abij", T, Vijka)
             + ein("abclki,kljc->abij", T, Vijka))
    return (2.0 * np.sum(Zai * T1) + 4.0 * np.sum(Zabij * T2)
            - 2.0 * np.sum(Zabij * ein("abji->abij", T2)))


def mp2_closed_form(gamma, eps_i, eps_a):
    """E_MP2 = sum_abij V_abij (2 V_abij - V_baij) / (e_i + e_j - e_a - e_b)."""
    No = eps_i.shape[0]
    Gai = gamma[:, No:, :No]
    V = ein("Gai,Gbj->abij", Gai, Gai)
    D = (eps_i[None, None, :, None] + eps_i[None, None, None, :]
         - eps_a[:, None, None, None] - eps_a[None, :, None, None])
    return float(np.sum(V * (2.0 * V - ein("baij->abij", V)) / D))
