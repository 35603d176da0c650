"""CPU checks of the algebra behind the refinement residual's ladder term.

The refinement residual evaluates the particle-particle ladder's change
V~(T) X(T) - V~(T_A) X(T_A) without per-slice dV~ products: the dressed
Gamma_ab changes only through T1 (D = -sum_k dT1[a,k] Gamma_bk), so with
S[c,k,b,d] = sum_G Gamma_ck^G Gamma~_bd^G(T_A) and V[c,d,k,l] = sum_G
Gamma_ck^G Gamma_dl^G the change is -dT1.(S.X) + its (a,i)<->(b,j) mirror +
dT1 dT1.(V.X) + V~(T_A) dX.  These tests restate both forms with numpy on
small random tensors (the ladder of CcsdEnergyFromCoulombIntegrals.cxx:336-453,
dressing at cxx:397-400) and check that they agree, and that the identity
needs the closed-shell symmetry T_abij = T_baji, which mixed runs enforce.
"""

import numpy as np
import pytest


def _system(seed, v=7, o=3, G=11, sym=True):
    rng = np.random.default_rng(seed)
    Gai = rng.standard_normal((G, v, o))
    Gab = rng.standard_normal((G, v, v))
    T1A = 0.1 * rng.standard_normal((v, o))
    d1 = 0.05 * rng.standard_normal((v, o))
    T2A = 0.1 * rng.standard_normal((v, v, o, o))
    d2 = 0.05 * rng.standard_normal((v, v, o, o))
    if sym:
        T2A = T2A + T2A.transpose(1, 0, 3, 2)
        d2 = d2 + d2.transpose(1, 0, 3, 2)
    return Gai, Gab, T1A, d1, T2A, d2


def _dressed(Gab, Gai, T1):
    # Gamma~[G,a,b] = Gamma_ab - sum_k Gamma_bk T1[a,k]  (cxx:397-400)
    return Gab - np.einsum("Gbk,ak->Gab", Gai, T1)


def _ladder(Gab, Gai, T1, T2):
    X = T2 + np.einsum("ai,bj->abij", T1, T1)
    Gd = _dressed(Gab, Gai, T1)
    Vt = np.einsum("Gac,Gbd->abcd", Gd, Gd)
    return np.einsum("abcd,cdij->abij", Vt, X)


def _low_rank_change(Gab, Gai, T1A, d1, T2A, d2):
    T1, T2 = T1A + d1, T2A + d2
    X = T2 + np.einsum("ai,bj->abij", T1, T1)
    XA = T2A + np.einsum("ai,bj->abij", T1A, T1A)
    GA = _dressed(Gab, Gai, T1A)
    S = np.einsum("Gck,Gbd->ckbd", Gai, GA)
    QA = np.einsum("ckbd,cdij->kbij", S, X)
    termA = -np.einsum("ak,kbij->abij", d1, QA)
    V = np.einsum("Gck,Gdl->cdkl", Gai, Gai)
    QW = np.einsum("cdkl,cdij->klij", V, X)
    termC = np.einsum("ak,bl,klij->abij", d1, d1, QW)
    VA = np.einsum("Gac,Gbd->abcd", GA, GA)
    return termA + termA.transpose(1, 0, 3, 2) + termC + np.einsum("abcd,cdij->abij", VA, X - XA)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_low_rank_ladder_change_equals_the_difference(seed):
    Gai, Gab, T1A, d1, T2A, d2 = _system(seed)
    full = _ladder(Gab, Gai, T1A + d1, T2A + d2) - _ladder(Gab, Gai, T1A, T2A)
    low = _low_rank_change(Gab, Gai, T1A, d1, T2A, d2)
    assert np.abs(low - full).max() <= 1e-12 * np.abs(full).max()


def test_low_rank_change_without_the_second_order_term_misses_it():
    """The D^T D term is second order in dT1: dropping it leaves an error of
    that size (a plausible slip the first check would not pass)."""
    Gai, Gab, T1A, d1, T2A, d2 = _system(4)
    full = _ladder(Gab, Gai, T1A + d1, T2A + d2) - _ladder(Gab, Gai, T1A, T2A)
    low = _low_rank_change(Gab, Gai, T1A, d1, T2A, d2)
    T1 = T1A + d1
    X = T2A + d2 + np.einsum("ai,bj->abij", T1, T1)
    V = np.einsum("Gck,Gdl->cdkl", Gai, Gai)
    termC = np.einsum("ak,bl,klij->abij", d1, d1, np.einsum("cdkl,cdij->klij", V, X))
    assert np.abs(termC).max() > 1e-6 * np.abs(full).max()
    assert np.abs(low - termC - full).max() > 1e3 * np.abs(low - full).max()


def test_low_rank_identity_needs_symmetric_doubles():
    """With an antisymmetric part in T2 the mirror term is no longer the
    (G_A^T D) X contribution, so the identity fails: mixed runs keep T2
    symmetric for that reason."""
    Gai, Gab, T1A, d1, T2A, d2 = _system(5, sym=False)
    full = _ladder(Gab, Gai, T1A + d1, T2A + d2) - _ladder(Gab, Gai, T1A, T2A)
    low = _low_rank_change(Gab, Gai, T1A, d1, T2A, d2)
    assert np.abs(low - full).max() > 1e-6 * np.abs(full).max()
