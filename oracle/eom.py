"""FP64 CPU restatement of the reference's IP-EOM-CCSD right-apply (test infrastructure).

UCcsdIPEquationOfMotionDavidson::run (src/algorithms/UCcsdIPEquationOfMotionDavidson.cxx:
43-205) diagonalises the CCSD similarity-transformed Hamiltonian in the
ionised (1h + 2h1p) space with SimilarityTransformedHamiltonian::
right_apply_CCSD_IP (SimilarityTransformedHamiltonian.cxx:64-68, CCSD
dressing, spin orbitals, canonical orbitals).  The right-apply is restated in
the intermediate form of Stanton and Gauss (J. Chem. Phys. 101, 8938 (1994));
any exact form has the same eigenvalues.  Integrals: the six antisymmetrized
classes of oracle.uccsd.

Pinned by a closed form: for two electrons CCSD and IP-EOM-CCSD are exact,
so the eigenvalues equal E(N-1) - E(N) from full CI (the one-electron
Hamiltonian's eigenvalues minus the two-electron ground state from
Slater-Condon rules); a sign error in one ring term of W_mbij moves them by
2e-3, so the check detects such mistakes.
"""

import numpy as np

from oracle import ccsd, uccsd

ein = ccsd.ein


def hbar_ip(ints, eps_i, eps_a, T1, T2):
    """One- and two-body blocks of the CCSD similarity-transformed Hamiltonian
    that the IP right-apply needs (Foo, Fvv, Fov, Woooo, Wovvo, Wooov, Wovoo),
    in Stanton's index orders (t_ia, t_ijab), canonical orbitals."""
    b = uccsd.blocks(ints)
    t1 = T1.T
    t2 = np.transpose(T2, (2, 3, 0, 1))
    oovv, ovvv, ooov = b["oovv"], b["ovvv"], b["ooov"]
    oooo, ovvo, ovoo = b["oooo"], b["ovvo"], b["ovoo"]
    fo, fv = np.diag(eps_i), np.diag(eps_a)
    o1 = ein("ia,jb->ijab", t1, t1)
    tau_t = t2 + 0.5 * (o1 - o1.transpose(0, 1, 3, 2))
    tau = t2 + o1 - o1.transpose(0, 1, 3, 2)
    Fme = ein("nf,mnef->me", t1, oovv)
    Fmi = fo + ein("ne,mnie->mi", t1, ooov) + 0.5 * ein("inef,mnef->mi", tau_t, oovv)
    Fae = fv + ein("mf,mafe->ae", t1, ovvv) - 0.5 * ein("mnaf,mnef->ae", tau_t, oovv)
    # Hbar one-body
    Foo = Fmi + 0.5 * ein("ie,me->mi", t1, Fme)
    Fvv = Fae - 0.5 * ein("ma,me->ae", t1, Fme)
    Fov = Fme
    # Hbar two-body
    Woooo = (oooo + (ein("je,mnie->mnij", t1, ooov) - ein("ie,mnje->mnij", t1, ooov))
             + 0.25 * ein("ijef,mnef->mnij", tau, oovv)) + 0.25 * ein("ijef,mnef->mnij", tau, oovv)
    mbef = -np.transpose(b["vovv"], (1, 0, 2, 3))
    mnej = -np.transpose(ooov, (0, 1, 3, 2))
    Wmbej = (ovvo + ein("jf,mbef->mbej", t1, mbef) - ein("nb,mnej->mbej", t1, mnej)
             - ein("jnfb,mnef->mbej", 0.5 * t2 + ein("jf,nb->jnfb", t1, t1), oovv))
    Wovvo = Wmbej - 0.5 * ein("jnfb,mnef->mbej", t2, oovv)
    Wooov = ooov + ein("if,mnfe->mnie", t1, oovv)
    # Wmbij
    Wovoo = (ovoo - ein("me,ijbe->mbij", Fov, t2) - ein("nb,mnij->mbij", t1, Woooo)
             + 0.5 * ein("mbef,ijef->mbij", mbef, tau))
    tmp = ein("mnie,jnbe->mbij", ooov, t2)
    Wovoo += tmp - tmp.transpose(0, 1, 3, 2)
    tmp = ein("ie,mbej->mbij", t1, ovvo) - ein("ie,njbf,mnef->mbij", t1, t2, oovv)
    Wovoo += tmp - tmp.transpose(0, 1, 3, 2)
    return dict(Foo=Foo, Fvv=Fvv, Fov=Fov, Woooo=Woooo, Wovvo=Wovvo, Wooov=Wooov, Wovoo=Wovoo,
                oovv=oovv, t2=t2)


def ip_sigma(H, r1, r2):
    """Hbar R on the IP space: r1[i] (1h), r2[i,j,a] (2h1p, antisymmetric in
    i, j); returns (sigma1, sigma2)."""
    Foo, Fvv, Fov = H["Foo"], H["Fvv"], H["Fov"]
    Hr1 = -ein("mi,m->i", Foo, r1)
    Hr1 += ein("me,mie->i", Fov, r2)
    Hr1 += -0.5 * ein("nmie,mne->i", H["Wooov"], r2)
    Hr2 = -ein("maji,m->ija", H["Wovoo"], r1)
    Hr2 += ein("ae,ije->ija", Fvv, r2)
    tmp1 = ein("mi,mja->ija", Foo, r2)
    Hr2 -= tmp1 - tmp1.transpose(1, 0, 2)
    Hr2 += 0.5 * ein("mnij,mna->ija", H["Woooo"], r2)
    tmp2 = ein("maei,mje->ija", H["Wovvo"], r2)
    Hr2 += tmp2 - tmp2.transpose(1, 0, 2)
    Hr2 += 0.5 * ein("mnef,mnf,ijae->ija", H["oovv"], r2, H["t2"])
    return Hr1, Hr2


def ip_matrix(H, no, nv):
    """Dense Hbar in the IP space (1h, then the 2h1p pairs i < j) for small
    systems, column by column from ip_sigma."""
    pairs = [(i, j) for i in range(no) for j in range(i + 1, no)]
    dim = no + len(pairs) * nv
    M = np.zeros((dim, dim))
    def unpack(x):
        r1 = x[:no].copy()
        r2 = np.zeros((no, no, nv))
        k = no
        for (i, j) in pairs:
            r2[i, j] = x[k:k + nv]
            r2[j, i] = -x[k:k + nv]
            k += nv
        return r1, r2
    def pack(r1, r2):
        out = [r1]
        for (i, j) in pairs:
            out.append(r2[i, j])
        return np.concatenate(out)
    for c in range(dim):
        x = np.zeros(dim); x[c] = 1.0
        r1, r2 = unpack(x)
        s1, s2 = ip_sigma(H, r1, r2)
        M[:, c] = pack(s1, s2)
    return M


def ip_eigenvalues(ints, eps_i, eps_a, T1, T2):
    """All IP-EOM-CCSD eigenvalues (real parts, ascending) by dense
    diagonalisation; small systems only."""
    H = hbar_ip(ints, eps_i, eps_a, T1, T2)
    M = ip_matrix(H, len(eps_i), len(eps_a))
    return np.sort(np.linalg.eigvals(M).real)
