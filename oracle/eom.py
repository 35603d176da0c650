"""FP64 CPU restatement of the reference's IP- and EA-EOM-CCSD right-applies (test
infrastructure).

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
2e-3, so the check detects such mistakes.  EA (hbar_ea, ea_sigma) is pinned
the same way: with two virtual spin orbitals CCSD is exact for any electron
count and the 1p + 2p1h space holds every (N+1)-electron determinant, so the
eigenvalues equal E(N+1) - E(N) from determinant full CI (to 4e-15; scaling
any one block of Hbar by 1 % moves them by 2e-5 or more).
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


# --- EA-EOM-CCSD (UCcsdEAEquationOfMotionDavidson) ---------------------------

def hbar_ea(ints, eps_i, eps_a, T1, T2):
    """Blocks of the CCSD similarity-transformed Hamiltonian that the EA
    right-apply needs (Foo, Fvv, Fov, Wovvo as for IP; Wvvvv, Wvovv, Wvvvo),
    Stanton's index orders, canonical orbitals.  The EA right-apply of the
    reference is SimilarityTransformedHamiltonian::right_apply_CCSD_EA, driven
    by UCcsdEAEquationOfMotionDavidson::run (UCcsdEAEquationOfMotionDavidson.cxx);
    restated in the intermediate form of Stanton and Gauss (J. Chem. Phys. 103,
    1064 (1995)), the particle-hole counterpart of hbar_ip."""
    H = hbar_ip(ints, eps_i, eps_a, T1, T2)
    b = uccsd.blocks(ints)
    t1 = T1.T
    t2 = H["t2"]
    oovv, ovvv, ooov, ovvo, vvvv = b["oovv"], b["ovvv"], b["ooov"], b["ovvo"], b["vvvv"]
    o1 = ein("ia,jb->ijab", t1, t1)
    tau = t2 + o1 - o1.transpose(0, 1, 3, 2)
    # W_abef = <ab||ef> - P(ab) t_mb <am||ef> + tau_mnab <mn||ef> / 2
    Wvvvv = (vvvv - ein("ma,mbef->abef", t1, ovvv) + ein("mb,maef->abef", t1, ovvv)
             + 0.5 * ein("mnab,mnef->abef", tau, oovv))
    # W_amef = <am||ef> - t_na <nm||ef>
    Wvovv = b["vovv"] - ein("na,nmef->amef", t1, oovv)
    # W_abei = <ab||ei> - F_me t_miab + t_if W_abef + <mn||ei> tau_mnab / 2
    #          - P(ab) <mb||ef> t_miaf - P(ab) t_ma (<mb||ei> - t_nibf <mn||ef>)
    mnei = -np.transpose(ooov, (0, 1, 3, 2))
    Wvvvo = (b["vvvo"] - ein("me,miab->abei", H["Fov"], t2) + ein("if,abef->abei", t1, Wvvvv)
             + 0.5 * ein("mnei,mnab->abei", mnei, tau))
    tmp = ein("mbef,miaf->abei", ovvv, t2)
    Wvvvo -= tmp - tmp.transpose(1, 0, 2, 3)
    tmp = ein("ma,mbei->abei", t1, ovvo) - ein("ma,nibf,mnef->abei", t1, t2, oovv)
    Wvvvo -= tmp - tmp.transpose(1, 0, 2, 3)
    H.update(Wvvvv=Wvvvv, Wvovv=Wvovv, Wvvvo=Wvvvo)
    return H


def ea_sigma(H, r1, r2):
    """Hbar R on the EA space: r1[a] (1p), r2[j,a,b] (2p1h, antisymmetric in
    a, b); returns (sigma1, sigma2)."""
    Fvv, Foo, Fov = H["Fvv"], H["Foo"], H["Fov"]
    Hr1 = ein("ac,c->a", Fvv, r1)
    Hr1 += ein("ld,lad->a", Fov, r2)
    Hr1 += 0.5 * ein("alcd,lcd->a", H["Wvovv"], r2)
    Hr2 = ein("abcj,c->jab", H["Wvvvo"], r1)
    tmp1 = ein("ac,jcb->jab", Fvv, r2)
    Hr2 += tmp1 - tmp1.transpose(0, 2, 1)
    Hr2 -= ein("lj,lab->jab", Foo, r2)
    tmp2 = ein("lbdj,lad->jab", H["Wovvo"], r2)
    Hr2 += tmp2 - tmp2.transpose(0, 2, 1)
    Hr2 += 0.5 * ein("abcd,jcd->jab", H["Wvvvv"], r2)
    Hr2 -= 0.5 * ein("klcd,lcd,kjab->jab", H["oovv"], r2, H["t2"])
    return Hr1, Hr2


def ea_matrix(H, no, nv):
    """Dense Hbar in the EA space (1p, then the 2p1h entries j, a < b)."""
    pairs = [(a, b) for a in range(nv) for b in range(a + 1, nv)]
    dim = nv + no * len(pairs)
    M = np.zeros((dim, dim))
    def unpack(x):
        r1 = x[:nv].copy()
        r2 = np.zeros((no, nv, nv))
        k = nv
        for j in range(no):
            for (a, b) in pairs:
                r2[j, a, b], r2[j, b, a] = x[k], -x[k]
                k += 1
        return r1, r2
    def pack(r1, r2):
        return np.concatenate([r1] + [np.array([r2[j, a, b] for (a, b) in pairs]) for j in range(no)])
    for c in range(dim):
        x = np.zeros(dim); x[c] = 1.0
        s1, s2 = ea_sigma(H, *unpack(x))
        M[:, c] = pack(s1, s2)
    return M


def ea_eigenvalues(ints, eps_i, eps_a, T1, T2):
    """All EA-EOM-CCSD eigenvalues (real parts, ascending); small systems."""
    H = hbar_ea(ints, eps_i, eps_a, T1, T2)
    return np.sort(np.linalg.eigvals(ea_matrix(H, len(eps_i), len(eps_a))).real)
