"""FP64 CPU restatement of the reference's spin-orbital CCSD (test infrastructure).

UccsdAmplitudesFromCoulombIntegrals::getResiduumTemplate
(src/algorithms/UccsdAmplitudesFromCoulombIntegrals.cxx:107-241) returns
R = <Phi_i^a| Hbar |Phi> and <Phi_ij^ab| Hbar |Phi> of the similarity-
transformed Hamiltonian (SimilarityTransformedHamiltonian::getAI / getABIJ
with GENERAL dressing) minus the diagonal Fock terms (cxx:217-240); with
canonical orbitals (eigenenergies, no Fock matrices, cxx:121-134) that is
the right-hand side of the spin-orbital CCSD equations without diagonal
Fock terms.  It is restated here in the intermediate form of Stanton and
Gauss (J. Chem. Phys. 94, 4334 (1991)), which the reference's `intermediates`
option also follows; any exact form gives the same R.  Driver, update and
mixer: ClusterSinglesDoublesAlgorithm.cxx:38-128, 302-331; energy: the
antisymmetrized branch of getEnergy (cxx:163-167):
E = 1/4 T_abij V_ijab + 1/2 T_ai T_bj V_ijab.

Integrals are antisymmetrized spin-orbital blocks <pq||rs> in the
reference's class index orders: PPHH V_abij, HHHH V_ijkl, PPPP V_abcd,
HHHP V_ijka, PPPH V_abci, PHPH V_aibj (real, so <pq||rs> = <rs||pq>).

Pinned by closed forms: on a closed-shell system written in spin orbitals
the converged energy equals the restricted CCSD energy of the pinned real
path, and for two electrons with generic spin-orbital integrals it equals
the full-CI correlation energy from an exact diagonalisation.
"""

import numpy as np

from oracle import ccsd

ein = ccsd.ein


def blocks(ints):
    """Derived blocks <..||..> in Stanton's index orders from the six classes."""
    V, Vabci, Vijka, Vaibj = ints["PPHH"], ints["PPPH"], ints["HHHP"], ints["PHPH"]
    return dict(
        oovv=np.transpose(V, (2, 3, 0, 1)),            # <mn||ef> = V_efmn
        ovvv=-np.transpose(Vabci, (3, 2, 0, 1)),       # <ma||fe> = -V_fe,am
        vovv=np.transpose(Vabci, (2, 3, 0, 1)),        # <am||ef> = V_ef,am
        vvvo=Vabci,                                    # <ab||ej>
        ooov=Vijka,                                    # <mn||ie>
        ovoo=np.transpose(Vijka, (2, 3, 0, 1)),        # <mb||ij> = V_ij,mb
        ovvo=-np.transpose(Vaibj, (1, 0, 2, 3)),       # <mb||ej> = -V_bm,ej
        ovov_t1=np.transpose(Vaibj, (1, 0, 3, 2)),     # <na||if> = V_an,fi
        oooo=ints["HHHH"], vvvv=ints["PPPP"])


def _P(x, a, b):
    """x - x with axes a and b exchanged."""
    return x - np.swapaxes(x, a, b)


def residual(it, T1, T2, ints, initial_doubles=False):
    """Returns (R1[a,i], R2[a,b,i,j]); iteration 0 without initial doubles
    gives R2 = V_abij, R1 = 0 (cxx:147-162)."""
    V = ints["PPHH"]
    if it == 0 and not initial_doubles:
        return np.zeros_like(T1), V.copy()
    b = blocks(ints)
    t1 = T1.T                                   # t_ia
    t2 = np.transpose(T2, (2, 3, 0, 1))         # t_ijab
    o1 = ein("ia,jb->ijab", t1, t1)
    tau_t = t2 + 0.5 * (o1 - np.swapaxes(o1, 2, 3))
    tau = t2 + o1 - np.swapaxes(o1, 2, 3)
    oovv, ovvv, ooov = b["oovv"], b["ovvv"], b["ooov"]
    Fae = ein("mf,mafe->ae", t1, ovvv) - 0.5 * ein("mnaf,mnef->ae", tau_t, oovv)
    Fmi = ein("ne,mnie->mi", t1, ooov) + 0.5 * ein("inef,mnef->mi", tau_t, oovv)
    Fme = ein("nf,mnef->me", t1, oovv)
    Wmnij = (b["oooo"] + _P(ein("je,mnie->mnij", t1, ooov), 2, 3)
             + 0.25 * ein("ijef,mnef->mnij", tau, oovv))
    Wabef = (b["vvvv"] - _P(ein("mb,amef->abef", t1, b["vovv"]), 0, 1)
             + 0.25 * ein("mnab,mnef->abef", tau, oovv))
    mbef = -np.transpose(b["vovv"], (1, 0, 2, 3))   # <mb||ef> = -<bm||ef>
    mnej = -np.transpose(ooov, (0, 1, 3, 2))        # <mn||ej> = -<mn||je>
    Wmbej = (b["ovvo"] + ein("jf,mbef->mbej", t1, mbef) - ein("nb,mnej->mbej", t1, mnej)
             - ein("jnfb,mnef->mbej", 0.5 * t2 + ein("jf,nb->jnfb", t1, t1), oovv))
    # T1 equations
    nmei = -np.transpose(ooov, (0, 1, 3, 2))        # <nm||ei> = -<nm||ie>
    r1 = (ein("ie,ae->ia", t1, Fae) - ein("ma,mi->ia", t1, Fmi) + ein("imae,me->ia", t2, Fme)
          - ein("nf,naif->ia", t1, b["ovov_t1"]) - 0.5 * ein("imef,maef->ia", t2, -np.transpose(b["vovv"], (1, 0, 2, 3)))
          - 0.5 * ein("mnae,nmei->ia", t2, nmei))
    # T2 equations
    r2 = np.transpose(V, (2, 3, 0, 1)).copy()     # <ij||ab>
    r2 += _P(ein("ijae,be->ijab", t2, Fae - 0.5 * ein("mb,me->be", t1, Fme)), 2, 3)
    r2 -= _P(ein("imab,mj->ijab", t2, Fmi + 0.5 * ein("je,me->mj", t1, Fme)), 0, 1)
    r2 += 0.5 * ein("mnab,mnij->ijab", tau, Wmnij)
    r2 += 0.5 * ein("ijef,abef->ijab", tau, Wabef)
    x = ein("imae,mbej->ijab", t2, Wmbej) - ein("ie,ma,mbej->ijab", t1, t1, b["ovvo"])
    r2 += _P(_P(x, 0, 1), 2, 3)
    r2 += _P(ein("ie,abej->ijab", t1, b["vvvo"]), 0, 1)
    r2 -= _P(ein("ma,mbij->ijab", t1, b["ovoo"]), 2, 3)
    return r1.T.copy(), np.transpose(r2, (2, 3, 0, 1)).copy()


def energy(T1, T2, V):
    """Antisymmetrized getEnergy (cxx:163-167)."""
    return 0.25 * np.sum(T2 * V) + 0.5 * np.sum(ein("ai,bj->abij", T1, T1) * V)


def solve(ints, eps_i, eps_a, *, mixer="DiisMixer", max_residua=4, mixing_ratio=1.0,
          max_iterations=16, amplitudes_convergence=1e-5, energy_convergence=1e-6,
          level_shift=0.0):
    """ClusterSinglesDoublesAlgorithm::run (cxx:38-128) on the spin-orbital residual."""
    No, Nv = eps_i.shape[0], eps_a.shape[0]
    T = [np.zeros((Nv, No)), np.zeros((Nv, Nv, No, No))]
    mix = ccsd.DiisMixer(max_residua) if mixer == "DiisMixer" else ccsd.LinearMixer(mixing_ratio)
    prev, energies, converged, i, e = 0.0, [], False, 0, 0.0
    while i < max_iterations:
        R1, R2 = residual(i, T[0], T[1], ints)
        E1, E2 = ccsd.estimate_amplitudes(R1, R2, T[0], T[1], eps_i, eps_a, level_shift)
        dT = [E1 - T[0], E2 - T[1]]
        mix.append([E1, E2], dT)
        T = mix.get()
        e = energy(T[0], T[1], ints["PPHH"])
        ratio = ccsd.fock_dot(dT, dT) / ccsd.fock_dot(T, T)
        energies.append(e)
        i += 1
        if (abs((e - prev) / e) < abs(energy_convergence)
                and abs(ratio) < abs(amplitudes_convergence ** 2)):
            converged = True
            break
        prev = e
    return dict(energy=e, T1=T[0], T2=T[1], energies=energies, iterations=i,
                converged=converged)


def spin_orbital_integrals(g, No, Nv):
    """Antisymmetrized spin-orbital classes of a closed-shell real vertex
    Gamma[G,P,Q] (spatial); spin orbitals: holes (I, alpha), (I, beta), then
    particles likewise, index = 2 * spatial + spin within each block.
    Returns (ints, spin-orbital index -> spatial map helpers)."""
    Np = No + Nv
    so = [(P, s) for P in range(No) for s in (0, 1)] + [(P, s) for P in range(No, Np) for s in (0, 1)]
    sp = np.array([p for p, _ in so])
    spin = np.array([s for _, s in so])
    eri = ein("Gpr,Gqs->pqrs", g, g)            # <PQ|RS> = (PR|QS), spatial
    full = eri[np.ix_(sp, sp, sp, sp)]
    d = spin[:, None] == spin[None, :]
    full = full * (d[:, None, :, None] & d[None, :, None, :])
    asym = full - np.swapaxes(full, 2, 3)       # <pq||rs>
    return classes(asym, 2 * No)


def classes(asym, no):
    """The six classes from a full antisymmetrized tensor <pq||rs> whose
    first `no` orbitals are holes."""
    h, p = slice(0, no), slice(no, asym.shape[0])
    return dict(PPHH=asym[p, p, h, h].copy(), HHHH=asym[h, h, h, h].copy(),
                PPPP=asym[p, p, p, p].copy(), HHHP=asym[h, h, h, p].copy(),
                PPPH=asym[p, p, p, h].copy(), PHPH=asym[p, h, p, h].copy())
