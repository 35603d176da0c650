"""FP64 CPU restatement of the reference's direct-ring CCD (test infrastructure).

Follows src/algorithms/DrccdEnergyFromCoulombIntegrals.cxx:31-80 (residual)
on the driver of ClusterSinglesDoublesAlgorithm.cxx:38-128 with its energy
(cxx:131-213, `unrestricted` sets spins = 1) and amplitude update
(cxx:302-331); integrals PPHH V_abij, PHHP V_aijb and HHPP V_ijab as
CoulombIntegralsFromVertex.cxx:486-520 derives them from V_abij for a real
vertex (V_ijab = V_abij^(ijab), V_aijb = V_abij^(abji)).

Pinned by closed forms: the converged direct energy of the restricted
non-linearized equations is the direct-RPA correlation energy
1/2 Tr(Omega - A) of the Casida problem (A = Delta + 2K, B = 2K; spins 1:
A = Delta + K, B = K), the linearized equations are a Sylvester equation,
and the adjacent-pairs exchange term is checked against explicit loops.
"""

import numpy as np

from oracle import ccsd

ein = ccsd.ein


def integrals(Vabij):
    """(V_abij, V_aijb, V_ijab) of a real vertex (cxx:491, 515)."""
    return Vabij, ein("abji->aijb", Vabij), ein("abij->ijab", Vabij)


def residual(it, T2, Vabij, Vaijb, Vijab, *, spins=2.0, linearized=False,
             adjacent_pairs_exchange=False, initial_doubles=False):
    """DrccdEnergyFromCoulombIntegrals::getResiduum (cxx:31-80); the singles
    residual is zero.  Returns (R1, R2)."""
    Nv, No = Vabij.shape[0], Vabij.shape[2]
    R1 = np.zeros((Nv, No))
    R2 = Vabij.copy()                                                # cxx:57, 77
    if it == 0 and not initial_doubles:
        return R1, R2
    R2 += spins * ein("akic,cbkj->abij", Vaijb, T2)                  # cxx:58
    R2 += spins * ein("bkjc,acik->abij", Vaijb, T2)                  # cxx:59
    if not linearized:
        W = spins * Vijab                                            # cxx:61-62
        if adjacent_pairs_exchange:
            W = W - ein("jiab->ijab", Vijab)                         # cxx:63-65
        C = spins * ein("klcd,acik->alid", W, T2)                    # cxx:68
        R2 += ein("alid,dblj->abij", C, T2)                          # cxx:69
    return R1, R2


def energy(T2, Vabij, spins=2.0, Vijab=None):
    """getEnergy (cxx:131-190) with T1 = 0: direct 1/2 spins^2 T V_ijab,
    exchange -1/2 spins T V_ijba, V_ijab = HHPP (V_abij^(ijab) for a real
    vertex).  Returns (E, direct, exchange), complex for complex integrals."""
    if Vijab is None:
        Vijab = ein("abij->ijab", Vabij)
    direct = 0.5 * spins * spins * ein("abij,ijab->", T2, Vijab)
    exchange = -0.5 * spins * ein("abij,ijba->", T2, Vijab)
    return direct + exchange, direct, exchange


def solve(Vabij, eps_i, eps_a, *, spins=2.0, linearized=False, adjacent_pairs_exchange=False,
          mixer="DiisMixer", max_residua=4, mixing_ratio=1.0, max_iterations=16,
          amplitudes_convergence=1e-5, energy_convergence=1e-6, level_shift=0.0,
          complex_integrals=None):
    """ClusterSinglesDoublesAlgorithm::run (cxx:38-128) on the DRCCD residual;
    complex_integrals = (V_abij, V_aijb, V_ijab) of the complex path (F =
    complex: complex energies, |dE/E| of the complex energies)."""
    No, Nv = eps_i.shape[0], eps_a.shape[0]
    if complex_integrals is not None:
        V, Vaijb, Vijab = complex_integrals
    else:
        V, Vaijb, Vijab = integrals(Vabij)
    dt = complex if np.iscomplexobj(V) else float
    T = [np.zeros((Nv, No), dt), np.zeros((Nv, Nv, No, No), dt)]
    mix = ccsd.DiisMixer(max_residua) if mixer == "DiisMixer" else ccsd.LinearMixer(mixing_ratio)
    prev, energies, converged, i = 0.0, [], False, 0
    parts = (0.0, 0.0, 0.0)
    while i < max_iterations:
        R1, R2 = residual(i, T[1], V, Vaijb, Vijab, spins=spins, linearized=linearized,
                          adjacent_pairs_exchange=adjacent_pairs_exchange)
        E1, E2 = ccsd.estimate_amplitudes(R1, R2, T[0], T[1], eps_i, eps_a, level_shift)
        dT = [E1 - T[0], E2 - T[1]]
        mix.append([E1, E2], dT)
        T = mix.get()
        parts = energy(T[1], V, spins, Vijab)
        e = parts[0]
        ratio = ccsd.fock_dot(dT, dT) / ccsd.fock_dot(T, T)
        energies.append(e)
        i += 1
        if (abs((e - prev) / e) < abs(energy_convergence)
                and abs(ratio) < abs(amplitudes_convergence ** 2)):
            converged = True
            break
        prev = e
    return dict(energy=np.real(parts[0]), energy_complex=parts[0], direct=np.real(parts[1]),
                exchange=np.real(parts[2]), T2=T[1],
                energies=energies, iterations=i, converged=converged)
