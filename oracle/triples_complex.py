"""FP64 CPU restatement of CcsdPerturbativeTriplesComplex (test infrastructure).

Follows src/algorithms/CcsdPerturbativeTriplesComplex.cxx:123-279 (the
Calculator<complex> branch) literally: the i0 <= i1 <= i2 loop, the six
simultaneous permutations with the distinct-permutation bookkeeping, the
spin/Fermi factors {+2, -4, 0, +8}, DV -> conj(DV / D), and
E += sum_abc DV Tabc.  Inputs are the complex CCSD amplitudes, the PPHH and
PHHH integrals of calculateComplexIntegrals and the vertex slices Gamma_ab
(conjugated elementwise, cxx:343-350) and Gamma_ai.
"""

import numpy as np

from .ccsd import ein
from .ccsd_complex import integrals_from_vertex
from .triples import (SPIN_FERMI, invariant_count, map_times_perm, permutation3,
                      sorted_blocks, str_times_perm)


class Inputs:
    def __init__(self, T1, T2, gamma, eps_i, eps_a):
        No, Nv = eps_i.shape[0], eps_a.shape[0]
        ints = integrals_from_vertex(gamma, No)
        self.T1, self.T2 = T1, T2
        self.Vabij, self.Valij = ints["PPHH"], ints["PHHH"]
        self.conjGab = gamma[:, No:, No:].conj()
        self.Gai = gamma[:, No:, :No]
        self.eps_i, self.eps_a = eps_i, eps_a
        self.No, self.Nv = No, Nv

    def singles(self, i):
        """getSinglesContribution (cxx:165-170): 0.5 T_a,i0 V_bc,i1i2."""
        return 0.5 * ein("a,bc->abc", self.T1[:, i[0]], self.Vabij[:, :, i[1], i[2]])

    def doubles(self, i):
        """getDoublesParticleContribution<complex> (cxx:370-376) and
        addDoublesHoleContribution (cxx:157-162)."""
        W = ein("ad,Fdb,Fc->abc", self.T2[:, :, i[0], i[1]], self.conjGab, self.Gai[:, :, i[2]])
        W -= ein("abl,cl->abc", self.T2[:, :, i[0], :], self.Valij[:, :, i[2], i[1]])
        return W

    def denominator(self, i):
        e = self.eps_i[i[0]] + self.eps_i[i[1]] + self.eps_i[i[2]]
        ea = self.eps_a
        return e - ea[:, None, None] - ea[None, :, None] - ea[None, None, :]


def block_energy(inp, i):
    """One sorted block (cxx:199-258); complex."""
    perms = [permutation3(p) for p in range(6)]
    piDV = [None] * 6
    distinct = [False] * 6
    DV = np.zeros((inp.Nv,) * 3, complex)
    for p in range(6):
        pi = perms[p]
        q = 0
        while q < p:
            if map_times_perm(i, perms[q]) == map_times_perm(i, pi):
                break
            q += 1
        if q < p:
            piDV[p] = piDV[q]
        else:
            distinct[p] = True
            piDV[p] = inp.doubles(map_times_perm(i, pi))
        DV += ein(str_times_perm("abc", pi) + "->abc", piDV[p])
    DV = np.conj(DV / inp.denominator(i))                     # cxx:226-230
    e = 0.0j
    for p in range(6):
        if not distinct[p]:
            continue
        pi = perms[p]
        S = inp.singles(map_times_perm(i, pi))
        Tabc = np.zeros_like(DV)
        for s in range(6):
            sigma = perms[s]
            sf = SPIN_FERMI[invariant_count(sigma)]
            idx = str_times_perm(str_times_perm("abc", sigma), pi)
            Tabc += sf * ein(idx + "->abc", piDV[p])
            Tabc += sf * ein(idx + "->abc", S)
        e += np.sum(DV * Tabc)
    return e


def triples_energy(T1, T2, gamma, eps_i, eps_a, blocks=None):
    """E_(T), complex; the reference adds its real part to the CCSD energy
    (cxx:73-84)."""
    inp = Inputs(T1, T2, gamma, eps_i, eps_a)
    if blocks is None:
        blocks = sorted_blocks(inp.No)
    return sum(block_energy(inp, b) for b in blocks)
