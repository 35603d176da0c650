"""FP64 CPU restatement of CcsdPerturbativeTriples (test infrastructure).

Follows src/algorithms/CcsdPerturbativeTriples.cxx:22-248 and
src/math/Permutation.hpp:52-100 literally: the i0 <= i1 <= i2 slice loop, the
six simultaneous permutations, the "distinct permutation" bookkeeping and the
spin/Fermi factors {+2, -4, 0, +8} indexed by the number of fixed points.
CTF index-string reads ``X[s]`` with ``s`` a permutation of "abc" are
restated with ``einsum(s + '->abc', X)`` (same label-matching semantics).
"""

import numpy as np

from .ccsd import ein, integrals_from_vertex


def permutation3(p):
    """Permutation<3>(p) (Permutation.hpp:52-62), recursive construction."""
    def perm(n, p):
        if n == 1:
            return [0]
        sub = perm(n - 1, p // n)
        images = [0] * n
        pos = p % n
        for i in range(pos):
            images[i] = sub[i] + 1
        images[pos] = 0
        for i in range(pos + 1, n):
            images[i] = sub[i - 1] + 1
        return images
    return perm(3, p)


def invariant_count(pi):
    return sum(1 for i in range(3) if pi[i] == i)


def str_times_perm(s, pi):
    """operator*(string, Permutation) (CcsdPerturbativeTriples.cxx:22-29)."""
    return "".join(s[pi[i]] for i in range(3))


def map_times_perm(m, pi):
    """operator*(Map, Permutation) (Permutation.hpp:93-98): g(k) = f(pi(k))."""
    return tuple(m[pi[k]] for k in range(3))


SPIN_FERMI = (2.0, -4.0, 0.0, 8.0)          # cxx:143


class TriplesInputs:
    """sliceTensors (cxx:32-79): the slices (T) reads, as plain arrays."""

    def __init__(self, T1, T2, gamma, eps_i, eps_a):
        No, Nv = eps_i.shape[0], eps_a.shape[0]
        ints = integrals_from_vertex(gamma, No)
        self.T1, self.T2 = T1, T2
        self.Vabij, self.Vijka = ints["PPHH"], ints["HHHP"]
        self.Gab = gamma[:, No:, No:]
        self.Gai = gamma[:, No:, :No]
        self.eps_i, self.eps_a = eps_i, eps_a
        self.No, self.Nv = No, Nv

    def singles(self, i):
        """getSinglesContribution (cxx:81-85): 0.5 T_a,i0 V_bc,i1i2."""
        return 0.5 * ein("a,bc->abc", self.T1[:, i[0]], self.Vabij[:, :, i[1], i[2]])

    def doubles(self, i):
        """getDoublesContribution (cxx:87-96)."""
        W = ein("ad,Fbd,Fc->abc", self.T2[:, :, i[0], i[1]], self.Gab, self.Gai[:, :, i[2]])
        W -= ein("abl,lc->abc", self.T2[:, :, i[0], :], self.Vijka[i[1], i[2], :, :])
        return W

    def denominator(self, i):
        """getEnergyDenominator (cxx:98-117)."""
        e = self.eps_i[i[0]] + self.eps_i[i[1]] + self.eps_i[i[2]]
        ea = self.eps_a
        return e - ea[:, None, None] - ea[None, :, None] - ea[None, None, :]


def block_energy(inp, i):
    """Energy contribution of one sorted block i = (i0 <= i1 <= i2)
    (cxx:159-215)."""
    perms = [permutation3(p) for p in range(6)]
    piDV = [None] * 6
    distinct = [False] * 6
    DV = np.zeros((inp.Nv,) * 3)
    for p in range(6):
        pi = perms[p]
        q = 0
        while q < p:
            if map_times_perm(i, perms[q]) == map_times_perm(i, pi):
                break
            q += 1
        if q < p:
            distinct[p] = False
            piDV[p] = piDV[q]
        else:
            distinct[p] = True
            piDV[p] = inp.doubles(map_times_perm(i, pi))
        DV += ein(str_times_perm("abc", pi) + "->abc", piDV[p])
    DV /= inp.denominator(i)
    e = 0.0
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
        e += float(np.sum(DV * Tabc))
    return e


def sorted_blocks(No):
    return [(i0, i1, i2) for i0 in range(No) for i1 in range(i0, No) for i2 in range(i1, No)]


def triples_energy(T1, T2, gamma, eps_i, eps_a, blocks=None):
    """CcsdPerturbativeTriples::run (cxx:119-248) without the CCSD energy:
    returns E_(T).  ``blocks`` restricts the loop to a subset (for sampling)."""
    inp = TriplesInputs(T1, T2, gamma, eps_i, eps_a)
    if blocks is None:
        blocks = sorted_blocks(inp.No)
    return sum(block_energy(inp, b) for b in blocks)
