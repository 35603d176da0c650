"""Philox-4x64-10 counter-based generator and the synthetic CCSD inputs.

Oracle restatement (test infrastructure only) of the product's input
generator.  The reference generates its random tensors with mt19937 seeded per
MPI rank (math/RandomTensor.hpp:38-58), which is not reproducible across
distributions; the workloads here use a counter-based generator instead so
that every element is a pure function of (seed, indices):

* Philox-4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11), the same bijection as
  ``numpy.random.Philox`` (pinned against it in the tests).
* Gamma[G,p,q] = Gamma[G,q,p] = z(G,p,q) * exp(-G/100) / sqrt(NG) for p <= q,
  z ~ N(0,1) by Box-Muller from the first two 64-bit words of
  philox(ctr=(G, p, q, 0), key=(seed, 0)).
* eps_i = -2.0 + 1.5 u,  u from philox(ctr=(i, 0, 0, 1), key=(seed, 0)).
* eps_a =  0.2 + 2.8 u,  u from philox(ctr=(a, 0, 0, 2), key=(seed, 0)).
* u = (word >> 11) * 2^-53 in [0, 1).
"""

import numpy as np

M0 = np.uint64(0xD2E7470EE14C6C93)
M1 = np.uint64(0xCA5A826395121157)
W0 = np.uint64(0x9E3779B97F4A7C15)
W1 = np.uint64(0xBB67AE8584CAA73B)
_MASK32 = np.uint64(0xFFFFFFFF)
_S32 = np.uint64(32)


def _mulhilo(a, b):
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    a0, a1 = a & _MASK32, a >> _S32
    b0, b1 = b & _MASK32, b >> _S32
    p00, p01, p10, p11 = a0 * b0, a0 * b1, a1 * b0, a1 * b1
    mid = (p00 >> _S32) + (p01 & _MASK32) + (p10 & _MASK32)
    hi = p11 + (p01 >> _S32) + (p10 >> _S32) + (mid >> _S32)
    with np.errstate(over="ignore"):
        lo = a * b
    return lo, hi


def philox4x64(ctr, key, rounds=10):
    """Vectorised Philox-4x64-R.  ctr: 4 uint64 arrays, key: 2 uint64 arrays."""
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint64) for c in ctr)
    k0, k1 = (np.asarray(k, dtype=np.uint64) for k in key)
    with np.errstate(over="ignore"):
        for r in range(rounds):
            if r > 0:
                k0 = k0 + W0
                k1 = k1 + W1
            lo0, hi0 = _mulhilo(M0, c0)
            lo1, hi1 = _mulhilo(M1, c2)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
    return c0, c1, c2, c3


def _uniform(word):
    return (np.asarray(word, dtype=np.uint64) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def coulomb_vertex(seed, NG, Np):
    """Real symmetric vertex Gamma[G,p,q] (logical order), float64."""
    G, p, q = np.meshgrid(np.arange(NG, dtype=np.uint64),
                          np.arange(Np, dtype=np.uint64),
                          np.arange(Np, dtype=np.uint64), indexing="ij")
    lo = np.minimum(p, q)
    hi = np.maximum(p, q)
    zero = np.zeros_like(G)
    key = (np.uint64(seed), np.uint64(0))
    r0, r1, _, _ = philox4x64((G, lo, hi, zero), key)
    u1 = _uniform(r0)
    u2 = _uniform(r1)
    z = np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(2.0 * np.pi * u2)
    scale = np.exp(-np.arange(NG, dtype=np.float64) / 100.0) / np.sqrt(float(NG))
    return z * scale[:, None, None]


def eigen_energies(seed, No, Nv):
    """(eps_i [No], eps_a [Nv]) with the distributions stated above."""
    key = (np.uint64(seed), np.uint64(0))
    i = np.arange(No, dtype=np.uint64)
    a = np.arange(Nv, dtype=np.uint64)
    zi = np.zeros_like(i)
    za = np.zeros_like(a)
    ri, _, _, _ = philox4x64((i, zi, zi, zi + np.uint64(1)), key)
    ra, _, _, _ = philox4x64((a, za, za, za + np.uint64(2)), key)
    return -2.0 + 1.5 * _uniform(ri), 0.2 + 2.8 * _uniform(ra)


def system(seed, No, Nv, NG):
    """Gamma, eps_i, eps_a for one synthetic system."""
    gamma = coulomb_vertex(seed, NG, No + Nv)
    eps_i, eps_a = eigen_energies(seed, No, Nv)
    return gamma, eps_i, eps_a
