"""Uniform electron gas vertex (test infrastructure: the oracle's restatement).

Restates src/algorithms/UegVertexGenerator.cxx:49-229 (grid, volume,
Madelung constant, Hartree-Fock eigenenergies, plane-wave vertex) and then
rewrites the plane-wave vertex in real orbitals, so that the real CCSD path
(which assumes Gamma[G,p,q] == Gamma[G,q,p], CcsdEnergyFromCoulombIntegrals.cxx:397-400)
applies to it:

* orbitals: each pair of plane waves (k, -k) becomes (cos, sin) =
  ((|k> + |-k>)/sqrt2, (|k> - |-k>)/(i sqrt2)); k = 0 stays;
* auxiliary index: each pair of momentum transfers (G, -G) becomes
  (sqrt2 Re M_G, sqrt2 Im M_G), where M_G is the plane-wave vertex of G in
  the real orbitals; G = 0 stays.

The rotation mixes only degenerate orbitals inside the hole block or inside
the particle block, so the Fock matrix stays diagonal and MP2, CCSD and (T)
energies are those of the plane-wave vertex.  Pinned by the reference's own
known answers for rs=1, No=7, Nv=26 (integration-tests/tests/cc4s/ueg/
rs1.0-7occ-26virt/cc4s.correct.out.yaml:128,153,169).

Within a shell of equal |k| the reference orders orbitals by std::sort on
|k|^2 alone (unspecified order); here ties break lexicographically on the
integer vector.  Energies do not depend on that order.
"""

import math

import numpy as np


def madelung(v):
    """UegVertexGenerator::evalMadelung (UegVertexGenerator.cxx:10-30): Ewald
    sum over a 13^3 block of image cells."""
    kappa = v ** (-1.0 / 3.0)
    term2 = math.pi / (kappa * kappa * v)
    term4 = 2.0 * kappa / math.sqrt(math.pi)
    box = 1.0 / kappa
    l = np.arange(-6, 7)
    n2 = (l[:, None, None] ** 2 + l[None, :, None] ** 2 + l[None, None, :] ** 2).ravel()
    n2 = n2[n2 > 0].astype(np.float64)
    modr = box * np.sqrt(n2)
    k2 = kappa * kappa * n2
    recip = -np.sum(1.0 / (math.pi * k2) * np.exp(-math.pi ** 2 * k2 / kappa / kappa) / v)
    real = -np.sum([math.erfc(kappa * r) / r for r in modr])
    return real + term2 + term4 + recip


def orbital_grid(No, Nv):
    """Integer momenta of the Np lowest plane waves (UegVertexGenerator.cxx:66-90)."""
    Np = No + Nv
    maxg = int((5.0 * Np) ** (1.0 / 3.0))
    r = np.arange(-maxg, maxg + 1)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    n2 = (g * g).sum(1)
    order = np.lexsort((g[:, 2], g[:, 1], g[:, 0], n2))
    g, n2 = g[order], n2[order]
    if len(g) < Np:
        raise ValueError("grid smaller than No + Nv")
    return g[:Np], n2


def plane_wave(No, Nv, rs, hartree_fock=True, madelung_constant=None):
    """Eigenenergies and the plane-wave vertex as the reference builds them.

    Returns dict(grid, eps, v, b, madelung, momenta, coeff) with
    coeff[G, p, r] = sqrt(w(G)) where G = g_r - g_p (UegVertexGenerator.cxx:212-226:
    the reference stores it at Gamma[G, r, p]); momenta lists the transfer
    vectors G in the order of the reference's momMap (cxx:151-160).
    """
    Np = No + Nv
    if No <= 0:
        raise ValueError("No larger zero please")
    if rs <= 0.0:
        raise ValueError("Invalid rs")
    g, n2 = orbital_grid(No, Nv)
    closed = n2[No] != n2[No - 1]
    if not closed and not hartree_fock:
        raise ValueError("Zero gap system! Either change No or use Hartree-Fock!")
    v = rs ** 3 / 3.0 * 4.0 * math.pi * No * 2                          # cxx:93
    b = 2.0 * math.pi / v ** (1.0 / 3.0)                                   # cxx:94-95
    mad = madelung(v) if madelung_constant is None or madelung_constant < 0 else madelung_constant
    k = b * g

    def w(q):                                                              # Vijji, cxx:35-39
        s = (q * q).sum(-1)
        return np.where(s < 1e-8, mad, 4.0 * math.pi / v / np.where(s < 1e-8, 1.0, s))

    eps = 0.5 * (k * k).sum(1)                                             # cxx:104-109
    if hartree_fock:
        eps = eps - np.array([w(k[p] - k[:No]).sum() for p in range(Np)])
    # momentum-transfer grid (cxx:133-160)
    # every transfer with |t|^2 <= max |g_r - g_p|^2; the reference bounds the
    # components by those of one longest transfer instead (cxx:145-150),
    # which is the same set whenever that transfer has the largest component
    d = g[None, :, :] - g[:, None, :]
    maxr = int((d * d).sum(-1).max())
    maxg = math.isqrt(maxr)
    r = np.arange(-maxg, maxg + 1)
    t = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    t = t[(t * t).sum(1) <= maxr]
    index = {tuple(x): n for n, x in enumerate(t)}
    NG = len(t)
    coeff = np.zeros((NG, Np, Np))
    for p in range(Np):
        for q in range(Np):
            dd = g[q] - g[p]
            s = int((dd * dd).sum())
            val = 4.0 * math.pi / v / (b * b * s) if s else mad               # cxx:221-222
            coeff[index[tuple(dd)], p, q] = math.sqrt(val)
    return dict(grid=g, eps=eps, v=v, b=b, madelung=mad, momenta=t, coeff=coeff)


def real_orbitals(grid, No):
    """Unitary U (rows: real orbitals) pairing |k> with |-k>.  Raises when a
    partner is missing or sits in the other block (open shells)."""
    Np = len(grid)
    where = {tuple(x): n for n, x in enumerate(grid)}
    U = np.zeros((Np, Np), complex)
    s2 = 1.0 / math.sqrt(2.0)
    for i, x in enumerate(grid):
        j = where.get(tuple(-x))
        if j is None or (i < No) != (j < No):
            raise ValueError("UEG real orbitals need closed shells at No and No + Nv")
        if j == i:
            U[i, i] = 1.0
        elif i < j:
            U[i, i] = U[i, j] = s2
            U[j, i], U[j, j] = -1j * s2, 1j * s2
    return U


def system(No, Nv, rs, hartree_fock=True, madelung_constant=None):
    """Real-orbital vertex Gamma[G, p, q] (symmetric in p, q) and eigenenergies.

    Auxiliary rows follow the reference's momMap order of transfers: one row
    for G = 0 and two, (cos G, sin G), for each G whose inverse comes later.
    """
    pw = plane_wave(No, Nv, rs, hartree_fock, madelung_constant)
    U = real_orbitals(pw["grid"], No)
    t = pw["momenta"]
    index = {tuple(x): n for n, x in enumerate(t)}
    M = np.einsum("Pp,Rr,Gpr->GPR", U.conj(), U, pw["coeff"], optimize=True)
    rows = []
    for n, x in enumerate(t):
        m = index[tuple(-x)]
        if m == n:
            rows.append(M[n].real)
        elif n < m:
            rows.append(math.sqrt(2.0) * M[n].real)
            rows.append(math.sqrt(2.0) * M[n].imag)
    gamma = np.ascontiguousarray(np.stack(rows))
    eps = pw["eps"]
    return gamma, eps[:No].copy(), eps[No:].copy(), pw
