"""FP64 CPU restatement of the reference's real CCSD path (test infrastructure).

Follows, term by term:

* src/algorithms/CoulombIntegralsFromVertex.cxx:23-148, 203-571 (real integrals)
* src/algorithms/CcsdEnergyFromCoulombIntegrals.cxx:27-494 (real residual)
* src/algorithms/ClusterSinglesDoublesAlgorithm.cxx:38-128 (driver),
  131-237 (energy), 302-365 (update), 887-915 (sliceIntoResiduum)
* src/mixers/DiisMixer.cxx:16-180, src/mixers/LinearMixer.cxx:17-55
* src/math/FockVector.hpp:136-236 (+=, -=, *=, dot)

The reference reads the vertex as complex and adds Im*Im partner products;
the vertex here is real, so those terms are identically zero and omitted.
The reference's real path assumes Gamma[G,k,b] == Gamma[G,b,k]
(CcsdEnergyFromCoulombIntegrals.cxx:397-400); the synthetic vertex is
symmetric in (p,q), so the restatement uses the same slices it does.
"""

import numpy as np

_EIN = dict(optimize=True)


def ein(spec, *ops):
    return np.einsum(spec, *ops, **_EIN)


# ---------------------------------------------------------------------------
# Vertex slices and integrals
# ---------------------------------------------------------------------------

def vertex_slices(gamma, No):
    """Gamma_ij, Gamma_ai, Gamma_ab (CcsdEnergyFromCoulombIntegrals.cxx:66-78;
    CoulombIntegralsFromVertex.cxx:121-136: a = No..Np-1, i = 0..No-1)."""
    Gij = gamma[:, :No, :No]
    Gai = gamma[:, No:, :No]
    Gab = gamma[:, No:, No:]
    return Gij, Gai, Gab


def integrals_from_vertex(gamma, No):
    """Real Coulomb integrals (CoulombIntegralsFromVertex.cxx:393-423, 485-496).

    Returns dict with PPHH V_abij, PHPH V_aibj, HHHH V_ijkl, HHHP V_ijka and
    HHPP V_ijab (= V_abij transposed).
    """
    Gij, Gai, Gab = vertex_slices(gamma, No)
    Vaibj = ein("Gab,Gij->aibj", Gab, Gij)          # cxx:395
    Vabij = ein("Gai,Gbj->abij", Gai, Gai)          # cxx:402
    Vijkl = ein("Gik,Gjl->ijkl", Gij, Gij)          # cxx:409
    Vijka = ein("Gik,Gaj->ijka", Gij, Gai)          # cxx:416
    Vijab = ein("abij->ijab", Vabij)                # cxx:485-496
    return dict(PPHH=Vabij, PHPH=Vaibj, HHHH=Vijkl, HHHP=Vijka, HHPP=Vijab)


def vabcd_from_vertex(gamma, No):
    """PPPP and PPPH classes (CoulombIntegralsFromVertex.cxx:423, 430); only
    used by the full-integral cross-formulation on tiny shapes."""
    _, Gai, Gab = vertex_slices(gamma, No)
    Vabcd = ein("Gac,Gbd->abcd", Gab, Gab)
    Vabci = ein("Gac,Gbi->abci", Gab, Gai)
    return Vabcd, Vabci


# ---------------------------------------------------------------------------
# Residual
# ---------------------------------------------------------------------------

def slice_bounds(Nv, slice_size):
    """Virtual slices of the ladder (CcsdEnergyFromCoulombIntegrals.cxx:374-416)."""
    n = int(np.ceil(Nv / slice_size))
    return [(v * slice_size, min((v + 1) * slice_size, Nv)) for v in range(n)]


def slice_into_residuum(Rxyij, a, b, R2):
    """ClusterSinglesDoublesAlgorithm.cxx:887-915: add the slice at (a,b,i,j)
    and, if a > b, its (y,x,j,i) transpose at (b,a)."""
    Nx, Ny = Rxyij.shape[0], Rxyij.shape[1]
    R2[a:a + Nx, b:b + Ny] += Rxyij
    if a > b:
        R2[b:b + Ny, a:a + Nx] += ein("xyij->yxji", Rxyij)


def ppl_ladder(dGab, X, slice_size, R2):
    """Particle-particle ladder from the dressed vertex, slice pair by slice
    pair (CcsdEnergyFromCoulombIntegrals.cxx:418-451)."""
    Nv = dGab.shape[1]
    bounds = slice_bounds(Nv, slice_size)
    for m in range(len(bounds)):
        for n in range(m, len(bounds)):
            xs, xe = bounds[n]
            ys, ye = bounds[m]
            Vxycd = ein("Gxc,Gyd->xycd", dGab[:, xs:xe, :], dGab[:, ys:ye, :])
            Rxyij = ein("xycd,cdij->xyij", Vxycd, X)
            slice_into_residuum(Rxyij, xs, ys, R2)


def factor_ladder(T1, X, Lambda, Pi, slice_size, R2):
    """sliceAmplitudesFromCoupledCoulombFactors for real amplitudes
    (ClusterSinglesDoublesAlgorithm.cxx:567-672) over all pairs of factor
    slices (CcsdEnergyFromCoulombIntegrals.cxx:357-369): Lambda[G,R] and
    Pi[p,R] complex, X = T2 + T1 T1.  Equals the vertex ladder when
    Gamma[G,p,q] = sum_R Lambda[G,R] Pi[p,R] Pi[q,R]."""
    Nv, No = T1.shape
    NR = Pi.shape[1]
    PiaR, PiiR = Pi[No:, :], Pi[:No, :]
    dPi = PiaR - ein("kR,ak->aR", PiiR, T1)                    # cxx:639-650
    for b in range(0, NR, slice_size):
        for a in range(0, NR, slice_size):
            L, Rr = slice(a, min(a + slice_size, NR)), slice(b, min(b + slice_size, NR))
            XRaij = ein("cdij,cR->Rdij", X, PiaR[:, L].conj())           # cxx:607-610
            XRSij = ein("Rdij,dS->RSij", XRaij, PiaR[:, Rr])            # cxx:613-614
            VRS = ein("GR,GS->RS", Lambda[:, L].conj(), Lambda[:, Rr])  # cxx:617-621
            XRSij = XRSij * VRS[:, :, None, None]                        # cxx:623
            XRbij = ein("RSij,bS->Rbij", XRSij, dPi[:, Rr])              # cxx:660
            R2 += np.real(ein("Rbij,aR->abij", XRbij, dPi[:, L].conj()))  # cxx:666-668


def residual(it, T1, T2, gamma, ints, slice_size=None, initial_doubles=False,
             ppl=True, distinguishable=False, factors=None):
    """CcsdEnergyFromCoulombIntegrals::getResiduum, real path (cxx:27-494),
    no CoulombFactors.  distinguishable (cxx:123-124) selects the
    distinguishable-cluster terms (cxx:150-160, 176-186, 244-246, 283-285,
    316-320); ppl = False (cxx:324-333) skips the particle-particle ladder
    and contracts the last hole-ladder term with T2 + T1 T1.  Returns
    (R1, R2).  factors = (Lambda, Pi, factors_slice_size) selects the
    factor-based ladder (cxx:340-372) in place of the vertex ladder."""
    Nv, No = T1.shape
    Vabij = ints["PPHH"]
    R1 = np.zeros_like(T1)
    R2 = np.zeros_like(T2)
    if it == 0 and not initial_doubles:                       # cxx:47-51
        R2[...] = Vabij
        return R1, R2
    Vaibj, Vijkl, Vijka = ints["PHPH"], ints["HHHH"], ints["HHHP"]
    Gij, Gai, Gab = vertex_slices(gamma, No)
    if slice_size is None:
        slice_size = No                                       # cxx:374-384

    X = T2 + ein("ai,bj->abij", T1, T1)                       # cxx:136-138

    # Kac, Lac, Kki, Lki (cxx:141-189)
    Kac = -2.0 * ein("cdkl,adkl->ac", Vabij, X)
    Kac += ein("dckl,adkl->ac", Vabij, X)
    if distinguishable:                                       # cxx:152-160
        Y = T2 + 2.0 * ein("ai,bj->abij", T1, T1)
        Lac = -1.0 * ein("cdkl,adkl->ac", Vabij, Y)
        Lac += 0.5 * ein("dckl,adkl->ac", Vabij, Y)
    else:
        Lac = Kac.copy()
    Lac += 2.0 * ein("Gca,Gdk,dk->ac", Gab, Gai, T1)
    Lac += -1.0 * ein("Gck,Gda,dk->ac", Gai, Gab, T1)
    Kki = 2.0 * ein("cdkl,cdil->ki", Vabij, X)
    Kki += -1.0 * ein("dckl,cdil->ki", Vabij, X)
    if distinguishable:                                       # cxx:178-186
        Y = T2 + 2.0 * ein("ai,bj->abij", T1, T1)
        Lki = ein("cdkl,cdil->ki", Vabij, Y)
        Lki += -0.5 * ein("dckl,cdil->ki", Vabij, Y)
    else:
        Lki = Kki.copy()
    Lki += 2.0 * ein("klic,cl->ki", Vijka, T1)
    Lki += -1.0 * ein("lkic,cl->ki", Vijka, T1)
    R2 += ein("ac,cbij->abij", Lac, T2)                       # cxx:192
    R2 += -1.0 * ein("ki,abkj->abij", Lki, T2)                # cxx:195

    # dressed Gamma_ai terms (cxx:198-216)
    dGai = Gai - ein("Gki,ak->Gai", Gij, T1)
    R2 += ein("Gai,Gbc,cj->abij", dGai, Gab, T1)
    R2 += -1.0 * ein("jika,bk->abij", Vijka, T1)
    R2 += -1.0 * ein("bk,acik,cj->abij", T1, Vabij, T1)

    # Xakic (cxx:218-252)
    dGai = Gai - ein("Gil,al->Gai", Gij, T1) + ein("Gad,di->Gai", Gab, T1)
    Xakic = ein("Gai,Gck->akic", dGai, Gai)
    Y = T2 + 2.0 * ein("ai,bj->abij", T1, T1)
    Xakic += -0.5 * ein("dclk,dail->akic", Vabij, Y)
    Xakic += ein("dclk,adil->akic", Vabij, T2)
    if not distinguishable:                                   # cxx:244-246
        Xakic += -0.5 * ein("cdlk,adil->akic", Vabij, T2)
    Y = 2.0 * T2 - ein("bckj->cbkj", T2)                      # cxx:248-249
    R2 += ein("akic,cbkj->abij", Xakic, Y)

    # Xakci (cxx:254-293)
    dGab = Gab - ein("Gcl,al->Gac", Gai, T1)
    dGij = Gij + ein("Gdk,di->Gki", Gai, T1)
    Xakci = ein("Gac,Gki->akci", dGab, dGij)
    if not distinguishable:                                   # cxx:283-285
        Xakci += -0.5 * ein("cdlk,dail->akci", Vabij, T2)
    R2 += -1.0 * ein("akci,cbkj->abij", Xakci, T2)
    R2 += -1.0 * ein("bkci,ackj->abij", Xakci, T2)
    R2 += ein("baji->abij", R2)                               # cxx:292

    R2 += Vabij                                               # cxx:301

    # Xklij (cxx:303-334)
    Xklij = Vijkl.copy()
    Xklij += ein("klic,cj->klij", Vijka, T1)
    Xklij += ein("lkjc,ci->klij", Vijka, T1)
    R2 += ein("klij,abkl->abij", Xklij, X)
    if distinguishable:                                       # cxx:316-320
        Xklij = ein("ci,cdkl,dj->klij", T1, Vabij, T1)
    else:
        Xklij = ein("cdkl,cdij->klij", Vabij, X)
    R2 += ein("klij,abkl->abij", Xklij, T2 if ppl else X)     # cxx:326-333

    # particle-particle ladder (cxx:336-453)
    if ppl and factors is not None:
        factor_ladder(T1, X, factors[0], factors[1], factors[2], R2)
    elif ppl:
        dGab = Gab - ein("Gbk,ak->Gab", Gai, T1)
        ppl_ladder(dGab, X, slice_size, R2)

    # T1 equations (cxx:457-492)
    R1 += ein("ac,ci->ai", Kac, T1)
    R1 += -1.0 * ein("ki,ak->ai", Kki, T1)
    Kck = 2.0 * ein("cdkl,dl->ck", Vabij, T1)
    Kck += -1.0 * ein("dckl,dl->ck", Vabij, T1)
    R1 += 2.0 * ein("ck,caki->ai", Kck, T2)
    R1 += -1.0 * ein("ck,caik->ai", Kck, T2)
    R1 += ein("ak,ck,ci->ai", T1, Kck, T1)
    R1 += 2.0 * ein("acik,ck->ai", Vabij, T1)
    R1 += -1.0 * ein("ciak,ck->ai", Vaibj, T1)
    R1 += 2.0 * ein("Gca,Gdk,cdik->ai", Gab, Gai, X)
    R1 += -1.0 * ein("Gda,Gck,cdik->ai", Gab, Gai, X)
    R1 += -2.0 * ein("klic,ackl->ai", Vijka, X)
    R1 += ein("lkic,ackl->ai", Vijka, X)
    return R1, R2


# ---------------------------------------------------------------------------
# Update, energy, mixers, driver
# ---------------------------------------------------------------------------

def denominators(eps_i, eps_a):
    """calculateExcitationEnergies (ClusterSinglesDoublesAlgorithm.cxx:343-364):
    D_ai = e_a - e_i, D_abij = e_a + e_b - e_i - e_j."""
    D1 = eps_a[:, None] - eps_i[None, :]
    D2 = (eps_a[:, None, None, None] + eps_a[None, :, None, None]
          - eps_i[None, None, :, None] - eps_i[None, None, None, :])
    return D1, D2


def estimate_amplitudes(R1, R2, T1, T2, eps_i, eps_a, level_shift=0.0):
    """estimateAmplitudesFromResiduum (cxx:302-331): R -= shift*T;
    T_est = -R / (D + shift)."""
    D1, D2 = denominators(eps_i, eps_a)
    R1 = R1 - level_shift * T1
    R2 = R2 - level_shift * T2
    return -R1 / (D1 + level_shift), -R2 / (D2 + level_shift)


def energy(T1, T2, Vabij):
    """getEnergy, restricted, not antisymmetrized (cxx:131-213).
    Returns (E, direct, exchange)."""
    tau = T2 + ein("ai,bj->abij", T1, T1)
    direct = 2.0 * np.sum(tau * Vabij)                        # 0.5*spins^2
    exchange = -1.0 * np.sum(tau * ein("baij->abij", Vabij))  # -0.5*spins
    return direct + exchange, direct, exchange


def energy_of(T1, T2, Vabij, spins=2.0, antisymmetrized=False, fia=None):
    """getEnergy with its arguments (ClusterSinglesDoublesAlgorithm.cxx:
    131-213): spins = 1 for `unrestricted` (cxx:133) in the direct and
    exchange terms (cxx:170-178); `antisymmetrize` evaluates
    0.25 T_abkl V_klab + 0.5 T_aj T_cl V_jlac (cxx:163-167) on
    V_ijab = V_abij (cxx:141-152); an HPFockMatrix f[i,a] adds
    spins * T_ai f_ia (cxx:199-211)."""
    Vijab = ein("abij->ijab", Vabij)
    if antisymmetrized:
        e = 0.25 * ein("abkl,klab->", T2, Vijab) + 0.5 * ein("aj,cl,jlac->", T1, T1, Vijab)
    else:
        dire = 0.5 * spins * spins * (ein("abij,ijab->", T2, Vijab)
                                      + ein("ai,bj,ijab->", T1, T1, Vijab))
        exce = -0.5 * spins * (ein("abij,ijba->", T2, Vijab) + ein("ai,bj,ijba->", T1, T1, Vijab))
        e = dire + exce
    if fia is not None:
        e += spins * ein("ai,ia->", T1, fia)
    return float(e)


def pair_energy(T1, T2, Vabij, antisymmetrized=False):
    """PairEnergy (cxx:218-235)."""
    if antisymmetrized:
        Vijab = ein("abij->ijab", Vabij)
        return (0.25 * ein("abij,ijab->ij", T2, Vijab)
                + 0.5 * ein("ai,bj,ijab->ij", T1, T1, Vijab))
    return pair_energy_restricted(T1, T2, Vabij)


def pair_energy_restricted(T1, T2, Vabij):
    """PairEnergy (cxx:218-235), non-antisymmetrized branch."""
    tau = T2 + ein("ai,bj->abij", T1, T1)
    return 2.0 * ein("abij,abij->ij", tau, Vabij) - ein("abij,baij->ij", tau, Vabij)


def fock_dot(x, y):
    """Real part of FockVector::dot (FockVector.hpp:224-236, dot(x,y) =
    conj(x) y, MathFunctions.hpp:45-48); the DIIS overlaps take 2 Re of it
    (DiisMixer.cxx:123) and |dT|^2/|T|^2 is real."""
    return sum(float(np.real(np.vdot(a, b))) for a, b in zip(x, y))


class DiisMixer:
    """DiisMixer<double> (DiisMixer.cxx:48-180)."""

    def __init__(self, max_residua=4):
        N = int(max_residua)
        self.N = N
        self.amplitudes = [None] * N
        self.residua = [None] * N
        self.next_index = 0
        self.count = 0
        self.B = np.zeros((N + 1, N + 1))
        self.B[0, 1:] = -1.0                                   # cxx:83-96
        self.B[1:, 0] = -1.0
        self.next = None
        self.weights = None
        self.lam = None

    def append(self, A, R):
        N = self.N
        self.amplitudes[self.next_index] = A
        self.residua[self.next_index] = R
        for i in range(N):                                     # cxx:113-146
            if self.residua[i] is not None:
                overlap = 2.0 * fock_dot(self.residua[i], R)
                self.B[self.next_index + 1, i + 1] = overlap
                self.B[i + 1, self.next_index + 1] = overlap
        if self.count < N:
            self.count += 1
        dim = self.count + 1
        rhs = np.zeros(dim)
        rhs[0] = -1.0
        column = np.linalg.solve(self.B[:dim, :dim], rhs)      # dsysv_, cxx:16-40
        self.lam = column[0]
        nxt = [np.zeros_like(c) for c in A]
        weights = []
        for j in range(self.count):                            # cxx:160-176
            i = (self.next_index + N - j) % N
            w = column[i + 1]
            weights.append(w)
            for c in range(len(nxt)):
                nxt[c] += w * self.amplitudes[i][c]
        self.weights = weights
        self.next = nxt
        self.next_index = (self.next_index + 1) % N

    def get(self):
        return self.next


class LinearMixer:
    """LinearMixer<double> (LinearMixer.cxx:13-55)."""

    def __init__(self, ratio=1.0):
        self.ratio = ratio
        self.last = None

    def append(self, nxt, _res):
        if self.last is not None:
            nxt = [self.ratio * n + (1.0 - self.ratio) * l
                   for n, l in zip(nxt, self.last)]
        self.last = nxt

    def get(self):
        return self.last


def solve(gamma, eps_i, eps_a, *, mixer="DiisMixer", max_residua=4,
          mixing_ratio=1.0, max_iterations=16, amplitudes_convergence=1e-5,
          energy_convergence=1e-6, level_shift=0.0, slice_size=None,
          initial=None, callback=None, ppl=True, distinguishable=False, factors=None,
          spins=2.0, antisymmetrized=False, fia=None):
    """ClusterSinglesDoublesAlgorithm::run<double> (cxx:38-128).

    Returns dict(energy, T1, T2, energies=[per iteration], iterations,
    converged, delta=[|dT|/|T| per iteration]).
    """
    No, Nv = eps_i.shape[0], eps_a.shape[0]
    ints = integrals_from_vertex(gamma, No)
    if initial is not None:
        T = [np.array(initial[0], dtype=np.float64), np.array(initial[1], dtype=np.float64)]
    else:
        T = [np.zeros((Nv, No)), np.zeros((Nv, Nv, No, No))]
    if mixer == "DiisMixer":
        mix = DiisMixer(max_residua)
    elif mixer == "LinearMixer":
        mix = LinearMixer(mixing_ratio)
    else:
        raise ValueError("Mixer not implemented: %s" % mixer)
    e, prev = 0.0, 0.0
    energies, deltas = [], []
    converged = False
    i = 0
    while i < max_iterations:
        R1, R2 = residual(i, T[0], T[1], gamma, ints, slice_size,
                          initial_doubles=initial is not None, ppl=ppl,
                          distinguishable=distinguishable, factors=factors)
        E1, E2 = estimate_amplitudes(R1, R2, T[0], T[1], eps_i, eps_a, level_shift)
        dT = [E1 - T[0], E2 - T[1]]
        mix.append([E1, E2], dT)
        T = mix.get()
        e = energy_of(T[0], T[1], ints["PPHH"], spins, antisymmetrized, fia)
        ratio = fock_dot(dT, dT) / fock_dot(T, T)
        energies.append(e)
        deltas.append(np.sqrt(abs(ratio)))
        if callback is not None:
            callback(i, e, T)
        if (abs((e - prev) / e) < abs(energy_convergence)
                and abs(ratio) < abs(amplitudes_convergence * amplitudes_convergence)):
            converged = True
            i += 1
            break
        prev = e
        i += 1
    if max_iterations == 0:
        e = energy_of(T[0], T[1], ints["PPHH"], spins, antisymmetrized, fia)
    return dict(energy=e, T1=T[0], T2=T[1], energies=energies, deltas=deltas,
                iterations=i, converged=converged, integrals=ints)


def mp2_energy(gamma, eps_i, eps_a):
    """MP2 energy the way the reference obtains it: iteration 0 of the CCSD
    driver (R2 = V_abij, T = -R/D) followed by getEnergy."""
    No = eps_i.shape[0]
    ints = integrals_from_vertex(gamma, No)
    _, D2 = denominators(eps_i, eps_a)
    T2 = -ints["PPHH"] / D2
    T1 = np.zeros((eps_a.shape[0], No))
    return energy(T1, T2, ints["PPHH"])[0]
