"""FP64 CPU restatement of the reference's complex CCSD path (test infrastructure).

Follows, term by term:

* src/algorithms/CoulombIntegralsFromVertex.cxx:613-805 (calculateComplexIntegrals,
  no antisymmetrization, no forceReal)
* src/algorithms/CcsdEnergyFromCoulombIntegrals.cxx:496-891 (complex residual)
* src/algorithms/ClusterSinglesDoublesAlgorithm.cxx:490-565
  (sliceCoupledCoulombIntegrals for complex amplitudes), 887-915
  (sliceIntoResiduum), 38-128 (driver with F = complex), 131-213 (getEnergy
  with the HHPP integrals)
* src/mixers/DiisMixer.cxx:103-180 with overlaps 2 Re<dR_i|dR_j>

Vertex slices (CoulombIntegralsFromVertex.cxx:120-136, holes first):
Gij[G,i,j] = Gamma[G,i,j], Gia[G,i,a] = Gamma[G,i,No+a], Gai[G,a,i] =
Gamma[G,No+a,i], Gab[G,a,b] = Gamma[G,No+a,No+b].  The reference's
conjugate transposes (conjTransposeGammaGai[G,a,i] = conj(Gia[G,i,a]) and
likewise for Gia, Gab, Gij) are written here as conj() of a transposed slice.
"""

import numpy as np

from . import ccsd

ein = ccsd.ein


def vertex_slices(gamma, No):
    Gij = gamma[:, :No, :No]
    Gia = gamma[:, :No, No:]
    Gai = gamma[:, No:, :No]
    Gab = gamma[:, No:, No:]
    return Gij, Gia, Gai, Gab


def integrals_from_vertex(gamma, No):
    """CoulombIntegralsFromVertex::calculateComplexIntegrals (cxx:762-805)."""
    Gij, Gia, Gai, Gab = vertex_slices(gamma, No)
    cGai = ein("Gia->Gai", Gia.conj())         # conjTransposeGammaGai
    cGia = ein("Gai->Gia", Gai.conj())         # conjTransposeGammaGia
    cGab = ein("Gba->Gab", Gab.conj())         # conjTransposeGammaGab
    cGij = ein("Gji->Gij", Gij.conj())         # conjTransposeGammaGij
    return dict(
        PPHH=ein("Gai,Gbj->abij", cGai, Gai),  # cxx:765
        PHHP=ein("Gaj,Gib->aijb", cGai, Gia),  # cxx:772
        HHPP=ein("Gia,Gjb->ijab", cGia, Gia),  # cxx:779
        PHPH=ein("Gab,Gij->aibj", cGab, Gij),  # cxx:786
        HHHH=ein("Gik,Gjl->ijkl", cGij, Gij),  # cxx:793
        HHHP=ein("Gik,Gja->ijka", cGij, Gia),  # cxx:800
        PHHH=ein("Gaj,Gik->aijk", cGai, Gij),  # cxx:807
    )


def ppl_ladder(left, right, X, slice_size, R2):
    """Ladder over slice pairs b <= a (CcsdEnergyFromCoulombIntegrals.cxx:
    836-858): Vxycd = sum_G left[G,x,c] right[G,y,d] with x in slice a,
    y in slice b (ClusterSinglesDoublesAlgorithm.cxx:549-557)."""
    Nv = left.shape[1]
    for b in range(0, Nv, slice_size):
        for a in range(b, Nv, slice_size):
            xe, ye = min(a + slice_size, Nv), min(b + slice_size, Nv)
            Vxycd = ein("Gxc,Gyd->xycd", left[:, a:xe, :], right[:, b:ye, :])
            Rxyij = ein("xycd,cdij->xyij", Vxycd, X)
            ccsd.slice_into_residuum(Rxyij, a, b, R2)


def factor_ladder(T1, X, Lambda, Pi, PiOut, slice_size, R2):
    """sliceAmplitudesFromCoupledCoulombFactors for complex amplitudes
    (ClusterSinglesDoublesAlgorithm.cxx:745-880) over all pairs of factor
    slices: Lambda[G,R], Pi[p,R] (FactorOrbitals), PiOut[p,R]
    (OutgoingFactorOrbitals), X = T2 + T1 T1."""
    Nv, No = T1.shape
    NR = Pi.shape[1]
    PiaR, PiiR = Pi[No:, :], Pi[:No, :]
    PicR, PijR = PiOut[No:, :], PiOut[:No, :]
    dressed = PicR - ein("kR,bk->bR", PijR, T1)                       # cxx:851-853
    conj_dressed = PiaR.conj() - ein("kR,ak->aR", PiiR.conj(), T1)    # cxx:855-857
    for b in range(0, NR, slice_size):
        for a in range(0, NR, slice_size):
            L, Rr = slice(a, min(a + slice_size, NR)), slice(b, min(b + slice_size, NR))
            XRaij = ein("cdij,cR->Rdij", X, PicR[:, L].conj())        # cxx:827
            XRSij = ein("Rdij,dS->RSij", XRaij, PiaR[:, Rr])          # cxx:830
            VRS = ein("GR,GS->RS", Lambda[:, L].conj(), Lambda[:, Rr])
            XRSij = XRSij * VRS[:, :, None, None]                      # cxx:834
            XRbij = ein("RSij,bS->Rbij", XRSij, dressed[:, Rr])        # cxx:867
            R2 += ein("Rbij,aR->abij", XRbij, conj_dressed[:, L])      # cxx:874


def residual(it, T1, T2, gamma, ints, slice_size=None, initial_doubles=False,
             ppl=True, distinguishable=False, factors=None):
    """CcsdEnergyFromCoulombIntegrals::getResiduum for complex amplitudes
    (cxx:496-891).  factors = (Lambda, Pi, PiOut, factors_slice_size) selects
    the factor-based ladder (cxx:786-813).  Returns (R1, R2), complex."""
    Nv, No = T1.shape
    Vabij = ints["PPHH"]
    R1 = np.zeros((Nv, No), complex)
    R2 = np.zeros((Nv, Nv, No, No), complex)
    if it == 0 and not initial_doubles:                               # cxx:516-521
        R2[...] = Vabij
        return R1, R2
    Vaijb, Vijab, Vaibj = ints["PHHP"], ints["HHPP"], ints["PHPH"]
    Vijkl, Vijka, Vaijk = ints["HHHH"], ints["HHHP"], ints["PHHH"]
    Gij, Gia, Gai, Gab = vertex_slices(gamma, No)
    cGai = ein("Gia->Gai", Gia.conj())
    cGia = ein("Gai->Gia", Gai.conj())
    cGab = ein("Gba->Gab", Gab.conj())
    cGij = ein("Gji->Gij", Gij.conj())
    if slice_size is None:
        slice_size = No                                                # cxx:822-832

    X = T2 + ein("ai,bj->abij", T1, T1)                               # cxx:584-586
    Y = T2 + 2.0 * ein("ai,bj->abij", T1, T1)

    # Kac, Lac, Kki, Lki (cxx:588-646)
    Kac = -2.0 * ein("klcd,adkl->ac", Vijab, X)
    Kac += ein("kldc,adkl->ac", Vijab, X)
    if distinguishable:
        Lac = -1.0 * ein("klcd,adkl->ac", Vijab, Y) + 0.5 * ein("kldc,adkl->ac", Vijab, Y)
    else:
        Lac = Kac.copy()
    Lac += 2.0 * ein("Gac,Gkd,dk->ac", cGab, Gia, T1)
    Lac += -1.0 * ein("Gad,Gkc,dk->ac", cGab, Gia, T1)
    Kki = 2.0 * ein("klcd,cdil->ki", Vijab, X)
    Kki += -1.0 * ein("kldc,cdil->ki", Vijab, X)
    if distinguishable:
        Lki = ein("klcd,cdil->ki", Vijab, Y) - 0.5 * ein("kldc,cdil->ki", Vijab, Y)
    else:
        Lki = Kki.copy()
    Lki += 2.0 * ein("klic,cl->ki", Vijka, T1)
    Lki += -1.0 * ein("lkic,cl->ki", Vijka, T1)
    R2 += ein("ac,cbij->abij", Lac, T2)                               # cxx:640
    R2 += -1.0 * ein("ki,abkj->abij", Lki, T2)                        # cxx:643

    # dressed conj-transposed Gamma_ai terms (cxx:646-659)
    dG = cGai - ein("Gki,ak->Gai", cGij, T1)
    R2 += ein("Gai,Gbc,cj->abij", dG, Gab, T1)
    R2 += -1.0 * ein("akij,bk->abij", Vaijk, T1)
    R2 += -1.0 * ein("bk,akic,cj->abij", T1, Vaijb, T1)

    # Xakic (cxx:661-694)
    dG = (cGai - ein("Gli,al->Gai", cGij, T1) + ein("Gad,di->Gai", cGab, T1)
          - 0.5 * ein("Gld,dail->Gai", cGia, Y))
    Xakic = ein("Gai,Gkc->akic", dG, Gia)
    Yv = ein("lkdc->dclk", Vijab)
    if not distinguishable:
        Yv = Yv - 0.5 * ein("lkcd->dclk", Vijab)
    Xakic += ein("dclk,adil->akic", Yv, T2)
    Y2 = 2.0 * T2 - ein("bckj->cbkj", T2)
    R2 += ein("akic,cbkj->abij", Xakic, Y2)

    # Xakci (cxx:696-727)
    dGab = cGab - ein("Glc,al->Gac", cGia, T1)
    dGij = Gij + ein("Gkd,di->Gki", Gia, T1)
    Xakci = ein("Gac,Gki->akci", dGab, dGij)
    if not distinguishable:
        Xakci += -0.5 * ein("lkcd,dail->akci", Vijab, T2)
    R2 += -1.0 * ein("akci,cbkj->abij", Xakci, T2)
    R2 += -1.0 * ein("bkci,ackj->abij", Xakci, T2)
    R2 += ein("baji->abij", R2)                                       # cxx:726

    R2 += Vabij                                                       # cxx:735

    # Xklij (cxx:737-768)
    Xklij = Vijkl.copy()
    Xklij += ein("klic,cj->klij", Vijka, T1)
    Xklij += ein("lkjc,ci->klij", Vijka, T1)
    R2 += ein("klij,abkl->abij", Xklij, X)
    if distinguishable:
        Xklij = ein("ci,klcd,dj->klij", T1, Vijab, T1)
    else:
        Xklij = ein("klcd,cdij->klij", Vijab, X)
    R2 += ein("klij,abkl->abij", Xklij, T2 if ppl else X)

    # particle-particle ladder from the dressed vertex
    # (ClusterSinglesDoublesAlgorithm.cxx:531-557)
    if ppl and factors is not None:
        factor_ladder(T1, X, factors[0], factors[1], factors[2], factors[3], R2)
    elif ppl:
        right = Gab - ein("Gkb,ak->Gab", Gia, T1)
        left = cGab - ein("Gkb,ak->Gab", cGia, T1)
        ppl_ladder(left, right, X, slice_size, R2)

    # T1 equations (cxx:861-891)
    R1 += ein("ac,ci->ai", Kac, T1)
    R1 += -1.0 * ein("ki,ak->ai", Kki, T1)
    Kck = 2.0 * ein("klcd,dl->ck", Vijab, T1)
    Kck += -1.0 * ein("kldc,dl->ck", Vijab, T1)
    R1 += 2.0 * ein("ck,caki->ai", Kck, T2)
    R1 += -1.0 * ein("ck,caik->ai", Kck, T2)
    R1 += ein("ak,ck,ci->ai", T1, Kck, T1)
    R1 += 2.0 * ein("akic,ck->ai", Vaijb, T1)
    R1 += -1.0 * ein("akci,ck->ai", Vaibj, T1)
    R1 += 2.0 * ein("Gac,Gkd,cdik->ai", cGab, Gia, X)
    R1 += -1.0 * ein("Gad,Gkc,cdik->ai", cGab, Gia, X)
    R1 += -2.0 * ein("klic,ackl->ai", Vijka, X)
    R1 += ein("lkic,ackl->ai", Vijka, X)
    return R1, R2


def energy(T1, T2, Vijab):
    """getEnergy with F = complex, restricted, not antisymmetrized
    (ClusterSinglesDoublesAlgorithm.cxx:131-189): direct and exchange from
    the HHPP integrals.  Returns (E, direct, exchange), complex."""
    tau = T2 + ein("ai,bj->abij", T1, T1)
    direct = 2.0 * ein("abij,ijab->", tau, Vijab)
    exchange = -1.0 * ein("abij,ijba->", tau, Vijab)
    return direct + exchange, direct, exchange


def pair_energy(T1, T2, Vijab):
    """PairEnergy with F = complex (ClusterSinglesDoublesAlgorithm.cxx:
    218-235, non-antisymmetrized branch)."""
    tau = T2 + ein("ai,bj->abij", T1, T1)
    return 2.0 * ein("abij,ijab->ij", tau, Vijab) - ein("abij,ijba->ij", tau, Vijab)


def solve(gamma, eps_i, eps_a, *, mixer="DiisMixer", max_residua=4,
          mixing_ratio=1.0, max_iterations=16, amplitudes_convergence=1e-5,
          energy_convergence=1e-6, level_shift=0.0, slice_size=None,
          initial=None, ppl=True, distinguishable=False, factors=None):
    """ClusterSinglesDoublesAlgorithm::run<complex> (cxx:38-128); energies
    are complex, the reported energy is their real part (cxx:30-33)."""
    No, Nv = eps_i.shape[0], eps_a.shape[0]
    ints = integrals_from_vertex(gamma, No)
    if initial is not None:
        T = [np.array(initial[0], dtype=complex), np.array(initial[1], dtype=complex)]
    else:
        T = [np.zeros((Nv, No), complex), np.zeros((Nv, Nv, No, No), complex)]
    if mixer == "DiisMixer":
        mix = ccsd.DiisMixer(max_residua)
    elif mixer == "LinearMixer":
        mix = ccsd.LinearMixer(mixing_ratio)
    else:
        raise ValueError("Mixer not implemented: %s" % mixer)
    e, prev = 0.0j, 0.0j
    energies, deltas = [], []
    converged = False
    i = 0
    while i < max_iterations:
        R1, R2 = residual(i, T[0], T[1], gamma, ints, slice_size,
                          initial_doubles=initial is not None, ppl=ppl,
                          distinguishable=distinguishable, factors=factors)
        E1, E2 = ccsd.estimate_amplitudes(R1, R2, T[0], T[1], eps_i, eps_a, level_shift)
        dT = [E1 - T[0], E2 - T[1]]
        mix.append([E1, E2], dT)
        T = mix.get()
        e = energy(T[0], T[1], ints["HHPP"])[0]
        ratio = ccsd.fock_dot(dT, dT) / ccsd.fock_dot(T, T)
        energies.append(e)
        deltas.append(np.sqrt(abs(ratio)))
        if (abs((e - prev) / e) < abs(energy_convergence)
                and abs(ratio) < abs(amplitudes_convergence * amplitudes_convergence)):
            converged = True
            i += 1
            break
        prev = e
        i += 1
    if max_iterations == 0:
        e = energy(T[0], T[1], ints["HHPP"])[0]
    return dict(energy=e.real, energy_complex=e, T1=T[0], T2=T[1], energies=energies,
                deltas=deltas, iterations=i, converged=converged, integrals=ints)


def gauge_transform(gamma, No, seed):
    """A complex vertex with the same physics as gamma: a random phase per
    orbital (Gamma[G,p,q] -> e^{-i t_p} Gamma[G,p,q] e^{i t_q}) and a random
    unitary mix of the auxiliary index.  Every Coulomb integral
    sum_G conj(Gamma[G,r,p]) Gamma[G,s,q] only picks up orbital phases, so
    energies are unchanged; amplitudes pick up e^{-i(t_a+t_b-t_i-t_j)}."""
    rng = np.random.default_rng(seed)
    NG, Np, _ = gamma.shape
    t = rng.uniform(0.0, 2.0 * np.pi, Np)
    q, r = np.linalg.qr(rng.normal(size=(NG, NG)) + 1j * rng.normal(size=(NG, NG)))
    U = q * (np.diag(r) / np.abs(np.diag(r)))
    g = np.exp(-1j * t)[None, :, None] * gamma * np.exp(1j * t)[None, None, :]
    return ein("GH,Hpq->Gpq", U, g), t
