"""IP-EOM-CCSD right-apply (SimilarityTransformedHamiltonian.cxx:64-68) through
the C ABI against the oracle's restatement: sigma vectors to 1e-11 of their
max norm on generic spin-orbital integrals, and for two electrons the
eigenvalues of the device's Hbar (built column by column from sigma) equal
E(N-1) - E(N) from full CI."""

import numpy as np
import pytest

from conftest import announce
from oracle import eom, uccsd

pytestmark = pytest.mark.gpu

CLASSES = ("PPHH", "HHHH", "PPPP", "HHHP", "PPPH", "PHPH")


def generic(seed, no, n, NG=12):
    rng = np.random.default_rng(seed)
    g = rng.normal(size=(NG, n, n)) * 0.25
    g = 0.5 * (g + np.transpose(g, (0, 2, 1)))
    eri = np.einsum("Gpr,Gqs->pqrs", g, g)
    asym = eri - np.swapaxes(eri, 2, 3)
    ei = np.sort(rng.uniform(-1.5, -0.6, no))
    ea = np.sort(rng.uniform(0.3, 2.0, n - no))
    return asym, uccsd.classes(asym, no), ei, ea


def spin_orbital_engine(ints, ei, ea):
    from sisi4s_amd import engine
    e = engine.Engine(len(ei), len(ea), 1)
    e.set_eigenenergies(ei, ea)
    for c in CLASSES:
        e.set_integrals(getattr(engine, c), ints[c])
    e.set_formulation(engine.FORMULATION_UCCSD)
    return e


def test_ip_sigma_matches_oracle():
    asym, ints, ei, ea = generic(3, 4, 12)
    no, nv = 4, 8
    r = uccsd.solve(ints, ei, ea, max_residua=6, max_iterations=200,
                    amplitudes_convergence=1e-11, energy_convergence=1e-13)
    assert r["converged"]
    e = spin_orbital_engine(ints, ei, ea)
    rng = np.random.default_rng(4)
    r1 = rng.normal(size=no)
    r2 = rng.normal(size=(no, no, nv))
    r2 = r2 - r2.transpose(1, 0, 2)
    announce("ip-eom sigma No=4 Nv=8 spin orbitals")
    s1, s2 = e.eom_ip_sigma(r["T1"], r["T2"], r1, r2)
    H = eom.hbar_ip(ints, ei, ea, r["T1"], r["T2"])
    o1, o2 = eom.ip_sigma(H, r1, r2)
    scale = max(float(np.abs(o1).max()), float(np.abs(o2).max()))
    assert float(np.abs(s1 - o1).max()) <= 1e-11 * scale
    assert float(np.abs(s2 - o2).max()) <= 1e-11 * scale


def test_ip_eigenvalues_exact_for_two_electrons():
    asym, ints, ei, ea = generic(7, 2, 8)
    n, nv = 8, 6
    r = uccsd.solve(ints, ei, ea, max_residua=6, max_iterations=300,
                    amplitudes_convergence=1e-12, energy_convergence=1e-14)
    assert r["converged"]
    e = spin_orbital_engine(ints, ei, ea)
    dim = 2 + nv
    M = np.zeros((dim, dim))
    announce("ip-eom two electrons, Hbar from device sigma")
    for c in range(dim):
        x = np.zeros(dim)
        x[c] = 1.0
        r1 = x[:2].copy()
        r2 = np.zeros((2, 2, nv))
        r2[0, 1] = x[2:]
        r2[1, 0] = -x[2:]
        s1, s2 = e.eom_ip_sigma(r["T1"], r["T2"], r1, r2)
        M[:2, c] = s1
        M[2:, c] = s2[0, 1]
    ev = np.sort(np.linalg.eigvals(M).real)
    eps = np.concatenate([ei, ea])
    h = np.diag(eps) - np.einsum("pkqk->pq", asym[:, :2, :, :2])
    dets = [(p, q) for p in range(n) for q in range(p + 1, n)]
    Hf = np.array([[h[p, s] * (q == t) - h[p, t] * (q == s) - h[q, s] * (p == t)
                    + h[q, t] * (p == s) + asym[p, q, s, t] for (s, t) in dets] for (p, q) in dets])
    exact = np.sort(np.linalg.eigvalsh(h) - np.linalg.eigvalsh(Hf)[0])
    assert np.abs(ev - exact).max() <= 1e-10


def test_ea_sigma_matches_oracle():
    """EA-EOM-CCSD right-apply (SimilarityTransformedHamiltonian.cxx:289-294)
    through the C ABI against the oracle's sigma, and for two virtual spin
    orbitals the eigenvalues of the device's Hbar against the oracle's, which
    equal full CI there."""
    asym, ints, ei, ea = generic(3, 4, 12)
    no, nv = 4, 8
    r = uccsd.solve(ints, ei, ea, max_residua=6, max_iterations=200,
                    amplitudes_convergence=1e-11, energy_convergence=1e-13)
    assert r["converged"]
    e = spin_orbital_engine(ints, ei, ea)
    rng = np.random.default_rng(5)
    r1 = rng.normal(size=nv)
    r2 = rng.normal(size=(no, nv, nv))
    r2 = r2 - r2.transpose(0, 2, 1)
    announce("ea-eom sigma No=4 Nv=8 spin orbitals")
    s1, s2 = e.eom_ea_sigma(r["T1"], r["T2"], r1, r2)
    H = eom.hbar_ea(ints, ei, ea, r["T1"], r["T2"])
    o1, o2 = eom.ea_sigma(H, r1, r2)
    scale = max(float(np.abs(o1).max()), float(np.abs(o2).max()))
    assert float(np.abs(s1 - o1).max()) <= 1e-11 * scale
    assert float(np.abs(s2 - o2).max()) <= 1e-11 * scale
    e.close()
    asym, ints, ei, ea = generic(9, 3, 5)
    r = uccsd.solve(ints, ei, ea, max_residua=6, max_iterations=300,
                    amplitudes_convergence=1e-12, energy_convergence=1e-14)
    e = spin_orbital_engine(ints, ei, ea)
    dim = 2 + 3
    M = np.zeros((dim, dim))
    announce("ea-eom two virtual spin orbitals, Hbar from device sigma")
    for c in range(dim):
        x = np.zeros(dim)
        x[c] = 1.0
        r1 = x[:2].copy()
        r2 = np.zeros((3, 2, 2))
        r2[:, 0, 1] = x[2:]
        r2[:, 1, 0] = -x[2:]
        s1, s2 = e.eom_ea_sigma(r["T1"], r["T2"], r1, r2)
        M[:2, c] = s1
        M[2:, c] = s2[:, 0, 1]
    ev = np.sort(np.linalg.eigvals(M).real)
    ref = eom.ea_eigenvalues(ints, ei, ea, r["T1"], r["T2"])
    assert np.abs(ev - ref).max() <= 1e-10
    e.close()
