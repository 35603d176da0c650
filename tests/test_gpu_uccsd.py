"""Spin-orbital CCSD (UccsdAmplitudesFromCoulombIntegrals.cxx:107-241) through
the C ABI against the oracle's restatement on generic (spin-symmetry-broken)
antisymmetrized integrals: residuals to 1e-11 relative to their max norm,
whole runs to 1e-10 Eh per iteration; on a closed-shell system in spin
orbitals the converged energy is the restricted CCSD energy, and for two
electrons it is full CI."""

import numpy as np
import pytest

from conftest import announce, random_amplitudes
from oracle import ccsd, philox, uccsd

pytestmark = pytest.mark.gpu

CLASSES = ("PPHH", "HHHH", "PPPP", "HHHP", "PPPH", "PHPH")


def generic_system(seed, no, nv, NG=14, scale=0.25):
    rng = np.random.default_rng(seed)
    n = no + nv
    g = rng.normal(size=(NG, n, n)) * scale
    g = 0.5 * (g + np.transpose(g, (0, 2, 1)))
    eri = np.einsum("Gpr,Gqs->pqrs", g, g)
    asym = eri - np.swapaxes(eri, 2, 3)
    ei = np.sort(rng.uniform(-2.0, -0.5, no))
    ea = np.sort(rng.uniform(0.2, 3.0, nv))
    return asym, uccsd.classes(asym, no), ei, ea


def make_engine(ints, ei, ea):
    from sisi4s_amd import engine
    e = engine.Engine(len(ei), len(ea), 1)
    e.set_eigenenergies(ei, ea)
    for c in CLASSES:
        e.set_integrals(getattr(engine, c), ints[c])
    e.set_formulation(engine.FORMULATION_UCCSD)
    return e


def test_uccsd_residual_matches_oracle():
    asym, ints, ei, ea = generic_system(3, 4, 10)
    e = make_engine(ints, ei, ea)
    T1, T2 = random_amplitudes(10, 4, seed=8, scale=0.05)
    T2 = T2 - np.transpose(T2, (1, 0, 2, 3))          # antisymmetric in (a,b) and (i,j)
    T2 = 0.5 * (T2 - np.transpose(T2, (0, 1, 3, 2)))
    announce("uccsd residual No=4 Nv=10 spin orbitals")
    R1, R2 = e.residual(1, T1, T2)
    O1, O2 = uccsd.residual(1, T1, T2, ints)
    scale = float(np.abs(O2).max())
    assert float(np.abs(R2 - O2).max()) <= 1e-11 * scale
    assert float(np.abs(R1 - O1).max()) <= 1e-11 * scale
    R1, R2 = e.residual(0, T1, T2)
    assert not R1.any() and float(np.abs(R2 - ints["PPHH"]).max()) == 0.0


def test_uccsd_solve_matches_oracle_and_mixed_precision():
    from sisi4s_amd import engine
    asym, ints, ei, ea = generic_system(4, 4, 12)
    e = make_engine(ints, ei, ea)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=80, amplitudes_convergence=1e-10,
              energy_convergence=1e-12)
    announce("uccsd solve")
    rep = e.ccsd_solve(engine.default_opts(**kw))
    ref = uccsd.solve(ints, ei, ea, **kw)
    assert rep["converged"] and ref["converged"] and rep["iterations"] == ref["iterations"]
    assert np.abs(e.energies() - np.array(ref["energies"])).max() <= 1e-10
    rmx = e.ccsd_solve(engine.default_opts(precision="bf16x3", switch_threshold=1e-5, **kw))
    assert rmx["converged"] and abs(rmx["energy"] - rep["energy"]) <= 1e-8


def test_uccsd_closed_shell_and_two_electron_closed_forms():
    from sisi4s_amd import engine
    No, Nv, NG = 3, 7, 20
    g, ei, ea = philox.system(12, No, Nv, NG)
    ints = uccsd.spin_orbital_integrals(g, No, Nv)
    e = make_engine(ints, np.repeat(ei, 2), np.repeat(ea, 2))
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=150, amplitudes_convergence=1e-10,
              energy_convergence=1e-12)
    announce("uccsd closed shell vs restricted")
    rep = e.ccsd_solve(engine.default_opts(**kw))
    rr = ccsd.solve(g, ei, ea, **kw)
    assert rep["converged"] and rr["converged"]
    assert abs(rep["energy"] - rr["energy"]) <= 1e-9 * abs(rr["energy"])
    # two electrons: full CI
    asym, ints2, ei2, ea2 = generic_system(9, 2, 8)
    e2 = make_engine(ints2, ei2, ea2)
    announce("uccsd two electrons vs full CI")
    r2 = e2.ccsd_solve(engine.default_opts(**kw))
    n = asym.shape[0]
    eps = np.concatenate([ei2, ea2])
    h = np.diag(eps) - np.einsum("pkqk->pq", asym[:, :2, :, :2])
    dets = [(p, q) for p in range(n) for q in range(p + 1, n)]
    H = np.array([[h[p, r] * (q == s) - h[p, s] * (q == r) - h[q, r] * (p == s)
                   + h[q, s] * (p == r) + asym[p, q, r, s] for (r, s) in dets] for (p, q) in dets])
    e_fci = np.linalg.eigvalsh(H)[0] - (h[0, 0] + h[1, 1] + asym[0, 1, 0, 1])
    assert r2["converged"] and abs(r2["energy"] - e_fci) <= 1e-9 * abs(e_fci)
