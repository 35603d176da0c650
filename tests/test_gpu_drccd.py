"""Direct-ring CCD (DrccdEnergyFromCoulombIntegrals.cxx:31-80) through the C
ABI against the oracle's restatement: residuals to 1e-11 relative to their
max norm for every argument combination, on the vertex engine and on an
integrals-only engine; whole runs to 1e-10 Eh per iteration; the converged
direct energy equals the direct-RPA correlation energy 1/2 Tr(Omega - A)."""

import numpy as np
import pytest

from conftest import announce, random_amplitudes
from oracle import ccsd, drccd, formulations, philox

pytestmark = pytest.mark.gpu

COMBOS = [(False, False, False), (True, False, False), (False, True, False),
          (False, False, True), (True, False, True)]


def system(seed=21, No=4, Nv=13, NG=40):
    g, ei, ea = philox.system(seed, No, Nv, NG)
    return g, ei, ea, ccsd.integrals_from_vertex(g, No)["PPHH"]


def symmetric_amplitudes(Nv, No, seed):
    _, T2 = random_amplitudes(Nv, No, seed=seed, scale=0.05)
    return 0.5 * (T2 + np.transpose(T2, (1, 0, 3, 2)))


@pytest.mark.parametrize("unrestricted,linearized,adjacent", COMBOS)
def test_drccd_residual_matches_oracle(unrestricted, linearized, adjacent):
    from sisi4s_amd import engine
    g, ei, ea, V = system()
    No, Nv = len(ei), len(ea)
    e = engine.Engine(No, Nv, g.shape[0])
    e.set_system(g, ei, ea)
    e.set_drccd(unrestricted, linearized, adjacent)
    # integrals-only engine: PPHH alone, then PHHP / HHPP uploaded as well
    f = engine.Engine(No, Nv, 1)
    f.set_eigenenergies(ei, ea)
    f.set_integrals(engine.PPHH, V)
    f.set_drccd(unrestricted, linearized, adjacent)
    T1 = np.zeros((Nv, No))
    T2 = symmetric_amplitudes(Nv, No, 5)
    _, Va, Vi = drccd.integrals(V)
    _, O2 = drccd.residual(1, T2, V, Va, Vi, spins=1.0 if unrestricted else 2.0,
                           linearized=linearized, adjacent_pairs_exchange=adjacent)
    scale = float(np.abs(O2).max())
    announce("drccd residual unrestricted=%d linearized=%d adjacent=%d"
             % (unrestricted, linearized, adjacent))
    for eng in (e, f):
        R1, R2 = eng.residual(1, T1, T2)
        assert not R1.any()
        assert float(np.abs(R2 - O2).max()) <= 1e-11 * scale
    f.set_integrals(engine.PHHP, Va)
    f.set_integrals(engine.HHPP, Vi)
    _, R2 = f.residual(1, T1, T2)
    assert float(np.abs(R2 - O2).max()) <= 1e-11 * scale
    _, R2 = e.residual(0, T1, T2)  # MP2 start
    assert np.array_equal(R2, V) or float(np.abs(R2 - V).max()) <= 1e-15 * np.abs(V).max()


@pytest.mark.parametrize("unrestricted", [False, True])
def test_drccd_solve_matches_oracle_and_rpa(unrestricted):
    from sisi4s_amd import engine
    g, ei, ea, V = system(22, 4, 14, 48)
    No, Nv = len(ei), len(ea)
    e = engine.Engine(No, Nv, g.shape[0])
    e.set_system(g, ei, ea)
    e.set_drccd(unrestricted=unrestricted)
    s = 1.0 if unrestricted else 2.0
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=100, amplitudes_convergence=1e-11,
              energy_convergence=1e-13)
    announce("drccd solve unrestricted=%d" % unrestricted)
    rep = e.ccsd_solve(engine.default_opts(**kw))
    ref = drccd.solve(V, ei, ea, spins=s, **kw)
    assert rep["converged"] and ref["converged"]
    assert rep["iterations"] == ref["iterations"]
    assert np.abs(e.energies() - np.array(ref["energies"])).max() <= 1e-10
    assert abs(rep["direct"] - ref["direct"]) <= 1e-10
    assert abs(rep["exchange"] - ref["exchange"]) <= 1e-10
    # direct-RPA correlation energy, independent of the oracle
    K = np.transpose(V, (0, 2, 1, 3)).reshape(Nv * No, Nv * No)
    delta = (ea[:, None] - ei[None, :]).reshape(-1)
    sq = np.sqrt(delta)
    omega = np.sqrt(np.linalg.eigvalsh(sq[:, None] * (np.diag(delta) + 2 * s * K) * sq[None, :]))
    e_rpa = 0.5 * (omega.sum() - np.trace(np.diag(delta) + s * K))
    assert abs(rep["direct"] - e_rpa) <= 1e-9 * abs(e_rpa)
    # iteration 1 is MP2 (restricted: the closed form)
    if not unrestricted:
        assert abs(e.energies()[0] - formulations.mp2_closed_form(g, ei, ea)) <= 1e-11


def test_drccd_mixed_precision_reaches_fp64_energy():
    """Split-BF16x3 bulk iterations, FP64 after |dT|/|T| <= 1e-5."""
    from sisi4s_amd import engine
    g, ei, ea, V = system(23, 5, 24, 96)
    e = engine.Engine(len(ei), len(ea), g.shape[0])
    e.set_system(g, ei, ea)
    e.set_drccd(adjacent_pairs_exchange=True)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=120, amplitudes_convergence=1e-9,
              energy_convergence=1e-11)
    announce("drccd fp64 vs mixed")
    r64 = e.ccsd_solve(engine.default_opts(**kw))
    rmx = e.ccsd_solve(engine.default_opts(precision="bf16x3", switch_threshold=1e-5, **kw))
    assert r64["converged"] and rmx["converged"]
    assert rmx["fp64_iterations"] < rmx["iterations"]
    assert abs(rmx["energy"] - r64["energy"]) <= 1e-8
    ref = drccd.solve(V, ei, ea, adjacent_pairs_exchange=True, **kw)
    assert abs(r64["energy"] - ref["energy"]) <= 1e-10


def test_complex_drccd_matches_oracle_and_real_path():
    """F = complex: residual against the oracle's complex restatement, and a
    whole run on a gauge-transformed vertex gives the real path's energies."""
    from sisi4s_amd import engine
    from oracle import ccsd_complex as cz
    g, ei, ea, V = system(24, 4, 12, 36)
    No, Nv = len(ei), len(ea)
    gz, _ = cz.gauge_transform(g.astype(complex), No, 25)
    iz = cz.integrals_from_vertex(gz, No)
    ez = engine.Engine(No, Nv, g.shape[0])
    ez.set_system(gz, ei, ea)
    assert ez.complex
    ez.set_drccd(adjacent_pairs_exchange=True)
    T2 = symmetric_amplitudes(Nv, No, 6) * (1 + 0.3j)
    _, O2 = drccd.residual(1, T2, iz["PPHH"], iz["PHHP"], iz["HHPP"], adjacent_pairs_exchange=True)
    announce("complex drccd residual and solve")
    R1, R2 = ez.residual(1, np.zeros((Nv, No), complex), T2)
    assert float(np.abs(R2 - O2).max()) <= 1e-11 * float(np.abs(O2).max()) and not R1.any()
    er = engine.Engine(No, Nv, g.shape[0])
    er.set_system(g, ei, ea)
    er.set_drccd(adjacent_pairs_exchange=True)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=60, amplitudes_convergence=1e-10,
              energy_convergence=1e-12)
    rz = ez.ccsd_solve(engine.default_opts(**kw))
    rr = er.ccsd_solve(engine.default_opts(**kw))
    assert rz["converged"] and rr["converged"] and rz["iterations"] == rr["iterations"]
    assert np.abs(ez.energies() - er.energies()).max() <= 1e-10
    assert abs(rz["energy_imag"]) <= 1e-10
