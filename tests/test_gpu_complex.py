"""Complex path on the GPU (CcsdEnergyFromCoulombIntegrals.cxx:496-891 and
calculateComplexIntegrals) through the C ABI against the oracle's complex
restatement, on genuinely complex vertices: a synthetic vertex after a random
orbital gauge and auxiliary-index rotation, and the plane-wave UEG vertex."""

import numpy as np
import pytest

from conftest import announce, random_amplitudes
from oracle import ccsd, ccsd_complex as cz, philox, ueg

pytestmark = pytest.mark.gpu


def complex_system(seed, No, Nv, NG):
    g, ei, ea = philox.system(seed, No, Nv, NG)
    gz, _ = cz.gauge_transform(g.astype(complex), No, seed + 1)
    return gz, ei, ea


def make_engine(gz, ei, ea):
    from sisi4s_amd import engine
    No, Nv, NG = len(ei), len(ea), gz.shape[0]
    e = engine.Engine(No, Nv, NG)
    e.set_system(gz, ei, ea)
    assert e.complex
    return e


def complex_amplitudes(Nv, No, seed, scale=0.1):
    a1, a2 = random_amplitudes(Nv, No, seed=seed, scale=scale)
    b1, b2 = random_amplitudes(Nv, No, seed=seed + 100, scale=scale)
    return a1 + 1j * b1, a2 + 1j * b2


def test_complex_integrals_match_oracle():
    from sisi4s_amd import engine
    gz, ei, ea = complex_system(3, 4, 11, 30)
    announce("complex integrals No=4 Nv=11 NG=30")
    e = make_engine(gz, ei, ea)
    ints = cz.integrals_from_vertex(gz, 4)
    for cls, name in [(engine.PPHH, "PPHH"), (engine.PHPH, "PHPH"), (engine.HHHH, "HHHH"),
                      (engine.HHHP, "HHHP"), (engine.HHPP, "HHPP"), (engine.PHHP, "PHHP"),
                      (engine.PHHH, "PHHH")]:
        got = e.get_integrals(cls)
        ref = ints[name]
        assert got.dtype == np.complex128 and got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), name


@pytest.mark.parametrize("ppl,dist", [(True, False), (False, False), (True, True)])
def test_complex_residual_matches_oracle(ppl, dist):
    No, Nv, NG = 4, 13, 30
    gz, ei, ea = complex_system(5, No, Nv, NG)
    e = make_engine(gz, ei, ea)
    ints = cz.integrals_from_vertex(gz, No)
    T1, T2 = complex_amplitudes(Nv, No, seed=7)
    e.set_residual_variant(ppl, dist)
    R1, R2 = e.residual(0, T1, T2)
    assert np.abs(R2 - ints["PPHH"]).max() <= 1e-13 * np.abs(ints["PPHH"]).max()
    assert np.abs(R1).max() == 0.0
    for s in (1, 4, 13):
        announce("complex residual ppl=%d dist=%d slice=%d" % (ppl, dist, s))
        R1, R2 = e.residual(1, T1, T2, slice_size=s)
        O1, O2 = cz.residual(1, T1, T2, gz, ints, slice_size=s, ppl=ppl, distinguishable=dist)
        scale = float(np.abs(O2).max())
        assert float(np.abs(R2 - O2).max()) <= 1e-11 * scale, s
        assert float(np.abs(R1 - O1).max()) <= 1e-11 * float(np.abs(O1).max()) + 1e-13 * scale


def test_complex_solve_matches_oracle_and_real_path():
    """Whole complex runs: per-iteration energies against the oracle's
    complex driver and against the real path on the untransformed vertex;
    energy and pair energies of the converged amplitudes."""
    from sisi4s_amd import engine
    No, Nv, NG = 5, 20, 80
    g, ei, ea = philox.system(7, No, Nv, NG)
    gz, _ = cz.gauge_transform(g.astype(complex), No, 8)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=100, slice_size=6,
              amplitudes_convergence=1e-8, energy_convergence=1e-10)
    announce("complex solve No=5 Nv=20 NG=80")
    e = make_engine(gz, ei, ea)
    rep = e.ccsd_solve(engine.default_opts(**kw))
    ref = cz.solve(gz, ei, ea, **kw)
    got = e.energies()
    assert rep["converged"] and ref["converged"] and len(got) == len(ref["energies"])
    assert np.abs(got - np.real(ref["energies"])).max() <= 1e-10
    assert abs(rep["energy_imag"]) <= 1e-10
    real = ccsd.solve(g, ei, ea, **kw)
    assert abs(rep["energy"] - real["energy"]) <= 1e-9
    T1, T2 = e.amplitudes()
    assert T1.dtype == np.complex128
    P = e.pair_energies()
    Pref = cz.pair_energy(T1, T2, cz.integrals_from_vertex(gz, No)["HHPP"])
    assert np.abs(P - Pref).max() <= 1e-11 * np.abs(Pref).max()
    assert abs(P.sum() - (rep["energy"] + 1j * rep["energy_imag"])) <= 1e-10


@pytest.mark.parametrize("gauge", [False, True])
def test_complex_ueg_plane_wave_known_answers(gauge):
    """The plane-wave UEG vertex through the complex path on the GPU: the
    reference's MP2 and ten-iteration CCSD energies
    (integration-tests/tests/cc4s/ueg/rs1.0-7occ-26virt/cc4s.correct.out.yaml:
    128, 153)."""
    from sisi4s_amd import engine
    pw = ueg.plane_wave(7, 26, 1.0)
    G = pw["coeff"].astype(complex)
    if gauge:
        G, _ = cz.gauge_transform(G, 7, 2)
    announce("complex ueg plane wave gauge=%d" % gauge)
    e = make_engine(G, pw["eps"][:7], pw["eps"][7:])
    assert abs(e.mp2_energy() - -0.36143028565969504) <= 1e-12
    rep = e.ccsd_solve(engine.default_opts(mixer="DiisMixer", max_residua=4, max_iterations=10,
                                           energy_convergence=1e-8, amplitudes_convergence=1e-8))
    assert abs(rep["energy"] - -0.39269658954585018) <= 1e-12
    assert abs(rep["energy_imag"]) <= 1e-12


def test_complex_mixed_precision_and_refusals():
    """Bulk BF16x3 iterations then FP64 on the complex path reach the FP64
    energy; refinement residuals refuse a complex vertex."""
    from sisi4s_amd import engine
    gz, ei, ea = complex_system(7, 5, 20, 80)
    e = make_engine(gz, ei, ea)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=120, slice_size=6,
              amplitudes_convergence=1e-8, energy_convergence=1e-10)
    announce("complex mixed precision")
    r64 = e.ccsd_solve(engine.default_opts(**kw))
    rm = e.ccsd_solve(engine.default_opts(precision="bf16x3", switch_threshold=1e-5, **kw))
    assert r64["converged"] and rm["converged"]
    assert abs(rm["energy"] - r64["energy"]) <= 1e-6
    with pytest.raises(engine.EngineError):
        e.ccsd_solve(engine.default_opts(precision="bf16x3", refine_interval=8, **kw))


@pytest.mark.parametrize("No,Nv,NG", [(3, 8, 20), (5, 13, 40)])
def test_complex_triples_matches_oracle(No, Nv, NG):
    """CcsdPerturbativeTriplesComplex on the GPU against the oracle's literal
    restatement, with complex amplitudes on a complex vertex (all blocks,
    including those with two equal indices), and its reduced-precision
    contractions."""
    gz, ei, ea = complex_system(13, No, Nv, NG)
    e = make_engine(gz, ei, ea)
    T1, T2 = complex_amplitudes(Nv, No, seed=17, scale=0.05)
    from oracle import triples_complex as tz
    announce("complex (T) No=%d Nv=%d" % (No, Nv))
    ref = tz.triples_energy(T1, T2, gz, ei, ea).real
    got = e.triples(T1, T2)
    assert abs(got - ref) <= 1e-11 * abs(ref), (got, ref)
    assert abs(e.triples(T1, T2, precision="fp32") - ref) <= 1e-5 * abs(ref)
    assert abs(e.triples(T1, T2, precision="bf16x3") - ref) <= 1e-4 * abs(ref)


def test_complex_ueg_plane_wave_triples_known_answer():
    """(T) of the plane-wave UEG vertex on the engine's ten-iteration complex
    amplitudes: the reference's value
    (integration-tests/tests/cc4s/ueg/rs1.0-7occ-26virt/cc4s.correct.out.yaml:169)."""
    from sisi4s_amd import engine
    pw = ueg.plane_wave(7, 26, 1.0)
    G, _ = cz.gauge_transform(pw["coeff"].astype(complex), 7, 4)
    announce("complex ueg plane wave (T)")
    e = make_engine(G, pw["eps"][:7], pw["eps"][7:])
    e.ccsd_solve(engine.default_opts(mixer="DiisMixer", max_residua=4, max_iterations=10,
                                     energy_convergence=1e-8, amplitudes_convergence=1e-8))
    assert abs(e.triples() - -0.0063019625641725016) <= 1e-13


def test_complex_factor_ladder_matches_oracle_and_vertex_ladder():
    """Factor-based ladder for complex amplitudes with OutgoingFactorOrbitals
    (ClusterSinglesDoublesAlgorithm.cxx:745-880): against the oracle and the
    engine's own vertex ladder on Gamma = sum_R Lambda Po Pi (Po = conj(Pi),
    Lambda real), and a whole run."""
    from sisi4s_amd import engine
    rng = np.random.default_rng(21)
    No, Nv, NG, NR = 4, 12, 16, 30

    def cn(*s):
        return rng.normal(size=s) + 1j * rng.normal(size=s)

    # a vertex Hermitian in (q,r), as the vertex ladder's slicing assumes
    Lam = rng.normal(size=(NG, NR)) / np.sqrt(2 * NR * NG)
    Pi = cn(No + Nv, NR) / 3
    Po = Pi.conj()
    g = np.einsum("GR,qR,rR->Gqr", Lam, Po, Pi)
    ei = np.sort(rng.uniform(-2.0, -0.5, No))
    ea = np.sort(rng.uniform(0.2, 3.0, Nv))
    e = make_engine(g, ei, ea)
    ints = cz.integrals_from_vertex(g, No)
    T1, T2 = complex_amplitudes(Nv, No, seed=22, scale=0.05)
    V1, V2 = e.residual(1, T1, T2, slice_size=5)
    for fs in (7, 30):
        announce("complex factor ladder slice=%d" % fs)
        e.set_coulomb_factors(Lam, Pi, slice_size=fs, PiOut=Po)
        R1, R2 = e.residual(1, T1, T2, slice_size=5)
        O1, O2 = cz.residual(1, T1, T2, g, ints, 5, factors=(Lam, Pi, Po, fs))
        scale = float(np.abs(O2).max())
        assert float(np.abs(R2 - O2).max()) <= 1e-11 * scale
        assert float(np.abs(R1 - O1).max()) <= 1e-11 * scale
        assert float(np.abs(R2 - V2).max()) <= 1e-11 * scale
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=80, slice_size=5,
              amplitudes_convergence=1e-9, energy_convergence=1e-11)
    rep = e.ccsd_solve(engine.default_opts(**kw))
    ref = cz.solve(g, ei, ea, factors=(Lam, Pi, Po, 30), **kw)
    assert rep["converged"] and ref["converged"]
    assert np.abs(e.energies() - np.real(ref["energies"])).max() <= 1e-10
