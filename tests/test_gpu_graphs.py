"""Residuals replayed from a captured hipGraph (opts.graphs = 1) against the
eager residual path: the replay runs the same kernels on the same inputs,
so every per-iteration energy agrees to 1e-13 Eh, over many replays, across
precision switches (recapture), vertex changes (graph dropped), the complex,
factor and full-integral paths, and the batched solver."""

import numpy as np
import pytest

from conftest import announce
from oracle import ccsd, ccsd_complex as cz, philox

pytestmark = pytest.mark.gpu

TOL = 1e-13


def run(e, **kw):
    from sisi4s_amd import engine
    rep = e.ccsd_solve(engine.default_opts(**kw))
    return rep, e.energies()


def fixed_iterations(n, **kw):
    return dict(mixer="DiisMixer", max_residua=6, max_iterations=n, amplitudes_convergence=0.0,
                energy_convergence=0.0, **kw)


def test_graphed_fp64_run_matches_eager_and_survives_reuse():
    from sisi4s_amd import engine
    No, Nv, NG = 4, 14, 56
    g, ei, ea = philox.system(3, No, Nv, NG)
    e = engine.Engine(No, Nv, NG)
    e.set_system(g, ei, ea)
    kw = fixed_iterations(15, slice_size=5)
    announce("graphs fp64 eager reference")
    _, E0 = run(e, **kw)
    for rep_no in range(2):
        announce("graphs fp64 graphed run %d" % rep_no)
        rep, E1 = run(e, graphs=1, **kw)
        assert rep["iterations"] == 15
        assert np.abs(E1 - E0).max() <= TOL, np.abs(E1 - E0).max()
        announce("graphs fp64 eager after graphed run %d" % rep_no)
        _, E2 = run(e, **kw)
        assert np.abs(E2 - E0).max() <= TOL
    # a new vertex on the same engine drops the graph; the next graphed run
    # follows the new system
    g2, ei2, ea2 = philox.system(4, No, Nv, NG)
    e.set_system(g2, ei2, ea2)
    announce("graphs fp64 after vertex change")
    _, F0 = run(e, **kw)
    _, F1 = run(e, graphs=1, **kw)
    assert np.abs(F1 - F0).max() <= TOL
    assert np.abs(F0 - E0).max() > 1e-6
    ref = ccsd.solve(g2, ei2, ea2, slice_size=5, **fixed_iterations(15))
    assert np.abs(F1 - np.array(ref["energies"])).max() <= 1e-10


def test_graphed_mixed_schedule_and_linear_mixer_match_eager():
    """BF16x3 bulk iterations, then FP64: the precision switch recaptures."""
    from sisi4s_amd import engine
    No, Nv, NG = 5, 20, 80
    g, ei, ea = philox.system(7, No, Nv, NG)
    e = engine.Engine(No, Nv, NG)
    e.set_system(g, ei, ea)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=80, amplitudes_convergence=1e-8,
              energy_convergence=1e-10, precision="bf16x3", switch_threshold=1e-5)
    announce("graphs mixed eager")
    r0, E0 = run(e, **kw)
    announce("graphs mixed graphed")
    r1, E1 = run(e, graphs=1, **kw)
    assert r0["converged"] and r1["converged"]
    assert 0 < r1["fp64_iterations"] < r1["iterations"] == r0["iterations"]
    assert np.abs(E1 - E0).max() <= TOL
    # plain linear mixing diverges on the random model after a few iterations
    kl = dict(mixer="LinearMixer", mixing_ratio=0.7, max_iterations=6,
              amplitudes_convergence=0.0, energy_convergence=0.0)
    announce("graphs linear mixer")
    _, L0 = run(e, **kl)
    _, L1 = run(e, graphs=1, **kl)
    assert np.abs(L1 - L0).max() <= TOL


def test_graphed_variant_complex_factor_and_full_integral_paths():
    from sisi4s_amd import engine
    No, Nv, NG = 4, 12, 40
    g, ei, ea = philox.system(11, No, Nv, NG)
    kw = fixed_iterations(10, slice_size=5)
    # residual variants (distinguishable cluster, no ladder)
    e = engine.Engine(No, Nv, NG)
    e.set_system(g, ei, ea)
    announce("graphs variants")
    for ppl, dist in ((1, 1), (0, 0)):
        _, A = run(e, ppl=ppl, distinguishable=dist, **kw)
        _, B = run(e, graphs=1, ppl=ppl, distinguishable=dist, **kw)
        assert np.abs(A - B).max() <= TOL
    # complex amplitudes on a gauge-transformed vertex
    gz, _ = cz.gauge_transform(g.astype(complex), No, 12)
    ez = engine.Engine(No, Nv, NG)
    ez.set_system(gz, ei, ea)
    assert ez.complex
    announce("graphs complex")
    _, A = run(ez, **kw)
    _, B = run(ez, graphs=1, **kw)
    assert np.abs(A - B).max() <= TOL
    # factor-based ladder
    rng = np.random.default_rng(13)
    NR = 24
    Lam = (rng.normal(size=(NG, NR)) / np.sqrt(NR * NG)).astype(complex)
    Pi = (rng.normal(size=(No + Nv, NR)) / 2).astype(complex)
    gf = np.einsum("GR,pR,qR->Gpq", Lam.real, Pi.real, Pi.real)
    ef = engine.Engine(No, Nv, NG)
    ef.set_system(gf, ei, ea)
    ef.set_coulomb_factors(Lam, Pi, slice_size=7)
    announce("graphs factor ladder")
    _, A = run(ef, **kw)
    _, B = run(ef, graphs=1, **kw)
    assert np.abs(A - B).max() <= TOL
    # full-integral formulation
    ints = ccsd.integrals_from_vertex(g, No)
    Vabcd, Vabci = ccsd.vabcd_from_vertex(g, No)
    ei_ = engine.Engine(No, Nv, 1)
    ei_.set_eigenenergies(ei, ea)
    for cls, V in [(engine.PPHH, ints["PPHH"]), (engine.PHPH, ints["PHPH"]),
                   (engine.HHHH, ints["HHHH"]), (engine.HHHP, ints["HHHP"]),
                   (engine.PPPP, Vabcd), (engine.PPPH, Vabci)]:
        ei_.set_integrals(cls, V)
    ei_.set_formulation(engine.FORMULATION_INTEGRALS)
    announce("graphs full-integral formulation")
    _, A = run(ei_, **kw)
    _, B = run(ei_, graphs=1, **kw)
    assert np.abs(A - B).max() <= TOL


def test_graphed_batched_solve_matches_eager():
    """Batched small systems (one engine and one graph per system)."""
    from sisi4s_amd import engine
    rng = np.random.default_rng(5)
    systems = []
    for s in range(6):
        No, Nv = int(rng.integers(3, 7)), int(rng.integers(10, 25))
        systems.append(philox.system(500 + s, No, Nv, 4 * Nv))
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=60, amplitudes_convergence=1e-8,
              energy_convergence=1e-10)
    announce("graphs batched eager")
    E0, r0 = engine.ccsd_solve_batched(systems, engine.default_opts(**kw), workers=3)
    announce("graphs batched graphed")
    E1, r1 = engine.ccsd_solve_batched(systems, engine.default_opts(graphs=1, **kw), workers=3)
    assert [r["iterations"] for r in r0] == [r["iterations"] for r in r1]
    assert np.abs(E1 - E0).max() <= TOL


def test_graphed_direct_ring_and_spin_orbital_residuals():
    from sisi4s_amd import engine
    from oracle import uccsd
    No, Nv, NG = 4, 12, 40
    g, ei, ea = philox.system(17, No, Nv, NG)
    kw = fixed_iterations(8)
    e = engine.Engine(No, Nv, NG)
    e.set_system(g, ei, ea)
    e.set_drccd(adjacent_pairs_exchange=True)
    announce("graphs direct-ring CCD")
    _, A = run(e, **kw)
    _, B = run(e, graphs=1, **kw)
    assert np.abs(A - B).max() <= TOL
    o, v = 2, 5
    gs, eis, eas = philox.system(18, o, v, 14)
    ints = uccsd.spin_orbital_integrals(gs, o, v)
    u = engine.Engine(2 * o, 2 * v, 1)
    u.set_eigenenergies(np.repeat(eis, 2), np.repeat(eas, 2))
    for c, V in ints.items():
        u.set_integrals(getattr(engine, c), V)
    u.set_formulation(engine.FORMULATION_UCCSD)
    announce("graphs spin-orbital CCSD")
    _, A = run(u, **kw)
    _, B = run(u, graphs=1, **kw)
    assert np.abs(A - B).max() <= TOL
