"""CCSD parity: the HIP path through the C ABI against the FP64 oracle on the
same seeded inputs.  Tolerances: FP64 path energies to 1e-8 Eh (the No=10, Nv=50, NG=300
target), residual elements to 1e-11 relative to the residual's max norm."""

import numpy as np
import pytest

from conftest import announce, random_amplitudes
from oracle import ccsd, formulations, philox

pytestmark = pytest.mark.gpu


def make_engine(seed, No, Nv, NG):
    from sisi4s_amd import engine
    g, ei, ea = philox.system(seed, No, Nv, NG)
    e = engine.Engine(No, Nv, NG)
    e.set_system(g, ei, ea)
    return e, g, ei, ea


def test_integrals_match_oracle():
    from sisi4s_amd import engine
    announce("integrals No=10 Nv=50 NG=300")
    e, g, ei, ea = make_engine(20261202, 10, 50, 300)
    ints = ccsd.integrals_from_vertex(g, 10)
    for cls, name in [(engine.PPHH, "PPHH"), (engine.PHPH, "PHPH"), (engine.HHHH, "HHHH"),
                      (engine.HHHP, "HHHP"), (engine.HHPP, "HHPP")]:
        got = e.get_integrals(cls)
        ref = ints[name]
        assert got.shape == ref.shape
        err, scale = float(np.abs(got - ref).max()), float(np.abs(ref).max())
        assert err <= 1e-13 * scale, (name, err, scale)


@pytest.mark.parametrize("No,Nv,NG,slices", [(4, 9, 20, (1, 2, 4, 9)), (3, 13, 17, (5, 13)),
                                             (10, 50, 300, (10, 16))])
def test_residual_matches_oracle(No, Nv, NG, slices):
    e, g, ei, ea = make_engine(31 + No, No, Nv, NG)
    ints = ccsd.integrals_from_vertex(g, No)
    T1, T2 = random_amplitudes(Nv, No, seed=No + Nv)
    announce("residual it=0 No=%d Nv=%d NG=%d" % (No, Nv, NG))
    R1, R2 = e.residual(0, T1, T2)
    from sisi4s_amd import engine
    np.testing.assert_array_equal(R2, e.get_integrals(engine.PPHH))
    assert float(np.abs(R2 - ints["PPHH"]).max()) <= 1e-13 * float(np.abs(ints["PPHH"]).max())
    assert np.abs(R1).max() == 0.0
    for s in slices:
        announce("residual it=1 No=%d Nv=%d NG=%d slice=%d" % (No, Nv, NG, s))
        R1, R2 = e.residual(1, T1, T2, slice_size=s)
        O1, O2 = ccsd.residual(1, T1, T2, g, ints, slice_size=s)
        scale = float(np.abs(O2).max())
        e2, e1 = float(np.abs(R2 - O2).max()), float(np.abs(R1 - O1).max())
        assert e2 <= 1e-11 * scale, (s, e2, scale)
        assert e1 <= 1e-11 * float(np.abs(O1).max()) + 1e-13 * scale, (s, e1)


def test_residual_initial_doubles_and_energy():
    No, Nv, NG = 4, 9, 20
    e, g, ei, ea = make_engine(5, No, Nv, NG)
    ints = ccsd.integrals_from_vertex(g, No)
    T1, T2 = random_amplitudes(Nv, No, seed=9)
    announce("residual initial doubles it=0")
    R1, R2 = e.residual(0, T1, T2, initial_doubles=True)
    O1, O2 = ccsd.residual(0, T1, T2, g, ints, initial_doubles=True)
    err = float(np.abs(R2 - O2).max())
    assert err <= 1e-11 * float(np.abs(O2).max()), err
    announce("energy")
    E, d, x = e.energy(T1, T2)
    Eo, do, xo = ccsd.energy(T1, T2, ints["PPHH"])
    assert abs(E - Eo) <= 1e-12 * abs(Eo) and abs(d - do) <= 1e-12 * abs(do)
    announce("mp2")
    assert abs(e.mp2_energy() - formulations.mp2_closed_form(g, ei, ea)) <= 1e-12 * abs(
        formulations.mp2_closed_form(g, ei, ea))


def test_no10_nv50_ng300_diis6_30_iterations():
    """No=10, Nv=50, NG=300, MP2 start, DIIS(6), 30 iterations;
    every iteration's energy within 1e-8 Eh of the oracle."""
    from sisi4s_amd import engine
    No, Nv, NG = 10, 50, 300
    e, g, ei, ea = make_engine(20261202, No, Nv, NG)
    opts = engine.default_opts(mixer="DiisMixer", max_residua=6, max_iterations=30,
                               amplitudes_convergence=1e-12, energy_convergence=1e-14)
    announce("No=10 Nv=50 NG=300 DIIS(6) 30 iterations")
    rep = e.ccsd_solve(opts)
    gpu = e.energies()
    ref = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_residua=6, max_iterations=30,
                     amplitudes_convergence=1e-12, energy_convergence=1e-14)
    assert rep["iterations"] == 30 and len(gpu) == 30
    diff = np.abs(gpu - np.array(ref["energies"]))
    print("No=10 Nv=50 energies gpu/oracle max |dE| = %.3e, final %.12f / %.12f"
          % (diff.max(), gpu[-1], ref["energies"][-1]), flush=True)
    assert float(diff.max()) <= 1e-8, list(diff)
    T1, T2 = e.amplitudes()
    err = float(np.abs(T2 - ref["T2"]).max())
    assert err <= 1e-8 * float(np.abs(ref["T2"]).max()), err


def test_pair_energies_match_oracle_and_sum_to_energy():
    """PairEnergy (ClusterSinglesDoublesAlgorithm.cxx:218-235) of the
    converged No=10, Nv=50, NG=300 amplitudes: the oracle's pair energies, summing to the
    correlation energy."""
    from sisi4s_amd import engine
    e, g, ei, ea = make_engine(20261202, 10, 50, 300)
    announce("pair energies No=10 Nv=50")
    rep = e.ccsd_solve(engine.default_opts(mixer="DiisMixer", max_residua=6, max_iterations=12))
    T1, T2 = e.amplitudes()
    P = e.pair_energies()
    ref = ccsd.pair_energy(T1, T2, ccsd.integrals_from_vertex(g, 10)["PPHH"])
    assert np.abs(P - ref).max() <= 1e-11 * np.abs(ref).max()
    assert abs(P.sum() - rep["energy"]) <= 1e-10 * abs(rep["energy"])


@pytest.mark.parametrize("ppl,dist", [(False, False), (True, True), (False, True)])
def test_residual_variants_match_oracle(ppl, dist):
    """ppl / distinguishable (CcsdEnergyFromCoulombIntegrals.cxx:123-124,
    150-187, 244-246, 283-285, 316-338): the FP64 residual against the
    oracle, the refinement residual's exactness in FP64, and the refusal of
    a refinement under another variant than its anchor."""
    from sisi4s_amd import engine
    No, Nv, NG = 4, 13, 30
    e, g, ei, ea = make_engine(71, No, Nv, NG)
    ints = ccsd.integrals_from_vertex(g, No)
    T1, T2 = random_amplitudes(Nv, No, seed=9)
    e.set_residual_variant(ppl, dist)
    for s in (3, 13):
        announce("residual variant ppl=%d dist=%d slice=%d" % (ppl, dist, s))
        R1, R2 = e.residual(1, T1, T2, slice_size=s)
        O1, O2 = ccsd.residual(1, T1, T2, g, ints, slice_size=s, ppl=ppl, distinguishable=dist)
        scale = float(np.abs(O2).max())
        assert float(np.abs(R2 - O2).max()) <= 1e-11 * scale
        assert float(np.abs(R1 - O1).max()) <= 1e-11 * float(np.abs(O1).max()) + 1e-13 * scale
    D1, D2 = random_amplitudes(Nv, No, seed=10, scale=1e-3)
    e.residual_anchor(T1, T2, slice_size=3)
    e.set_precision("fp64")
    Q1, Q2 = e.residual_refine(T1 + D1, T2 + D2, slice_size=3)
    P1, P2 = e.residual(1, T1 + D1, T2 + D2, slice_size=3)
    scale = float(np.abs(P2).max())
    assert float(np.abs(Q2 - P2).max()) <= 1e-12 * scale
    assert float(np.abs(Q1 - P1).max()) <= 1e-12 * scale
    e.set_residual_variant(not ppl, dist)
    with pytest.raises(engine.EngineError):
        e.residual_refine(T1 + D1, T2 + D2, slice_size=3)


@pytest.mark.parametrize("ppl,dist", [(True, True), (False, False)])
def test_variant_solves_match_oracle_and_mixed_schedule(ppl, dist):
    """Whole runs of the variants: per-iteration energies against the oracle's
    driver (FP64), and the mixed schedule with refinement residuals within
    1e-6 Eh of the FP64 run."""
    from sisi4s_amd import engine
    No, Nv, NG = 5, 20, 80
    e, g, ei, ea = make_engine(7, No, Nv, NG)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=150,
              amplitudes_convergence=1e-8, energy_convergence=1e-10, slice_size=6)
    announce("variant solve ppl=%d dist=%d fp64" % (ppl, dist))
    r64 = e.ccsd_solve(engine.default_opts(ppl=int(ppl), distinguishable=int(dist), **kw))
    ref = ccsd.solve(g, ei, ea, ppl=ppl, distinguishable=dist, **kw)
    got = e.energies()
    assert r64["converged"] and ref["converged"] and len(got) == len(ref["energies"])
    assert np.abs(np.array(got) - ref["energies"]).max() <= 1e-10
    announce("variant solve ppl=%d dist=%d mixed" % (ppl, dist))
    rr = e.ccsd_solve(engine.default_opts(precision="bf16x3", switch_threshold=1e-5,
                                          refine_interval=8, ppl=int(ppl),
                                          distinguishable=int(dist), **kw))
    assert rr["converged"] and abs(rr["energy"] - r64["energy"]) <= 1e-6


def test_linear_mixer_and_convergence_flag():
    from sisi4s_amd import engine
    No, Nv, NG = 4, 9, 20
    e, g, ei, ea = make_engine(77, No, Nv, NG)
    opts = engine.default_opts(mixer="LinearMixer", mixing_ratio=0.7, max_iterations=6,
                               slice_size=3)
    announce("linear mixer")
    rep = e.ccsd_solve(opts)
    ref = ccsd.solve(g, ei, ea, mixer="LinearMixer", mixing_ratio=0.7, max_iterations=6,
                     slice_size=3)
    # this mixing diverges on the random model: compare relative to the oracle
    np.testing.assert_allclose(e.energies(), ref["energies"], rtol=1e-10, atol=1e-10)
    assert rep["converged"] == int(ref["converged"])


def test_initial_amplitudes_and_zero_iterations():
    from sisi4s_amd import engine
    No, Nv, NG = 4, 9, 20
    e, g, ei, ea = make_engine(78, No, Nv, NG)
    T1, T2 = random_amplitudes(Nv, No, seed=1, scale=0.02)
    announce("initial amplitudes, 3 iterations")
    opts = engine.default_opts(mixer="DiisMixer", max_iterations=3)
    e.ccsd_solve(opts, T1=T1, T2=T2)
    ref = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_iterations=3, initial=(T1, T2))
    np.testing.assert_allclose(e.energies(), ref["energies"], rtol=0, atol=1e-10)
    announce("maxIterations 0")
    rep = e.ccsd_solve(engine.default_opts(max_iterations=0), T1=T1, T2=T2)
    Eo = ccsd.energy(T1, T2, ccsd.integrals_from_vertex(g, No)["PPHH"])[0]
    assert abs(rep["energy"] - Eo) <= 1e-12 * abs(Eo)


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_reduced_precision_residual_error(prec, tol):
    No, Nv, NG = 10, 50, 300
    e, g, ei, ea = make_engine(20261202, No, Nv, NG)
    ints = ccsd.integrals_from_vertex(g, No)
    T1, T2 = random_amplitudes(Nv, No, seed=2, scale=0.02)
    e.set_precision(prec)
    announce("residual precision=%s" % prec)
    R1, R2 = e.residual(1, T1, T2)
    O1, O2 = ccsd.residual(1, T1, T2, g, ints)
    err, scale = float(np.abs(R2 - O2).max()), float(np.abs(O2).max())
    print("precision %s residual max rel error %.3e" % (prec, err / scale), flush=True)
    assert err <= tol * scale, (err, scale)
    e.set_precision("fp64")


def test_mixed_precision_solve_reaches_fp64_energy():
    """FP32/BF16x3 bulk iterations, FP64 refinement: the converged energy
    matches the FP64 run to 1 uEh."""
    from sisi4s_amd import engine
    No, Nv, NG = 5, 20, 80   # NG = 4 Nv: converges to 1e-8 in ~60 DIIS(6) iterations
    e, g, ei, ea = make_engine(7, No, Nv, NG)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=80,
              amplitudes_convergence=1e-8, energy_convergence=1e-10)
    announce("fp64 solve small")
    r64 = e.ccsd_solve(engine.default_opts(**kw))
    assert r64["converged"] == 1
    announce("bf16x3 mixed solve small")
    rmx = e.ccsd_solve(engine.default_opts(precision="bf16x3", switch_threshold=1e-5, **kw))
    assert rmx["converged"] == 1
    assert rmx["fp64_iterations"] < rmx["iterations"]
    print("mixed: %d iterations (%d fp64), fp64: %d iterations, |dE| = %.2e"
          % (rmx["iterations"], rmx["fp64_iterations"], r64["iterations"],
             abs(rmx["energy"] - r64["energy"])), flush=True)
    assert abs(rmx["energy"] - r64["energy"]) <= 1e-6


def uneven_small_systems(n, base=20261202, no=(3, 6), nv=(10, 24)):
    rng = np.random.default_rng(base)
    out = []
    for s in range(n):
        No = int(rng.integers(no[0], no[1] + 1))
        Nv = int(rng.integers(nv[0], nv[1] + 1))
        out.append((base + s, No, Nv, 4 * Nv))
    return out


def test_batched_ccsd_matches_per_system_oracle():
    """Batched CCSD over uneven small systems: each system's
    10-iteration DIIS(6) energy equals the oracle's to 1e-8 Eh."""
    from sisi4s_amd import engine
    specs = uneven_small_systems(8)
    systems = [philox.system(seed, No, Nv, NG) for seed, No, Nv, NG in specs]
    opts = engine.default_opts(mixer="DiisMixer", max_residua=6, max_iterations=10,
                               amplitudes_convergence=0.0, energy_convergence=0.0)
    announce("batched ccsd 8 systems, 4 workers")
    energies, reports = engine.ccsd_solve_batched(systems, opts, workers=4)
    for (seed, No, Nv, NG), (g, ei, ea), E, rep in zip(specs, systems, energies, reports):
        ref = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_residua=6, max_iterations=10,
                         amplitudes_convergence=0.0, energy_convergence=0.0)
        assert rep["iterations"] == 10
        assert abs(E - ref["energy"]) <= 1e-8, (seed, No, Nv, E, ref["energy"])


def test_batched_mixed_precision_within_1uEh():
    from sisi4s_amd import engine
    specs = uneven_small_systems(6, base=99)
    systems = [philox.system(seed, No, Nv, NG) for seed, No, Nv, NG in specs]
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=200,
              amplitudes_convergence=1e-8, energy_convergence=1e-10)
    announce("batched fp64 vs bf16x3 mixed, 6 systems")
    e64, r64 = engine.ccsd_solve_batched(systems, engine.default_opts(**kw), workers=3)
    emx, rmx = engine.ccsd_solve_batched(
        systems, engine.default_opts(precision="bf16x3", switch_threshold=1e-5, **kw), workers=3)
    assert all(r["converged"] for r in r64) and all(r["converged"] for r in rmx)
    dmax = float(np.max(np.abs(e64 - emx)))
    print("batched mixed vs fp64 max |dE| = %.2e" % dmax, flush=True)
    assert dmax <= 1e-6


def test_batched_reports_failure_of_one_system():
    from sisi4s_amd import engine
    g, ei, ea = philox.system(1, 3, 8, 32)
    bad = ea.copy()
    bad[0] = np.nan
    announce("batched with one invalid system")
    with pytest.raises(engine.EngineError) as ex:
        engine.ccsd_solve_batched([(g, ei, ea), (g, ei, bad)],
                                  engine.default_opts(max_iterations=2), workers=2)
    assert "system 1" in str(ex.value)


def test_library_backend_residuals():
    """With the hipBLASLt backend for the split-BF16x3 / bf16 products (the
    planner's permuted outputs go through its FP64 epilogue), the split-BF16x3
    residual and a plain-bf16 refinement residual keep their error bounds."""
    from sisi4s_amd import engine
    No, Nv, NG = 10, 50, 300
    e, g, ei, ea = make_engine(20261202, No, Nv, NG)
    TA1, TA2 = random_amplitudes(Nv, No, seed=2, scale=0.02)
    D1, D2 = random_amplitudes(Nv, No, seed=3, scale=2e-6)
    P1, P2 = e.residual(1, TA1 + D1, TA2 + D2)
    scale = float(np.abs(P2).max())
    engine.set_bf16_backend(True)
    try:
        announce("library backend residuals")
        e.set_precision("bf16x3")
        Q1, Q2 = e.residual(1, TA1 + D1, TA2 + D2)
        assert float(np.abs(Q2 - P2).max()) <= 1e-4 * scale
        e.set_precision("fp64")
        e.residual_anchor(TA1, TA2)
        e.set_precision("bf16")
        R1, R2 = e.residual_refine(TA1 + D1, TA2 + D2)
        assert float(np.abs(R2 - P2).max()) <= 1e-8 * scale
    finally:
        engine.set_bf16_backend(False)
        e.set_precision("fp64")


def test_refinement_residual_is_exact_in_fp64():
    """Anchor at T_A, refinement at T_A + d (FP64): equals the plain residual
    at T_A + d (the expansion is exact algebra)."""
    No, Nv, NG = 5, 17, 40
    e, g, ei, ea = make_engine(61, No, Nv, NG)
    TA1, TA2 = random_amplitudes(Nv, No, seed=21)
    D1, D2 = random_amplitudes(Nv, No, seed=22, scale=1e-3)
    announce("refinement residual fp64 exactness")
    e.residual_anchor(TA1, TA2, slice_size=4)
    e.set_precision("fp64")
    R1, R2 = e.residual_refine(TA1 + D1, TA2 + D2, slice_size=4)
    P1, P2 = e.residual(1, TA1 + D1, TA2 + D2, slice_size=4)
    scale = float(np.abs(P2).max())
    e2, e1 = float(np.abs(R2 - P2).max()), float(np.abs(R1 - P1).max())
    assert e2 <= 1e-12 * scale and e1 <= 1e-12 * scale, (e1, e2, scale)


def test_refinement_residual_reduced_precision_error_scales_with_change():
    No, Nv, NG = 10, 50, 300
    e, g, ei, ea = make_engine(20261202, No, Nv, NG)
    TA1, TA2 = random_amplitudes(Nv, No, seed=2, scale=0.02)
    D1, D2 = random_amplitudes(Nv, No, seed=3, scale=2e-6)
    announce("refinement residual bf16x3 error")
    P1, P2 = e.residual(1, TA1 + D1, TA2 + D2)          # FP64 reference
    e.residual_anchor(TA1, TA2)
    e.set_precision("bf16x3")
    R1, R2 = e.residual_refine(TA1 + D1, TA2 + D2)
    Q1, Q2 = e.residual(1, TA1 + D1, TA2 + D2)          # plain bf16x3
    e.set_precision("fp64")
    scale = float(np.abs(P2).max())
    err_ref, err_plain = float(np.abs(R2 - P2).max()), float(np.abs(Q2 - P2).max())
    print("refinement err %.2e, plain bf16x3 err %.2e (rel. to |R| %.2e)"
          % (err_ref / scale, err_plain / scale, scale), flush=True)
    assert err_ref <= 1e-3 * err_plain and err_ref <= 1e-9 * scale


@pytest.mark.parametrize("prec", ["bf16", "bf16x3"])
def test_refinement_residual_with_anchor_ladder_planes(prec):
    """Ladder slice pairs large enough for the bf16 kernels take V~(T_A) from
    the bf16 planes the FP64 anchor kept (No=16, Nv=160, slices of 40), with
    dV~ X(T) in the low-rank form from the anchor's S planes (the default) or
    from dV~ formed per slice pair in FP64 and split (ladder planes off): the
    refinement residual matches the FP64 residual to 1e-9 of its max norm in
    split-BF16x3 for a change of 2e-6 (a wrong plane layout would leave an
    error of the size of the change times |V~|), and in plain bf16 to 1e-6
    and a thousandth of a plain bf16 residual's error."""
    No, Nv, NG = 16, 160, 640
    e, g, ei, ea = make_engine(11, No, Nv, NG)
    TA1, TA2 = random_amplitudes(Nv, No, seed=4, scale=0.02)
    D1, D2 = random_amplitudes(Nv, No, seed=5, scale=2e-6)
    announce("refinement residual with anchor ladder planes, %s" % prec)
    P1, P2 = e.residual(1, TA1 + D1, TA2 + D2, slice_size=40)
    e.residual_anchor(TA1, TA2, slice_size=40)
    e.set_precision(prec)
    R1, R2 = e.residual_refine(TA1 + D1, TA2 + D2, slice_size=40)   # low-rank form
    e.set_ladder_planes(False)
    S1, S2 = e.residual_refine(TA1 + D1, TA2 + D2, slice_size=40)   # FP64 dV~, split
    e.set_ladder_planes(True)
    Q1, Q2 = e.residual(1, TA1 + D1, TA2 + D2, slice_size=40)  # plain reduced precision
    e.set_precision("fp64")
    scale = float(np.abs(P2).max())
    err, err_plain = float(np.abs(R2 - P2).max()), float(np.abs(Q2 - P2).max())
    err_s = float(np.abs(S2 - P2).max())
    print("%s refinement with planes: err %.2e (stored dV~ %.2e), plain %.2e of |R| %.2e"
          % (prec, err / scale, err_s / scale, err_plain / scale, scale), flush=True)
    tol = 1e-9 if prec == "bf16x3" else 1e-6
    for E1, E2 in ((R1, R2), (S1, S2)):
        assert float(np.abs(E2 - P2).max()) <= tol * scale
        assert float(np.abs(E2 - P2).max()) <= 1e-3 * err_plain
        assert float(np.abs(E1 - P1).max()) <= tol * scale


@pytest.mark.parametrize("scale", [2e-6, 1e-3])
def test_refinement_low_rank_with_ragged_slices(scale):
    """The low-rank refinement ladder (anchor S planes) with a ragged last
    slice (No=20, Nv=224, NG=900, slices of 64: 64+64+64+32) against the
    slice-pair form and the FP64 residual, for a small change and for a
    change large enough that the second-order D^T D term matters: both
    forms within the same bf16 bound of the FP64 residual."""
    No, Nv, NG = 20, 224, 900
    e, g, ei, ea = make_engine(31, No, Nv, NG)
    TA1, TA2 = random_amplitudes(Nv, No, seed=6, scale=0.02)
    D1, D2 = random_amplitudes(Nv, No, seed=7, scale=scale)
    announce("low-rank refinement, ragged slices, change %g" % scale)
    P1, P2 = e.residual(1, TA1 + D1, TA2 + D2, slice_size=64)
    e.residual_anchor(TA1, TA2, slice_size=64)
    e.set_precision("bf16")
    R1, R2 = e.residual_refine(TA1 + D1, TA2 + D2, slice_size=64)
    e.set_ladder_planes(False)
    S1, S2 = e.residual_refine(TA1 + D1, TA2 + D2, slice_size=64)
    e.set_ladder_planes(True)
    e.set_precision("fp64")
    A1, A2 = e.residual(1, TA1, TA2, slice_size=64)
    sc = float(np.abs(P2).max())
    dr = float(np.abs(P2 - A2).max())   # size of the residual change
    err, err_s = float(np.abs(R2 - P2).max()), float(np.abs(S2 - P2).max())
    print("change %g: |dR| %.2e, low-rank err %.2e, slice-pair err %.2e (of |R| %.2e)"
          % (scale, dr, err, err_s, sc), flush=True)
    assert err <= 1e-2 * dr and err <= 2.0 * err_s + 1e-12 * sc
    assert float(np.abs(R1 - P1).max()) <= 1e-2 * dr


def test_refined_mixed_solve_low_rank_matches_fp64():
    """A mixed run whose refinement residuals take the low-rank ladder (No=16,
    Nv=160, NG=640, slices of 40, anchors every 8) converges to the FP64
    energy within 1 uEh."""
    from sisi4s_amd import engine
    No, Nv, NG = 16, 160, 640
    e, g, ei, ea = make_engine(11, No, Nv, NG)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=200, slice_size=40,
              level_shift=2.0, amplitudes_convergence=1e-8, energy_convergence=1e-10)
    announce("low-rank refined mixed solve No=16 Nv=160 NG=640")
    r64 = e.ccsd_solve(engine.default_opts(**kw))
    rm = e.ccsd_solve(engine.default_opts(precision="bf16x3", switch_threshold=1e-5,
                                          refine_interval=8, refine_precision="bf16", **kw))
    print("low-rank mixed: fp64 %d it, mixed %d it (%d fp64), |dE| %.2e"
          % (r64["iterations"], rm["iterations"], rm["fp64_iterations"],
             abs(rm["energy"] - r64["energy"])), flush=True)
    assert r64["converged"] and rm["converged"]
    assert abs(rm["energy"] - r64["energy"]) <= 1e-6


def test_bulk_refined_mixed_solve_matches_fp64():
    """Bulk phase on split-BF16x3 anchors every 6 iterations with plain-bf16
    refinement residuals between, then FP64 anchors every 8 (No=16, Nv=160,
    NG=640, slices of 40: the low-rank ladder in both phases): the run
    converges to the FP64 energy within 1 uEh."""
    from sisi4s_amd import engine
    No, Nv, NG = 16, 160, 640
    e, g, ei, ea = make_engine(11, No, Nv, NG)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=200, slice_size=40,
              level_shift=2.0, amplitudes_convergence=1e-8, energy_convergence=1e-10)
    announce("bulk-refined mixed solve No=16 Nv=160 NG=640")
    r64 = e.ccsd_solve(engine.default_opts(**kw))
    rb = e.ccsd_solve(engine.default_opts(precision="bf16x3", switch_threshold=1e-5,
                                          refine_interval=8, refine_precision="bf16",
                                          bulk_refine_interval=6, **kw))
    print("bulk-refined mixed: fp64 %d it, mixed %d it (%d fp64), |dE| %.2e"
          % (r64["iterations"], rb["iterations"], rb["fp64_iterations"],
             abs(rb["energy"] - r64["energy"])), flush=True)
    assert r64["converged"] and rb["converged"]
    assert abs(rb["energy"] - r64["energy"]) <= 1e-6


def test_refined_mixed_solve_matches_fp64():
    from sisi4s_amd import engine
    No, Nv, NG = 5, 20, 80
    e, g, ei, ea = make_engine(7, No, Nv, NG)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=120,
              amplitudes_convergence=1e-8, energy_convergence=1e-10)
    announce("fp64 vs refined mixed solve")
    r64 = e.ccsd_solve(engine.default_opts(**kw))
    rr = e.ccsd_solve(engine.default_opts(precision="bf16x3", switch_threshold=1e-5,
                                          refine_interval=8, **kw))
    print("refined mixed: %d iterations (%d fp64) vs fp64 %d, |dE| %.2e"
          % (rr["iterations"], rr["fp64_iterations"], r64["iterations"],
             abs(rr["energy"] - r64["energy"])), flush=True)
    assert r64["converged"] and rr["converged"]
    assert rr["fp64_iterations"] <= rr["iterations"] // 4 + 2
    assert abs(rr["energy"] - r64["energy"]) <= 1e-6


def test_refined_mixed_solve_with_bf16_refinement_matches_fp64():
    """Refinement residuals in plain bf16 between FP64 anchors: their error
    scales with the change since the anchor, so the run still converges to
    the FP64 energy."""
    from sisi4s_amd import engine
    No, Nv, NG = 5, 20, 80
    e, g, ei, ea = make_engine(7, No, Nv, NG)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=150,
              amplitudes_convergence=1e-8, energy_convergence=1e-10)
    announce("fp64 vs bf16-refined mixed solve")
    r64 = e.ccsd_solve(engine.default_opts(**kw))
    rr = e.ccsd_solve(engine.default_opts(precision="bf16x3", switch_threshold=1e-5,
                                          refine_interval=8, refine_precision="bf16", **kw))
    print("bf16-refined mixed: %d iterations (%d fp64) vs fp64 %d, |dE| %.2e"
          % (rr["iterations"], rr["fp64_iterations"], r64["iterations"],
             abs(rr["energy"] - r64["energy"])), flush=True)
    assert r64["converged"] and rr["converged"]
    assert abs(rr["energy"] - r64["energy"]) <= 1e-6


def test_mixed_schedule_at_mid_size_matches_fp64():
    """No=20, Nv=200, NG=800: large enough that the ladder and ring products
    run in split-BF16x3 (bulk) and plain bf16 (refinement) rather than
    falling back to FP32; the mixed run reaches the FP64 energy to 1 uEh
    with FP64 only on the anchors."""
    from sisi4s_amd import engine
    No, Nv, NG = 20, 200, 800
    e, g, ei, ea = make_engine(20261202, No, Nv, NG)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=200, slice_size=32,
              level_shift=2.0, amplitudes_convergence=1e-8, energy_convergence=1e-10)
    announce("mid-size fp64 vs mixed (bf16x3 bulk, bf16 refinement)")
    r64 = e.ccsd_solve(engine.default_opts(**kw))
    rm = e.ccsd_solve(engine.default_opts(precision="bf16x3", switch_threshold=1e-5,
                                          refine_interval=8, refine_precision="bf16", **kw))
    print("mid-size: fp64 %d it, mixed %d it (%d fp64), |dE| %.2e"
          % (r64["iterations"], rm["iterations"], rm["fp64_iterations"],
             abs(rm["energy"] - r64["energy"])), flush=True)
    assert r64["converged"] and rm["converged"]
    assert rm["fp64_iterations"] <= rm["iterations"] // 4 + 2
    assert abs(rm["energy"] - r64["energy"]) <= 1e-6
    # with a plain-bf16 coarse stage down to |dT|/|T| = 1e-3 before the bulk
    rc = e.ccsd_solve(engine.default_opts(precision="bf16x3", switch_threshold=1e-5,
                                          refine_interval=8, refine_precision="bf16",
                                          coarse_threshold=1e-3, **kw))
    print("mid-size with coarse stage: %d it (%d fp64), |dE| %.2e"
          % (rc["iterations"], rc["fp64_iterations"], abs(rc["energy"] - r64["energy"])),
          flush=True)
    assert rc["converged"] and abs(rc["energy"] - r64["energy"]) <= 1e-6


@pytest.mark.gpu
def test_ueg_known_answers():
    """The reference's own known answers for the uniform electron gas, rs=1,
    No=7, Nv=26 (integration-tests/tests/cc4s/ueg/rs1.0-7occ-26virt/
    cc4s.correct.out.yaml:128 MP2, :153 CCSD, :169 (T)), with the vertex from
    the engine's generator: DIIS(4), ten iterations as the reference ran."""
    from sisi4s_amd import engine
    announce("ueg rs1 No7 Nv26 known answers")
    g, ei, ea = engine.generate_ueg(7, 26, 1.0)
    e = engine.Engine(7, 26, g.shape[0])
    try:
        e.set_system(g, ei, ea)
        assert abs(e.mp2_energy() - -0.36143028565969504) <= 1e-13
        rep = e.ccsd_solve(engine.default_opts(mixer="DiisMixer", max_residua=4, max_iterations=10,
                                               amplitudes_convergence=1e-8, energy_convergence=1e-8))
        assert rep["iterations"] == 10
        assert abs(rep["energy"] - -0.39269658954585018) <= 1e-12
        assert abs(e.triples() - -0.0063019625641725016) <= 1e-14
    finally:
        e.close()


def factor_system(seed, No, Nv, NG, NR, complex_factors):
    rng = np.random.default_rng(seed)
    Lam = rng.normal(size=(NG, NR)) / np.sqrt(NR * NG)
    Pi = rng.normal(size=(No + Nv, NR)) / 2
    g = np.einsum("GR,pR,qR->Gpq", Lam, Pi, Pi)
    if complex_factors:
        Lam = Lam + 0.3j * rng.normal(size=Lam.shape) / np.sqrt(NR * NG)
        Pi = Pi + 0.3j * rng.normal(size=Pi.shape) / 2
    ea = np.sort(rng.uniform(0.2, 3.0, Nv))
    ei = np.sort(rng.uniform(-2.0, -0.5, No))
    return g, ei, ea, Lam.astype(complex), Pi.astype(complex)


@pytest.mark.parametrize("complex_factors", [False, True])
def test_factor_ladder_matches_oracle_and_vertex_ladder(complex_factors):
    """CoulombFactors / FactorOrbitals (ClusterSinglesDoublesAlgorithm.cxx:
    567-672): the residual with the factor-based ladder against the oracle
    for several factor slice sizes; with real factors of the vertex it is the
    vertex-ladder residual; a whole run with the factors; removing them
    restores the vertex ladder."""
    from sisi4s_amd import engine
    No, Nv, NG, NR = 4, 12, 16, 30
    g, ei, ea, Lam, Pi = factor_system(5, No, Nv, NG, NR, complex_factors)
    e = engine.Engine(No, Nv, NG)
    e.set_system(g, ei, ea)
    ints = ccsd.integrals_from_vertex(g, No)
    T1, T2 = random_amplitudes(Nv, No, seed=6, scale=0.05)
    V1, V2 = e.residual(1, T1, T2, slice_size=5)
    for fs in (7, 30):
        announce("factor ladder complex=%d slice=%d" % (complex_factors, fs))
        e.set_coulomb_factors(Lam, Pi, slice_size=fs)
        R1, R2 = e.residual(1, T1, T2, slice_size=5)
        O1, O2 = ccsd.residual(1, T1, T2, g, ints, 5, factors=(Lam, Pi, fs))
        scale = float(np.abs(O2).max())
        assert float(np.abs(R2 - O2).max()) <= 1e-11 * scale
        assert float(np.abs(R1 - O1).max()) <= 1e-11 * scale
        if not complex_factors:
            assert float(np.abs(R2 - V2).max()) <= 1e-11 * scale
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=60, slice_size=5,
              amplitudes_convergence=1e-9, energy_convergence=1e-11)
    rep = e.ccsd_solve(engine.default_opts(**kw))
    ref = ccsd.solve(g, ei, ea, factors=(Lam, Pi, 30), **kw)
    assert rep["converged"] and ref["converged"]
    assert np.abs(e.energies() - np.array(ref["energies"])).max() <= 1e-10
    with pytest.raises(engine.EngineError):
        e.ccsd_solve(engine.default_opts(precision="bf16x3", refine_interval=8, **kw))
    e.set_coulomb_factors()
    R1, R2 = e.residual(1, T1, T2, slice_size=5)
    assert np.array_equal(R2, V2)


def test_full_integral_formulation_matches_oracle_and_vertex_path():
    """CcsdEnergyFromCoulombIntegralsReference (cxx:29-295) on an engine that
    only holds uploaded integrals: PPPP/PPPH built from the vertex match the
    oracle, the full-integral residual and its onlyPPL branch match the
    oracle's restatements, and a whole run gives the vertex formulation's
    energies."""
    from sisi4s_amd import engine
    from oracle import formulations
    No, Nv, NG = 4, 11, 30
    ev, g, ei, ea = make_engine(41, No, Nv, NG)
    Vabcd, Vabci = ccsd.vabcd_from_vertex(g, No)
    announce("full-integral formulation No=4 Nv=11")
    assert np.abs(ev.get_integrals(engine.PPPP) - Vabcd).max() <= 1e-13 * np.abs(Vabcd).max()
    assert np.abs(ev.get_integrals(engine.PPPH) - Vabci).max() <= 1e-13 * np.abs(Vabci).max()
    ints = ccsd.integrals_from_vertex(g, No)
    ef = engine.Engine(No, Nv, 1)
    ef.set_eigenenergies(ei, ea)
    for cls, V in [(engine.PPHH, ints["PPHH"]), (engine.PHPH, ints["PHPH"]),
                   (engine.HHHH, ints["HHHH"]), (engine.HHHP, ints["HHHP"]),
                   (engine.PPPP, Vabcd), (engine.PPPH, Vabci)]:
        ef.set_integrals(cls, V)
    ef.set_formulation(engine.FORMULATION_INTEGRALS)
    T1, T2 = random_amplitudes(Nv, No, seed=12)
    R1, R2 = ef.residual(1, T1, T2)
    O1, O2 = formulations.residual_full_integrals(1, T1, T2, g, No)
    scale = float(np.abs(O2).max())
    assert float(np.abs(R2 - O2).max()) <= 1e-11 * scale
    assert float(np.abs(R1 - O1).max()) <= 1e-11 * scale
    ef.set_formulation(engine.FORMULATION_ONLY_PPL)
    R1, R2 = ef.residual(0, T1, T2)
    _, P2 = formulations.residual_only_ppl(T1, T2, g, No)
    assert float(np.abs(R2 - P2).max()) <= 1e-11 * float(np.abs(P2).max()) and not R1.any()
    ef.set_formulation(engine.FORMULATION_INTEGRALS)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=120, amplitudes_convergence=1e-9,
              energy_convergence=1e-11)
    rf = ef.ccsd_solve(engine.default_opts(**kw))
    rv = ev.ccsd_solve(engine.default_opts(**kw))
    assert rf["converged"] and rv["converged"]
    assert np.abs(ef.energies() - ev.energies()).max() <= 1e-10


@pytest.mark.parametrize("unrestricted,anti,fock", [(True, False, False), (False, True, False),
                                                    (False, False, True), (True, False, True)])
def test_energy_variants_match_oracle(unrestricted, anti, fock):
    """getEnergy's arguments (ClusterSinglesDoublesAlgorithm.cxx:131-226):
    `unrestricted` (spins = 1), `antisymmetrize`, and the HPFockMatrix term:
    every iteration's energy within 1e-10 Eh of the oracle, pair energies to
    1e-11 relative, and the non-canonical term itself."""
    from sisi4s_amd import engine
    No, Nv, NG = 5, 20, 80
    e, g, ei, ea = make_engine(7, No, Nv, NG)
    f = 0.05 * np.random.default_rng(2).standard_normal((No, Nv)) if fock else None
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=12, amplitudes_convergence=0.0,
              energy_convergence=0.0)
    announce("energy variants unrestricted=%d antisymmetrize=%d fock=%d" % (unrestricted, anti, fock))
    try:
        e.set_energy_variant(unrestricted, anti)
        if fock:
            e.set_hp_fock(f)
        e.ccsd_solve(engine.default_opts(**kw))
        got = e.energies()
        spins = 1.0 if unrestricted else 2.0
        ref = ccsd.solve(g, ei, ea, spins=spins, antisymmetrized=anti, fia=f, **kw)
        assert float(np.abs(got - np.array(ref["energies"])).max()) <= 1e-10
        T1, T2 = e.amplitudes()
        V = ccsd.integrals_from_vertex(g, No)["PPHH"]
        P = e.pair_energies()
        Pr = ccsd.pair_energy(T1, T2, V, antisymmetrized=anti)
        assert np.abs(P - Pr).max() <= 1e-11 * np.abs(Pr).max()
        if fock:
            nc = spins * float(np.sum(T1 * f.T))
            assert abs(e.noncanonical_energy() - nc) <= 1e-12 * max(abs(nc), 1e-300)
    finally:
        e.set_energy_variant(False, False)
        e.set_hp_fock(None)


@pytest.mark.parametrize("prec", ["bf16x3", "bf16"])
def test_ladder_plane_output_matches_stored_slice_pairs(prec):
    """The reduced-precision ladder with plane output (the V~ product writes
    the bf16 planes the X contraction reads) against the same residual with
    the FP64 slice pair stored and split: the same products at the same
    precision, so equal up to summation order (1e-6 of the residual scale;
    a layout error would be O(1)), and both within the precision's bound of
    the FP64 residual.  No=16, Nv=96, NG=384, slices of 48 (every pair large
    enough for the plane path)."""
    from sisi4s_amd import engine
    No, Nv, NG = 16, 96, 384
    e, g, ei, ea = make_engine(20261203, No, Nv, NG)
    T1, T2 = random_amplitudes(Nv, No, seed=4, scale=0.02)
    announce("ladder plane output %s No=16 Nv=96 NG=384" % prec)
    R1f, R2f = e.residual(1, T1, T2, slice_size=48)
    e.set_precision(prec)
    e.set_ladder_planes(True)
    R1p, R2p = e.residual(1, T1, T2, slice_size=48)
    e.set_ladder_planes(False)
    R1s, R2s = e.residual(1, T1, T2, slice_size=48)
    e.set_ladder_planes(True)
    e.set_precision("fp64")
    scale = float(np.abs(R2f).max())
    d = float(np.abs(R2p - R2s).max())
    print("%s plane vs stored ladder: max |dR2| = %.3e of %.3e (bitwise %s)"
          % (prec, d, scale, bool(np.array_equal(R2p, R2s))), flush=True)
    assert d <= 1e-6 * scale
    assert float(np.abs(R1p - R1s).max()) <= 1e-6 * float(np.abs(R1f).max())
    tol = 2e-5 if prec == "bf16x3" else 2e-2
    assert float(np.abs(R2p - R2f).max()) <= tol * scale
