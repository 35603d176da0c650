"""Pins the CPU oracle (test infrastructure) against independent checks:
numpy's Philox, the reference's full-integral CCSD formulation
(CcsdEnergyFromCoulombIntegralsReference.cxx), the unsliced 6-index (T)
(PerturbativeTriples.cxx:19-97), the MP2 closed form, DIIS identities,
symmetries and limits, and the reference's own known answers for the
uniform electron gas (rs=1, No=7, Nv=26)."""

import itertools

import numpy as np
import pytest

from conftest import random_amplitudes
from oracle import ccsd, formulations, philox, triples, ueg


def test_philox_matches_numpy_bit_generator():
    # numpy.random.Philox is Philox-4x64-10; it increments its counter
    # before each block, so counter c-1 yields philox(c).
    for key, ctr in [((5, 7), (10, 0, 0, 0)), ((2**63 + 11, 3), (123456789, 42, 7, 1)),
                     ((0, 0), (1, 0, 0, 0))]:
        bg = np.random.Philox(key=np.array(key, dtype=np.uint64),
                              counter=np.array((ctr[0] - 1,) + ctr[1:], dtype=np.uint64))
        ref = bg.random_raw(4)
        out = philox.philox4x64(tuple(np.uint64(c) for c in ctr), tuple(np.uint64(k) for k in key))
        assert [int(x) for x in out] == [int(x) for x in ref]


def test_generator_properties():
    g, ei, ea = philox.system(3, 5, 12, 40)
    assert g.shape == (40, 17, 17)
    np.testing.assert_array_equal(g, g.transpose(0, 2, 1))
    assert ei.min() >= -2.0 and ei.max() < -0.5
    assert ea.min() >= 0.2 and ea.max() < 3.0
    # N(0,1) scaled by exp(-G/100)/sqrt(NG): check the row variance trend
    z = g * np.sqrt(40.0) / np.exp(-np.arange(40) / 100.0)[:, None, None]
    assert abs(z.std() - 1.0) < 0.05 and abs(z.mean()) < 0.05


@pytest.mark.parametrize("No,Nv,NG", [(3, 7, 11), (4, 9, 20)])
def test_residual_matches_full_integral_formulation(No, Nv, NG):
    g, ei, ea = philox.system(11, No, Nv, NG)
    T1, T2 = random_amplitudes(Nv, No, seed=No)
    ints = ccsd.integrals_from_vertex(g, No)
    F1, F2 = formulations.residual_full_integrals(1, T1, T2, g, No)
    scale = np.abs(F2).max()
    for s in sorted({1, 2, 3, No, Nv}):
        R1, R2 = ccsd.residual(1, T1, T2, g, ints, slice_size=s)
        assert np.abs(R1 - F1).max() <= 1e-12 * scale
        assert np.abs(R2 - F2).max() <= 1e-12 * scale


def test_residual_planted_error_is_detected():
    # A dropped PPL transpose (the a > b branch of sliceIntoResiduum) must be
    # caught by the full-integral comparison.
    No, Nv, NG = 3, 7, 11
    g, _, _ = philox.system(11, No, Nv, NG)
    T1, T2 = random_amplitudes(Nv, No, seed=2)
    ints = ccsd.integrals_from_vertex(g, No)
    orig = ccsd.slice_into_residuum

    def broken(Rxyij, a, b, R2):
        R2[a:a + Rxyij.shape[0], b:b + Rxyij.shape[1]] += Rxyij

    ccsd.slice_into_residuum = broken
    try:
        _, R2 = ccsd.residual(1, T1, T2, g, ints, slice_size=2)
    finally:
        ccsd.slice_into_residuum = orig
    _, F2 = formulations.residual_full_integrals(1, T1, T2, g, No)
    assert np.abs(R2 - F2).max() > 1e-3 * np.abs(F2).max()


def test_residual_limits_and_symmetry():
    No, Nv, NG = 3, 7, 11
    g, _, _ = philox.system(5, No, Nv, NG)
    ints = ccsd.integrals_from_vertex(g, No)
    zero1, zero2 = np.zeros((Nv, No)), np.zeros((Nv, Nv, No, No))
    # MP2 branch and the T = 0 limit both give R2 = V_abij, R1 = 0
    for it in (0, 3):
        R1, R2 = ccsd.residual(it, zero1, zero2, g, ints)
        np.testing.assert_allclose(R2, ints["PPHH"], rtol=0, atol=1e-14)
        assert np.abs(R1).max() == 0.0
    T1, T2 = random_amplitudes(Nv, No, seed=3)
    _, R2 = ccsd.residual(2, T1, T2, g, ints, slice_size=2)
    np.testing.assert_allclose(R2, R2.transpose(1, 0, 3, 2), rtol=0, atol=1e-13 * np.abs(R2).max())


@pytest.mark.parametrize("No,Nv,NG", [(3, 5, 9), (4, 6, 13)])
def test_triples_sliced_matches_unsliced(No, Nv, NG):
    g, ei, ea = philox.system(21, No, Nv, NG)
    T1, T2 = random_amplitudes(Nv, No, seed=4)
    e_sl = triples.triples_energy(T1, T2, g, ei, ea)
    e_un = formulations.triples_unsliced(T1, T2, g, ei, ea)
    assert abs(e_sl - e_un) <= 1e-12 * abs(e_un)


def test_triples_diagonal_blocks_vanish_and_relabel_invariance():
    No, Nv, NG = 3, 5, 9
    g, ei, ea = philox.system(22, No, Nv, NG)
    T1, T2 = random_amplitudes(Nv, No, seed=5)
    inp = triples.TriplesInputs(T1, T2, g, ei, ea)
    for k in range(No):
        assert abs(triples.block_energy(inp, (k, k, k))) < 1e-12
    e0 = triples.triples_energy(T1, T2, g, ei, ea)
    perm = np.array([2, 0, 1])
    gp = g.copy()
    gp[:, :No, :] = g[:, perm, :]
    gp2 = gp.copy()
    gp2[:, :, :No] = gp[:, :, perm]
    e1 = triples.triples_energy(T1[:, perm], T2[:, :, perm][:, :, :, perm], gp2, ei[perm], ea)
    assert abs(e1 - e0) <= 1e-12 * abs(e0)


def test_mp2_driver_matches_closed_form():
    g, ei, ea = philox.system(9, 4, 10, 30)
    e_drv = ccsd.mp2_energy(g, ei, ea)
    e_cf = formulations.mp2_closed_form(g, ei, ea)
    assert abs(e_drv - e_cf) <= 1e-13 * abs(e_cf)
    r = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_residua=4, max_iterations=1)
    assert abs(r["energies"][0] - e_cf) <= 1e-13 * abs(e_cf)


def test_pair_energies_sum_to_energy():
    # e_ij = sum_ab tau (2V - V') over pairs must add up to direct + exchange
    g, ei, ea = philox.system(4, 4, 7, 12)
    T1, T2 = random_amplitudes(7, 4, seed=1)
    V = ccsd.integrals_from_vertex(g, 4)["PPHH"]
    P = ccsd.pair_energy(T1, T2, V)
    E = ccsd.energy(T1, T2, V)[0]
    assert P.shape == (4, 4) and abs(P.sum() - E) <= 1e-12 * abs(E)
    # a pair energy only sees its own (i,j) amplitudes
    T2b = T2.copy()
    T2b[:, :, 1, 2] += 0.1
    Pb = ccsd.pair_energy(T1, T2b, V)
    changed = np.argwhere(np.abs(Pb - P) > 0)
    assert changed.tolist() == [[1, 2]]


def test_distinguishable_cluster_is_exact_for_two_electrons():
    """The distinguishable-cluster residual (distinguishable = 1) keeps CCSD's
    exactness for two electrons: at No = 1 both converge to the same (full
    CI) energy, while for more electrons the two models differ."""
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=80,
              amplitudes_convergence=1e-13, energy_convergence=1e-15)
    for Nv, NG in [(6, 20), (9, 30)]:
        g, ei, ea = philox.system(5, 1, Nv, NG)
        a = ccsd.solve(g, ei, ea, **kw)
        b = ccsd.solve(g, ei, ea, distinguishable=True, **kw)
        assert a["converged"] and b["converged"]
        assert abs(a["energy"] - b["energy"]) <= 1e-13 * abs(a["energy"])
    g, ei, ea = philox.system(5, 3, 8, 20)
    a = ccsd.solve(g, ei, ea, **kw)
    b = ccsd.solve(g, ei, ea, distinguishable=True, **kw)
    assert abs(a["energy"] - b["energy"]) > 1e-3


def test_distinguishable_changes_only_nonlinear_terms():
    # R_dist(lam T) - R_ccsd(lam T) has no constant or linear part in lam
    g, ei, ea = philox.system(8, 3, 7, 15)
    ints = ccsd.integrals_from_vertex(g, 3)
    T1, T2 = random_amplitudes(7, 3, seed=4)

    def diff(lam):
        a = ccsd.residual(1, lam * T1, lam * T2, g, ints, 2, distinguishable=True)
        b = ccsd.residual(1, lam * T1, lam * T2, g, ints, 2)
        return np.concatenate([(a[0] - b[0]).ravel(), (a[1] - b[1]).ravel()])

    d1, d2 = diff(1e-3), diff(2e-3)
    assert np.linalg.norm(d1) > 0
    assert abs(np.linalg.norm(d2) / np.linalg.norm(d1) - 4.0) < 0.05


@pytest.mark.parametrize("slice_size", [1, 3, 7])
def test_ppl0_drops_exactly_the_ladder(slice_size):
    """ppl = 0 leaves out the particle-particle ladder, written here with the
    stored PPPP/PPPH integrals instead of sliced dressed vertices:
    sum_cd (V_abcd - T_ak V_bkdc - T_bl V_adcl + T_ak T_bl V_cdkl) X_cdij,
    and moves the T1 T1 part of the hole-ladder term to the residual."""
    No, Nv, NG = 3, 7, 15
    g, ei, ea = philox.system(9, No, Nv, NG)
    ints = ccsd.integrals_from_vertex(g, No)
    Vabcd, Vabci = ccsd.vabcd_from_vertex(g, No)
    V = ints["PPHH"]
    T1, T2 = random_amplitudes(Nv, No, seed=6)
    X = T2 + np.einsum("ai,bj->abij", T1, T1)
    dressed = (Vabcd - np.einsum("ak,bcdk->abcd", T1, Vabci)
               - np.einsum("bl,adcl->abcd", T1, Vabci)
               + np.einsum("ak,bl,cdkl->abcd", T1, T1, V))
    ladder = np.einsum("abcd,cdij->abij", dressed, X)
    hole = np.einsum("klij,ak,bl->abij", np.einsum("cdkl,cdij->klij", V, X), T1, T1)
    R1a, R2a = ccsd.residual(1, T1, T2, g, ints, slice_size)
    R1b, R2b = ccsd.residual(1, T1, T2, g, ints, slice_size, ppl=False)
    assert np.abs(R1a - R1b).max() == 0.0
    err = np.abs(R2a - R2b - ladder + hole).max()
    assert err <= 1e-13 * np.abs(ladder).max(), err


def test_diis_weights():
    rng = np.random.default_rng(0)
    mix = ccsd.DiisMixer(3)
    vecs = [[rng.standard_normal(5), rng.standard_normal(7)] for _ in range(5)]
    res = [[rng.standard_normal(5), rng.standard_normal(7)] for _ in range(5)]
    mix.append(vecs[0], res[0])
    assert mix.weights == [1.0]
    np.testing.assert_allclose(mix.get()[0], vecs[0][0])
    for k in range(1, 5):
        mix.append(vecs[k], res[k])
        assert abs(sum(mix.weights) - 1.0) < 1e-12
    # brute force: minimise |sum_j w_j r_j|^2 subject to sum w = 1 over the
    # stored (ring) vectors, newest first
    stored = [res[4], res[3], res[2]]
    Rm = np.array([np.concatenate(r) for r in stored])
    Bm = Rm @ Rm.T
    w = np.linalg.solve(Bm, np.ones(3))
    w /= w.sum()
    np.testing.assert_allclose(mix.weights, w, rtol=1e-10, atol=1e-12)


def test_level_shift_keeps_the_fixed_point(monkeypatch):
    """estimateAmplitudesFromResiduum with a level shift s
    (ClusterSinglesDoublesAlgorithm.cxx:302-331) maps T to
    -(R - sT)/(D + s), whose fixed point T D = -R does not depend on s: the
    shifted and unshifted runs converge to the same energy, and the shifted
    update leaves the converged amplitudes unchanged.  A shift of the wrong
    sign (R + sT) has the fixed point T (D + 2s) = -R and is detected."""
    g, ei, ea = philox.system(7, 4, 12, 48)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=200,
              amplitudes_convergence=1e-11, energy_convergence=1e-13)
    r0 = ccsd.solve(g, ei, ea, level_shift=0.0, **kw)
    r2 = ccsd.solve(g, ei, ea, level_shift=2.0, **kw)
    assert r0["converged"] and r2["converged"]
    assert abs(r0["energy"] - r2["energy"]) <= 1e-10 * abs(r0["energy"])
    T1, T2 = r0["T1"], r0["T2"]
    R1, R2 = ccsd.residual(5, T1, T2, g, r0["integrals"])
    E1, E2 = ccsd.estimate_amplitudes(R1, R2, T1, T2, ei, ea, level_shift=2.0)
    assert np.abs(E2 - T2).max() <= 1e-8 * np.abs(T2).max()
    assert np.abs(E1 - T1).max() <= 1e-8 * np.abs(T2).max()

    def wrong_sign(R1, R2, T1, T2, eps_i, eps_a, level_shift=0.0):
        D1, D2 = ccsd.denominators(eps_i, eps_a)
        return (-(R1 + level_shift * T1) / (D1 + level_shift),
                -(R2 + level_shift * T2) / (D2 + level_shift))

    monkeypatch.setattr(ccsd, "estimate_amplitudes", wrong_sign)
    rw = ccsd.solve(g, ei, ea, level_shift=2.0, **kw)
    assert abs(rw["energy"] - r0["energy"]) > 1e-4 * abs(r0["energy"])


def test_golden_fixture_regression():
    d = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                           "tiny_system.npz"))
    No, Nv, NG = (int(x) for x in d["dims"])
    g, ei, ea = philox.system(int(d["seed"]), No, Nv, NG)
    np.testing.assert_allclose(g, d["gamma"], rtol=0, atol=1e-15)
    ints = ccsd.integrals_from_vertex(g, No)
    R1, R2 = ccsd.residual(1, d["T1"], d["T2"], g, ints, slice_size=int(d["slice"]))
    np.testing.assert_allclose(R1, d["R1"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(R2, d["R2"], rtol=1e-12, atol=1e-14)
    et = triples.triples_energy(d["T1"], d["T2"], g, ei, ea)
    assert abs(et - float(d["E_T"])) <= 1e-12 * abs(float(d["E_T"]))
    r = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_residua=6, max_iterations=8,
                   amplitudes_convergence=0.0, energy_convergence=0.0)
    np.testing.assert_allclose(r["energies"], d["energies"], rtol=1e-11)


# Known answers of integration-tests/tests/cc4s/ueg/rs1.0-7occ-26virt/
# cc4s.correct.out.yaml: MP2 (:128, secondOrder), CCSD (:153, correlation of
# the Ccsd step: DiisMixer, maxResidua 4, energyConvergence 1e-8) and (T) on
# those amplitudes (:169).  The reference's CCSD stopped after iteration 9
# (0-based), so the driver here runs exactly 10 iterations.
UEG_MP2 = -0.36143028565969504
UEG_CCSD = -0.39269658954585018
UEG_TRIPLES = -0.0063019625641725016


def test_ueg_vertex_properties():
    g, ei, ea, pw = ueg.system(7, 26, 1.0)
    assert g.shape == (257, 33, 33)
    np.testing.assert_array_equal(g, g.transpose(0, 2, 1))
    # HOMO/LUMO and Madelung constant as the reference prints them
    assert ei.max() < ea.min()
    # the rotation to real orbitals preserves every Coulomb integral:
    # sum_G Gamma_Gpr Gamma_Gqs (real) == U* U U* U (pr|qs) (plane waves)
    U = ueg.real_orbitals(pw["grid"], 7)
    c = pw["coeff"]
    t = pw["momenta"]
    inv = np.array([np.flatnonzero((t == -x).all(1))[0] for x in t])
    Vpw = np.einsum("Gpr,Gqs->prqs", c, c[inv])
    Vrot = np.einsum("Pp,Rr,Qq,Ss,prqs->PRQS", U.conj(), U, U.conj(), U, Vpw, optimize=True)
    Vreal = np.einsum("Gpr,Gqs->prqs", g, g)
    assert np.abs(Vrot.imag).max() < 1e-14
    np.testing.assert_allclose(Vreal, Vrot.real, rtol=0, atol=1e-14)


def test_ueg_known_answers():
    g, ei, ea, _ = ueg.system(7, 26, 1.0)
    assert abs(ccsd.mp2_energy(g, ei, ea) - UEG_MP2) < 1e-14
    r = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_residua=4, max_iterations=10,
                   amplitudes_convergence=1e-8, energy_convergence=1e-8)
    assert abs(r["energies"][0] - UEG_MP2) < 1e-14
    assert abs(r["energy"] - UEG_CCSD) < 1e-13
    et = triples.triples_energy(r["T1"], r["T2"], g, ei, ea)
    assert abs(et - UEG_TRIPLES) < 1e-15


def test_ueg_open_shell_is_rejected():
    with pytest.raises(ValueError):
        ueg.system(5, 28, 1.0)


# ---------------------------------------------------------------------------
# complex path (oracle/ccsd_complex.py)
# ---------------------------------------------------------------------------

def test_complex_path_equals_real_path_for_real_vertex():
    """On a real vertex symmetric in (p,q) the complex residual
    (CcsdEnergyFromCoulombIntegrals.cxx:496-891) and the real one
    (cxx:27-494) are the same equations."""
    from oracle import ccsd_complex as cz
    g, ei, ea = philox.system(5, 3, 8, 20)
    T1, T2 = random_amplitudes(8, 3, seed=3)
    ints, zints = ccsd.integrals_from_vertex(g, 3), cz.integrals_from_vertex(g.astype(complex), 3)
    for ppl, dist in [(True, False), (False, False), (True, True)]:
        R1, R2 = ccsd.residual(1, T1, T2, g, ints, 3, ppl=ppl, distinguishable=dist)
        Z1, Z2 = cz.residual(1, T1.astype(complex), T2.astype(complex), g.astype(complex),
                             zints, 3, ppl=ppl, distinguishable=dist)
        assert np.abs(Z2 - R2).max() <= 1e-13 * np.abs(R2).max()
        assert np.abs(Z1 - R1).max() <= 1e-13 * np.abs(R2).max()


def test_complex_path_is_gauge_covariant():
    """A random phase per orbital and a unitary mix of the auxiliary index
    make the vertex genuinely complex without changing the physics: the
    residual transforms like the amplitudes (phase e^{-i(t_a+t_b-t_i-t_j)})
    and every iteration's energy is the real path's, with zero imaginary
    part.  A conjugate dropped or added anywhere in the term list breaks
    this."""
    from oracle import ccsd_complex as cz
    No, Nv = 3, 8
    g, ei, ea = philox.system(5, No, Nv, 20)
    gz, t = cz.gauge_transform(g.astype(complex), No, 1)
    T1, T2 = random_amplitudes(Nv, No, seed=3)
    ph1 = np.exp(-1j * (t[No:, None] - t[None, :No]))
    ph2 = np.exp(-1j * (t[No:, None, None, None] + t[None, No:, None, None]
                        - t[None, None, :No, None] - t[None, None, None, :No]))
    R1, R2 = ccsd.residual(1, T1, T2, g, ccsd.integrals_from_vertex(g, No), 3)
    Z1, Z2 = cz.residual(1, ph1 * T1, ph2 * T2, gz, cz.integrals_from_vertex(gz, No), 3)
    assert np.abs(Z2 - ph2 * R2).max() <= 1e-12 * np.abs(R2).max()
    assert np.abs(Z1 - ph1 * R1).max() <= 1e-12 * np.abs(R2).max()
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=12, slice_size=3,
              amplitudes_convergence=1e-10, energy_convergence=1e-12)
    a = ccsd.solve(g, ei, ea, **kw)
    b = cz.solve(gz, ei, ea, **kw)
    assert np.abs(np.array(b["energies"]) - a["energies"]).max() <= 1e-12
    assert np.abs(np.imag(b["energies"])).max() <= 1e-13


@pytest.mark.parametrize("gauge", [False, True])
def test_complex_path_ueg_plane_wave_known_answers(gauge):
    """The plane-wave UEG vertex (not symmetric in p,q; complex after a gauge
    transformation) through the complex path: the reference's MP2 and
    ten-iteration CCSD energies (integration-tests/tests/cc4s/ueg/
    rs1.0-7occ-26virt/cc4s.correct.out.yaml:128, 153)."""
    from oracle import ccsd_complex as cz
    pw = ueg.plane_wave(7, 26, 1.0)
    G = pw["coeff"].astype(complex)
    if gauge:
        G, _ = cz.gauge_transform(G, 7, 2)
    ei, ea = pw["eps"][:7], pw["eps"][7:]
    r = cz.solve(G, ei, ea, mixer="DiisMixer", max_residua=4, max_iterations=10,
                 energy_convergence=1e-8, amplitudes_convergence=1e-8)
    assert abs(r["energies"][0] - -0.36143028565969504) <= 1e-13
    assert abs(r["energy"] - -0.39269658954585018) <= 1e-13
    assert abs(r["energy_complex"].imag) <= 1e-13


def test_complex_triples_equals_real_for_real_vertex_and_is_gauge_invariant():
    """CcsdPerturbativeTriplesComplex (CcsdPerturbativeTriplesComplex.cxx:
    172-279): on a real symmetric vertex it is the real (T); on the
    gauge-transformed vertex with that vertex's converged complex amplitudes
    the energy is unchanged and real."""
    from oracle import ccsd_complex as cz, triples_complex as tz
    g, ei, ea = philox.system(5, 3, 8, 20)
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=40, slice_size=3,
              amplitudes_convergence=1e-10, energy_convergence=1e-12)
    a = ccsd.solve(g, ei, ea, **kw)
    et = triples.triples_energy(a["T1"], a["T2"], g, ei, ea)
    ez = tz.triples_energy(a["T1"].astype(complex), a["T2"].astype(complex), g.astype(complex),
                           ei, ea)
    assert abs(ez - et) <= 1e-14 * abs(et)
    gz, _ = cz.gauge_transform(g.astype(complex), 3, 1)
    c = cz.solve(gz, ei, ea, **kw)
    eg = tz.triples_energy(c["T1"], c["T2"], gz, ei, ea)
    assert abs(eg.real - et) <= 1e-12 * abs(et) and abs(eg.imag) <= 1e-13


def test_complex_triples_ueg_plane_wave_known_answer():
    """(T) of the plane-wave UEG vertex on the ten-iteration amplitudes: the
    reference's value (integration-tests/tests/cc4s/ueg/rs1.0-7occ-26virt/
    cc4s.correct.out.yaml:169)."""
    from oracle import ccsd_complex as cz, triples_complex as tz
    pw = ueg.plane_wave(7, 26, 1.0)
    G = pw["coeff"].astype(complex)
    ei, ea = pw["eps"][:7], pw["eps"][7:]
    r = cz.solve(G, ei, ea, mixer="DiisMixer", max_residua=4, max_iterations=10,
                 energy_convergence=1e-8, amplitudes_convergence=1e-8)
    et = tz.triples_energy(r["T1"], r["T2"], G, ei, ea)
    assert abs(et.real - -0.0063019625641725016) <= 1e-15


@pytest.mark.parametrize("fs", [1, 7, 20])
def test_factor_ladder_equals_vertex_ladder(fs):
    """The factor-based ladder (ClusterSinglesDoublesAlgorithm.cxx:567-672)
    equals the vertex ladder when Gamma[G,p,q] = sum_R Lambda_GR Pi_pR Pi_qR,
    for every factor slice size; with complex factors it equals the real
    path's Re sum_G conj(dGamma_G,ac) dGamma_G,bd X_cdij written densely."""
    rng = np.random.default_rng(3)
    No, Nv, NG, NR = 3, 8, 12, 20
    Lam = rng.normal(size=(NG, NR)) / np.sqrt(NR)
    Pi = rng.normal(size=(No + Nv, NR)) / 2
    g = np.einsum("GR,pR,qR->Gpq", Lam, Pi, Pi)
    ints = ccsd.integrals_from_vertex(g, No)
    T1, T2 = random_amplitudes(Nv, No, seed=4)
    R1, R2 = ccsd.residual(1, T1, T2, g, ints, 3)
    F1, F2 = ccsd.residual(1, T1, T2, g, ints, 3, factors=(Lam.astype(complex),
                                                           Pi.astype(complex), fs))
    assert np.abs(F2 - R2).max() <= 1e-13 * np.abs(R2).max() and np.abs(F1 - R1).max() == 0.0
    Lz = Lam + 1j * rng.normal(size=Lam.shape) / np.sqrt(NR)
    Pz = Pi + 1j * rng.normal(size=Pi.shape) / 2
    gz = np.einsum("GR,pR,qR->Gpq", Lz, Pz, Pz)
    X = T2 + np.einsum("ai,bj->abij", T1, T1)
    dG = gz[:, No:, No:] - np.einsum("Gbk,ak->Gab", gz[:, No:, :No], T1)
    ref = np.real(np.einsum("Gac,Gbd,cdij->abij", dG.conj(), dG, X))
    out = np.zeros_like(X)
    ccsd.factor_ladder(T1, X, Lz, Pz, fs, out)
    assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()


def test_only_ppl_is_the_ladder_of_the_vertex_path_without_singles():
    """onlyPPL (CcsdEnergyFromCoulombIntegralsReference.cxx:56-73) with T1 = 0
    is the particle-particle ladder the vertex path adds (its residual with
    ppl = 1 minus with ppl = 0), independent of the slice size."""
    No, Nv = 3, 8
    g, ei, ea = philox.system(9, No, Nv, 15)
    ints = ccsd.integrals_from_vertex(g, No)
    _, T2 = random_amplitudes(Nv, No, seed=6)
    T1 = np.zeros((Nv, No))
    _, P2 = formulations.residual_only_ppl(T1, T2, g, No)
    for s in (2, 8):
        _, A = ccsd.residual(1, T1, T2, g, ints, s)
        _, B = ccsd.residual(1, T1, T2, g, ints, s, ppl=False)
        assert np.abs((A - B) - P2).max() <= 1e-13 * np.abs(P2).max()


def test_complex_factor_ladder_equals_complex_vertex_ladder():
    """The factor-based ladder for complex amplitudes (ClusterSinglesDoublesAlgorithm.cxx:
    745-880) with FactorOrbitals Pi and OutgoingFactorOrbitals Po equals the
    complex vertex ladder when Gamma[G,q,r] = sum_R Lambda_GR Po_qR Pi_rR: on
    the ladder alone for any factors, and on the whole residual (whose vertex
    ladder evaluates slice pairs a >= b and adds their transposes, i.e.
    assumes a vertex Hermitian in (q,r)) for Po = conj(Pi), Lambda real, for
    several factor slice sizes."""
    from oracle import ccsd_complex as cz
    rng = np.random.default_rng(3)
    No, Nv, NG, NR = 3, 7, 10, 16

    def cn(*s):
        return rng.normal(size=s) + 1j * rng.normal(size=s)

    T1, T2 = 0.1 * cn(Nv, No), 0.1 * cn(Nv, Nv, No, No)
    T2 = T2 + T2.transpose(1, 0, 3, 2)
    X = T2 + np.einsum("ai,bj->abij", T1, T1)
    Lam, Pi, Po = cn(NG, NR) / np.sqrt(NR * NG), cn(No + Nv, NR) / 2, cn(No + Nv, NR) / 2
    g = np.einsum("GR,qR,rR->Gqr", Lam, Po, Pi)
    out = np.zeros_like(X)
    cz.factor_ladder(T1, X, Lam, Pi, Po, 5, out)
    right = g[:, No:, No:] - np.einsum("Gkb,ak->Gab", g[:, :No, No:], T1)
    left = (np.einsum("Gba->Gab", g[:, No:, No:].conj())
            - np.einsum("Gkb,ak->Gab", np.einsum("Gai->Gia", g[:, No:, :No].conj()), T1))
    full = np.einsum("Gxc,Gyd,cdij->xyij", left, right, X)
    assert np.abs(out - full).max() <= 1e-13 * np.abs(full).max()
    Lam = Lam.real
    Po = Pi.conj()
    g = np.einsum("GR,qR,rR->Gqr", Lam, Po, Pi)
    ints = cz.integrals_from_vertex(g, No)
    R1, R2 = cz.residual(1, T1, T2, g, ints, 3)
    for fs in (1, 5, 16):
        F1, F2 = cz.residual(1, T1, T2, g, ints, 3, factors=(Lam, Pi, Po, fs))
        assert np.abs(F2 - R2).max() <= 1e-13 * np.abs(R2).max() and np.abs(F1 - R1).max() == 0.0


# --- direct-ring CCD (DrccdEnergyFromCoulombIntegrals) ------------------------

def _drccd_system(seed=31, No=3, Nv=8, NG=24):
    from oracle import drccd
    g, ei, ea = philox.system(seed, No, Nv, NG)
    V = ccsd.integrals_from_vertex(g, No)["PPHH"]
    return drccd, V, ei, ea


def _ring_matrices(V, ei, ea):
    """K[(a,i),(b,j)] = V_abij and the orbital-energy differences (a,i)."""
    Nv, No = V.shape[0], V.shape[2]
    K = np.transpose(V, (0, 2, 1, 3)).reshape(Nv * No, Nv * No)
    delta = (ea[:, None] - ei[None, :]).reshape(-1)
    return K, delta


@pytest.mark.parametrize("spins", [2.0, 1.0])
def test_drccd_direct_energy_is_the_rpa_correlation_energy(spins):
    """Converged ring CCD gives the direct-RPA correlation energy
    1/2 Tr(Omega - A), Omega^2 = eig((A-B)^1/2 (A+B) (A-B)^1/2), with
    A = Delta + s K, B = s K (s = 2 restricted, 1 unrestricted)."""
    drccd, V, ei, ea = _drccd_system()
    r = drccd.solve(V, ei, ea, spins=spins, max_residua=6, max_iterations=200,
                    amplitudes_convergence=1e-12, energy_convergence=1e-14)
    assert r["converged"]
    K, delta = _ring_matrices(V, ei, ea)
    A = np.diag(delta) + spins * K
    sq = np.sqrt(delta)
    M = sq[:, None] * (np.diag(delta) + 2.0 * spins * K) * sq[None, :]
    omega = np.sqrt(np.linalg.eigvalsh(M))
    e_rpa = 0.5 * (omega.sum() - np.trace(A))
    assert abs(r["direct"] - e_rpa) <= 1e-10 * abs(e_rpa), (r["direct"], e_rpa)
    # a planted error in the quadratic term moves the energy off the closed form
    Vb = V.copy()
    r2 = drccd.solve(Vb, ei, ea, spins=spins, max_residua=6, max_iterations=200,
                     amplitudes_convergence=1e-12, energy_convergence=1e-14,
                     adjacent_pairs_exchange=True)
    assert abs(r2["direct"] - e_rpa) > 1e-6 * abs(e_rpa)


def test_linearized_drccd_solves_the_sylvester_equation():
    """linearized = 1: K + A T + T A = 0 with A = Delta + 2K (matrix form over
    compound indices (a,i)), solved independently by scipy."""
    import scipy.linalg
    drccd, V, ei, ea = _drccd_system(seed=32)
    r = drccd.solve(V, ei, ea, linearized=True, max_residua=6, max_iterations=200,
                    amplitudes_convergence=1e-12, energy_convergence=1e-14)
    assert r["converged"]
    K, delta = _ring_matrices(V, ei, ea)
    A = np.diag(delta) + 2.0 * K
    T = scipy.linalg.solve_sylvester(A, A, -K)
    Nv, No = V.shape[0], V.shape[2]
    Tm = np.transpose(r["T2"], (0, 2, 1, 3)).reshape(Nv * No, Nv * No)
    assert np.abs(Tm - T).max() <= 1e-10 * np.abs(T).max()


def test_drccd_adjacent_pairs_exchange_term_by_loops_and_mp2_start():
    drccd, V, ei, ea = _drccd_system(seed=33, No=2, Nv=3, NG=7)
    _, Vaijb, Vijab = drccd.integrals(V)
    T1, T2 = random_amplitudes(3, 2, seed=4)
    T2 = 0.5 * (T2 + np.transpose(T2, (1, 0, 3, 2)))
    for spins in (2.0, 1.0):
        _, Ra = drccd.residual(1, T2, V, Vaijb, Vijab, spins=spins, adjacent_pairs_exchange=True)
        _, Rn = drccd.residual(1, T2, V, Vaijb, Vijab, spins=spins)
        ref = np.zeros_like(Rn)
        Nv, No = 3, 2
        for a in range(Nv):
            for b in range(Nv):
                for i in range(No):
                    for j in range(No):
                        s = 0.0
                        for k in range(No):
                            for l in range(No):
                                for c in range(Nv):
                                    for d in range(Nv):
                                        s += V[c, d, l, k] * T2[a, c, i, k] * T2[d, b, l, j]
                        ref[a, b, i, j] = -spins * s
        assert np.abs((Ra - Rn) - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())
    # iteration 0: MP2 start, energy = closed-form MP2
    g, ei2, ea2 = philox.system(34, 3, 8, 24)
    V2 = ccsd.integrals_from_vertex(g, 3)["PPHH"]
    r = drccd.solve(V2, ei2, ea2, max_iterations=1)
    assert abs(r["energy"] - formulations.mp2_closed_form(g, ei2, ea2)) <= 1e-12


# --- spin-orbital CCSD (UccsdAmplitudesFromCoulombIntegrals) ------------------

def test_uccsd_closed_shell_equals_restricted_ccsd():
    """A closed-shell system written in spin orbitals: iteration 1 is MP2 and
    the converged energy is the restricted CCSD energy of the pinned real path."""
    from oracle import uccsd
    No, Nv, NG = 2, 5, 14
    g, ei, ea = philox.system(41, No, Nv, NG)
    ints = uccsd.spin_orbital_integrals(g, No, Nv)
    kw = dict(max_residua=6, max_iterations=200, amplitudes_convergence=1e-11,
              energy_convergence=1e-13)
    ru = uccsd.solve(ints, np.repeat(ei, 2), np.repeat(ea, 2), **kw)
    rr = ccsd.solve(g, ei, ea, **kw)
    assert ru["converged"] and rr["converged"]
    assert abs(ru["energies"][0] - formulations.mp2_closed_form(g, ei, ea)) <= 1e-13
    assert abs(ru["energy"] - rr["energy"]) <= 1e-10 * abs(rr["energy"])


def _two_electron_fci(asym, eps, no=2):
    """Full-CI correlation energy of a two-electron Hamiltonian whose Fock
    matrix with respect to the first `no` = 2 spin orbitals is diag(eps)."""
    n = asym.shape[0]
    f = np.diag(eps)
    h = f - np.einsum("pkqk->pq", asym[:, :no, :, :no])
    dets = [(p, q) for p in range(n) for q in range(p + 1, n)]
    H = np.zeros((len(dets), len(dets)))
    for I, (p, q) in enumerate(dets):
        for J, (r, s) in enumerate(dets):
            H[I, J] = (h[p, r] * (q == s) - h[p, s] * (q == r) - h[q, r] * (p == s)
                       + h[q, s] * (p == r) + asym[p, q, r, s])
    e_ref = h[0, 0] + h[1, 1] + asym[0, 1, 0, 1]
    return np.linalg.eigvalsh(H)[0] - e_ref


def test_uccsd_is_exact_for_two_electrons():
    """Generic (spin-symmetry-broken) real spin-orbital integrals, two
    electrons: the converged energy equals full CI from Slater-Condon rules
    and an exact diagonalisation; a planted error in one term is detected."""
    from oracle import uccsd
    rng = np.random.default_rng(7)
    n, NG = 8, 12
    g = rng.normal(size=(NG, n, n)) * 0.25
    g = 0.5 * (g + np.transpose(g, (0, 2, 1)))
    eri = np.einsum("Gpr,Gqs->pqrs", g, g)
    asym = eri - np.swapaxes(eri, 2, 3)
    eps = np.concatenate([[-1.1, -0.8], np.sort(rng.uniform(0.3, 2.0, n - 2))])
    ints = uccsd.classes(asym, 2)
    r = uccsd.solve(ints, eps[:2], eps[2:], max_residua=6, max_iterations=300,
                    amplitudes_convergence=1e-11, energy_convergence=1e-13)
    assert r["converged"]
    e_fci = _two_electron_fci(asym, eps)
    assert abs(r["energy"] - e_fci) <= 1e-10 * abs(e_fci), (r["energy"], e_fci)
    # plant: drop the T1 <na||if> term's sign (a plausible index mistake)
    bad = dict(ints)
    bad["PHPH"] = -ints["PHPH"]
    rb = uccsd.solve(bad, eps[:2], eps[2:], max_residua=6, max_iterations=300,
                     amplitudes_convergence=1e-11, energy_convergence=1e-13)
    assert abs(rb["energy"] - e_fci) > 1e-6


def test_complex_drccd_is_gauge_covariant():
    """F = complex direct-ring CCD on a gauge-transformed vertex (random
    orbital phases and a unitary on the auxiliary index): every iteration's
    energy equals the real path's, with zero imaginary part."""
    from oracle import ccsd_complex as cz, drccd
    No, Nv, NG = 3, 8, 24
    g, ei, ea = philox.system(44, No, Nv, NG)
    V = ccsd.integrals_from_vertex(g, No)["PPHH"]
    gz, _ = cz.gauge_transform(g.astype(complex), No, 45)
    iz = cz.integrals_from_vertex(gz, No)
    kw = dict(max_residua=6, max_iterations=40, amplitudes_convergence=1e-10,
              energy_convergence=1e-12, adjacent_pairs_exchange=True)
    rr = drccd.solve(V, ei, ea, **kw)
    rz = drccd.solve(None, ei, ea, complex_integrals=(iz["PPHH"], iz["PHHP"], iz["HHPP"]), **kw)
    ez = np.array(rz["energies"])
    assert rr["iterations"] == rz["iterations"]
    assert np.abs(ez.real - np.array(rr["energies"])).max() <= 1e-12
    assert np.abs(ez.imag).max() <= 1e-12


def test_linear_mixer_fixed_point_and_unit_ratio():
    """LinearMixer (LinearMixer.cxx:30-55): with ratio 1 it is the unmixed
    update, i.e. DIIS with one vector (weight 1) iteration for iteration; with
    ratio 0.6 on a weakly correlated system it converges to the same energy
    as DIIS (mixing changes the path, not the fixed point)."""
    g, ei, ea = philox.system(51, 3, 10, 30)
    g = 0.5 * g   # weak coupling: the plain update converges
    kw = dict(max_iterations=12, amplitudes_convergence=0.0, energy_convergence=0.0)
    r1 = ccsd.solve(g, ei, ea, mixer="LinearMixer", mixing_ratio=1.0, **kw)
    d1 = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_residua=1, **kw)
    assert np.abs(np.array(r1["energies"]) - np.array(d1["energies"])).max() <= 1e-14
    conv = dict(amplitudes_convergence=1e-11, energy_convergence=1e-13)
    rl = ccsd.solve(g, ei, ea, mixer="LinearMixer", mixing_ratio=0.6, max_iterations=300, **conv)
    rd = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_residua=6, max_iterations=300, **conv)
    assert rl["converged"] and rd["converged"]
    assert abs(rl["energy"] - rd["energy"]) <= 1e-10 * abs(rd["energy"])
    assert rl["iterations"] != rd["iterations"]


# --- IP-EOM-CCSD right-apply (UCcsdIPEquationOfMotionDavidson) ---------------

def test_ip_eom_is_exact_for_two_electrons():
    """Two electrons, generic spin-orbital integrals: every IP-EOM-CCSD
    eigenvalue equals E(N-1) - E(N) from full CI."""
    from oracle import eom, uccsd
    rng = np.random.default_rng(7)
    n, NG = 8, 12
    g = rng.normal(size=(NG, n, n)) * 0.25
    g = 0.5 * (g + np.transpose(g, (0, 2, 1)))
    eri = np.einsum("Gpr,Gqs->pqrs", g, g)
    asym = eri - np.swapaxes(eri, 2, 3)
    eps = np.concatenate([[-1.1, -0.8], np.sort(rng.uniform(0.3, 2.0, n - 2))])
    ints = uccsd.classes(asym, 2)
    r = uccsd.solve(ints, eps[:2], eps[2:], max_residua=6, max_iterations=300,
                    amplitudes_convergence=1e-12, energy_convergence=1e-14)
    assert r["converged"]
    ev = eom.ip_eigenvalues(ints, eps[:2], eps[2:], r["T1"], r["T2"])
    h = np.diag(eps) - np.einsum("pkqk->pq", asym[:, :2, :, :2])
    dets = [(p, q) for p in range(n) for q in range(p + 1, n)]
    H = np.array([[h[p, s] * (q == t) - h[p, t] * (q == s) - h[q, s] * (p == t)
                   + h[q, t] * (p == s) + asym[p, q, s, t] for (s, t) in dets] for (p, q) in dets])
    exact = np.sort(np.linalg.eigvalsh(h) - np.linalg.eigvalsh(H)[0])
    assert np.abs(ev - exact).max() <= 1e-11


def test_ip_eom_closed_shell_spin_degeneracy():
    """A closed-shell system in spin orbitals: the IP spectrum comes in
    degenerate spin pairs (doublets), and the lowest root lies near minus the
    highest occupied orbital energy (Koopmans) for a weak interaction."""
    from oracle import eom, uccsd
    g, ei, ea = philox.system(61, 2, 4, 12)
    g = 0.3 * g
    ints = uccsd.spin_orbital_integrals(g, 2, 4)
    epi, epa = np.repeat(ei, 2), np.repeat(ea, 2)
    r = uccsd.solve(ints, epi, epa, max_residua=6, max_iterations=200,
                    amplitudes_convergence=1e-12, energy_convergence=1e-14)
    ev = eom.ip_eigenvalues(ints, epi, epa, r["T1"], r["T2"])
    low = ev[:4]
    assert abs(low[0] - low[1]) <= 1e-9 and abs(low[2] - low[3]) <= 1e-9
    assert abs(low[0] + ei.max()) <= 0.1 * abs(ei.max())


# --- EA-EOM-CCSD right-apply (UCcsdEAEquationOfMotionDavidson) ---------------

def _fci_energies(h, asym, k):
    """Full CI in the k-electron determinant space of H = sum h_pq p+ q +
    1/4 sum <pq||rs> p+ q+ s r, by applying the operators to occupation bit
    strings (the sign of each operator: (-1) to the number of occupied
    orbitals below it).  Small n only."""
    n = h.shape[0]
    dets = [sum(1 << p for p in c) for c in itertools.combinations(range(n), k)]
    index = {d: i for i, d in enumerate(dets)}

    def op(d, p, create):
        if ((d >> p) & 1) == create:
            return None, 0
        return d ^ (1 << p), (-1) ** bin(d & ((1 << p) - 1)).count("1")

    H = np.zeros((len(dets), len(dets)))
    for j, d in enumerate(dets):
        for p in range(n):
            for q in range(n):
                d1, s1 = op(d, q, 0)
                d2, s2 = op(d1, p, 1) if d1 is not None else (None, 0)
                if d2 is not None:
                    H[index[d2], j] += s1 * s2 * h[p, q]
        for r in range(n):
            for s in range(r + 1, n):
                d1, s1 = op(d, r, 0)
                d2, s2 = op(d1, s, 0) if d1 is not None else (None, 0)
                if d2 is None:
                    continue
                for p in range(n):
                    for q in range(p + 1, n):
                        d3, s3 = op(d2, q, 1)
                        d4, s4 = op(d3, p, 1) if d3 is not None else (None, 0)
                        if d4 is not None:
                            H[index[d4], j] += s1 * s2 * s3 * s4 * asym[p, q, r, s]
    return np.linalg.eigvalsh(H)


def _generic_spin_orbital_system(no, n, seed):
    rng = np.random.default_rng(seed)
    g = rng.normal(size=(12, n, n)) * 0.25
    g = 0.5 * (g + np.transpose(g, (0, 2, 1)))
    eri = np.einsum("Gpr,Gqs->pqrs", g, g)
    asym = eri - np.swapaxes(eri, 2, 3)
    eps = np.concatenate([np.linspace(-1.3, -0.8, no), np.sort(rng.uniform(0.3, 2.0, n - no))])
    h = np.diag(eps) - np.einsum("pkqk->pq", asym[:, :no, :, :no])  # Fock matrix = diag(eps)
    return asym, eps, h


def test_fci_helper_closed_forms():
    """The determinant full CI used as the EOM pin: one electron gives the
    eigenvalues of h; two electrons give the Slater-Condon matrix of the IP
    test; the Hartree-Fock determinant's energy is E_ref = sum h_ii +
    1/2 sum <ij||ij> when it is an eigenvector (no interaction)."""
    asym, eps, h = _generic_spin_orbital_system(2, 6, 3)
    assert np.abs(_fci_energies(h, asym, 1) - np.linalg.eigvalsh(h)).max() <= 1e-12
    dets = [(p, q) for p in range(6) for q in range(p + 1, 6)]
    H = np.array([[h[p, s] * (q == t) - h[p, t] * (q == s) - h[q, s] * (p == t)
                   + h[q, t] * (p == s) + asym[p, q, s, t] for (s, t) in dets] for (p, q) in dets])
    assert np.abs(_fci_energies(h, asym, 2) - np.linalg.eigvalsh(H)).max() <= 1e-12
    e = np.array([-1.0, -0.5, -0.2, 0.4, 0.9])
    assert abs(_fci_energies(np.diag(e), np.zeros((5,) * 4), 3)[0] - e[:3].sum()) <= 1e-15


@pytest.mark.parametrize("no", [2, 3, 4])
def test_ea_eom_is_exact_with_two_virtual_spin_orbitals(no):
    """With two virtual spin orbitals CCSD is exact for any electron count and
    the EA space (1p + 2p1h) spans every (N+1)-electron determinant, so every
    EA-EOM-CCSD eigenvalue equals E(N+1) - E(N) from full CI.  Scaling any
    one block of Hbar by 1 % moves the eigenvalues by 2e-5 to 2e-2."""
    from oracle import eom, uccsd
    asym, eps, h = _generic_spin_orbital_system(no, no + 2, 5)
    ints = uccsd.classes(asym, no)
    r = uccsd.solve(ints, eps[:no], eps[no:], max_residua=6, max_iterations=300,
                    amplitudes_convergence=1e-12, energy_convergence=1e-14)
    assert r["converged"]
    E0 = _fci_energies(h, asym, no)[0]
    e_ref = np.trace(h[:no, :no]) + 0.5 * np.einsum("ijij->", asym[:no, :no, :no, :no])
    assert abs(e_ref + r["energy"] - E0) <= 1e-11  # CCSD exact
    ev = eom.ea_eigenvalues(ints, eps[:no], eps[no:], r["T1"], r["T2"])
    exact = np.sort(_fci_energies(h, asym, no + 1) - E0)
    assert len(ev) == len(exact) == no + 2
    assert np.abs(ev - exact).max() <= 1e-11


def test_ea_eom_closed_shell_spin_degeneracy():
    """Closed shell in spin orbitals: the EA roots come in spin doublets and
    the lowest lies near the lowest virtual orbital energy for a weak
    interaction (Koopmans)."""
    from oracle import eom, uccsd
    g, ei, ea = philox.system(61, 2, 3, 12)
    g = 0.3 * g
    ints = uccsd.spin_orbital_integrals(g, 2, 3)
    epi, epa = np.repeat(ei, 2), np.repeat(ea, 2)
    r = uccsd.solve(ints, epi, epa, max_residua=6, max_iterations=200,
                    amplitudes_convergence=1e-12, energy_convergence=1e-14)
    ev = eom.ea_eigenvalues(ints, epi, epa, r["T1"], r["T2"])
    assert abs(ev[0] - ev[1]) <= 1e-9 and abs(ev[2] - ev[3]) <= 1e-9
    assert abs(ev[0] - ea.min()) <= 0.1 * abs(ea.min())


def test_energy_variants_closed_forms():
    """getEnergy's arguments (ClusterSinglesDoublesAlgorithm.cxx:131-213) on a
    closed shell written in spin orbitals, against the restricted energy of
    the pinned real path: the antisymmetrized energy of the spin-orbital
    amplitudes and <ij||ab> equals it; the unrestricted (spins = 1) direct
    minus exchange energy of the spin-orbital amplitudes and the plain <ij|ab>
    equals twice its doubles part plus its singles part (the spin sums give
    4 and 2 for the doubles' direct and exchange terms, 4 and 2 against 2
    and 1 for the singles').  The HPFockMatrix term is the element sum
    spins * T_ai f_ia.  Planted mistakes (spins^2 in the exchange term, the
    singles term of the antisymmetrized energy dropped) are detected."""
    from oracle import uccsd
    No, Nv, NG = 2, 5, 14
    g, ei, ea = philox.system(41, No, Nv, NG)
    r = ccsd.solve(g, ei, ea, max_residua=6, max_iterations=8, amplitudes_convergence=0.0,
                   energy_convergence=0.0)
    T1, T2 = r["T1"], r["T2"]
    V = r["integrals"]["PPHH"]
    E = ccsd.energy(T1, T2, V)[0]
    assert abs(ccsd.energy_of(T1, T2, V) - E) <= 1e-13 * abs(E)
    # spin orbital P = 2 p + s (alpha 0, beta 1) inside each block
    sp_o, sp_v = np.tile([0, 1], No), np.tile([0, 1], Nv)
    o2, v2 = np.repeat(np.arange(No), 2), np.repeat(np.arange(Nv), 2)
    d = lambda x, y: (x[:, None] == y[None, :]).astype(float)  # noqa: E731
    S1 = T1[np.ix_(v2, o2)] * d(sp_v, sp_o)
    Tdir = T2[np.ix_(v2, v2, o2, o2)] * (d(sp_v, sp_o)[:, None, :, None]
                                         * d(sp_v, sp_o)[None, :, None, :])
    S2 = Tdir - Tdir.transpose(1, 0, 2, 3)
    Vplain = V[np.ix_(v2, v2, o2, o2)] * (d(sp_v, sp_o)[:, None, :, None]
                                          * d(sp_v, sp_o)[None, :, None, :])
    Vanti = uccsd.spin_orbital_integrals(g, No, Nv)["PPHH"]
    assert np.abs(Vanti - (Vplain - Vplain.transpose(1, 0, 2, 3))).max() <= 1e-14
    Ea = ccsd.energy_of(S1, S2, Vanti, antisymmetrized=True)
    assert abs(Ea - E) <= 1e-12 * abs(E)
    Eu = ccsd.energy_of(S1, S2, Vplain, spins=1.0)
    E2 = ccsd.energy_of(0.0 * T1, T2, V)
    assert abs(Eu - (2.0 * E2 + (E - E2))) <= 1e-12 * abs(E)
    assert abs(E - E2) > 1e-6 * abs(E)  # the singles part is not negligible here
    # planted: the antisymmetrized energy without its singles term
    Vij = ccsd.ein("abij->ijab", Vanti)
    assert abs(0.25 * ccsd.ein("abkl,klab->", S2, Vij) - E) > 1e-6 * abs(E)
    # planted: spins^2 in the exchange term of the unrestricted energy (equals
    # the correct one at spins = 1, so compare at spins = 2 with the restricted E)
    Vr = ccsd.ein("abij->ijab", V)
    tau = T2 + ccsd.ein("ai,bj->abij", T1, T1)
    wrong = 2.0 * ccsd.ein("abij,ijab->", tau, Vr) - 2.0 * ccsd.ein("abij,ijba->", tau, Vr)
    assert abs(wrong - E) > 1e-6 * abs(E)
    rng = np.random.default_rng(3)
    f = rng.standard_normal((No, Nv))
    nc = sum(T1[a, i] * f[i, a] for a in range(Nv) for i in range(No))
    for spins in (1.0, 2.0):
        base = ccsd.energy_of(T1, T2, V, spins=spins)
        assert abs(ccsd.energy_of(T1, T2, V, spins=spins, fia=f) - base - spins * nc) <= 1e-13
    P = ccsd.pair_energy(S1, S2, Vanti, antisymmetrized=True)
    assert abs(P.sum() - Ea) <= 1e-13 * abs(Ea)
