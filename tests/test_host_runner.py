"""The C++ step runner (host mirror of the plugin surface): input parsing on
the CPU; end-to-end CCSD + (T) at No=10, Nv=50, NG=300 on the GPU against the oracle."""

import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, announce

RUNNER = os.environ.get("SISI_RUNNER") or os.path.join(ROOT, "sisi4s_amd", "sisi4s-mi355x")
INPUT = os.path.join(ROOT, "tests", "inputs", "no10_nv50_ccsd_t.yaml")
UEG_INPUT = os.path.join(ROOT, "tests", "inputs", "ueg_rs1_no7_nv26.yaml")


def run(args, **kw):
    return subprocess.run([RUNNER] + args, capture_output=True, text=True, timeout=600, **kw)


def test_runner_lists_plugins():
    r = run(["--list"])
    assert r.returncode == 0
    names = r.stdout.split()
    for n in ("CoulombIntegralsFromVertex", "CcsdEnergyFromCoulombIntegrals",
              "CcsdPerturbativeTriples", "GenerateSyntheticSystem", "UegVertexGenerator"):
        assert n in names


def test_runner_parses_reference_style_input():
    r = run(["-i", INPUT, "--dry"])
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert [l.split()[0] for l in lines] == ["Nop", "GenerateSyntheticSystem",
                                             "CoulombIntegralsFromVertex",
                                             "CcsdEnergyFromCoulombIntegrals",
                                             "CcsdPerturbativeTriples",
                                             "CcsdEnergyFromCoulombIntegrals"]
    ccsd = lines[3]
    # aliases resolved, symbols kept
    assert 'mixer: "DiisMixer"' in ccsd and 'maxResidua: "6"' in ccsd
    assert "CoulombVertex: $CoulombVertex" in ccsd and "CcsdEnergy: $CcsdEnergy" in ccsd
    assert "(fallible)" in lines[5]


def test_runner_parses_ueg_input():
    r = run(["-i", UEG_INPUT, "--dry"])
    assert r.returncode == 0, r.stderr
    names = [l.split()[0] for l in r.stdout.strip().splitlines()]
    assert names == ["Nop", "UegVertexGenerator", "CoulombIntegralsFromVertex",
                     "CcsdEnergyFromCoulombIntegrals", "CcsdPerturbativeTriples"]


def test_runner_rejects_unknown_algorithm(tmp_path):
    p = tmp_path / "bad.yaml"
    p.write_text("- name: NoSuchAlgorithm\n  in:\n    x: 1\n")
    r = run(["-i", str(p), "--dry"])
    assert r.returncode == 1 and "Algorithm not implemented" in r.stderr


def test_runner_rejects_malformed_yaml(tmp_path):
    p = tmp_path / "bad.yaml"
    p.write_text("- name: Nop\n  in:\n    a: *undefined\n")
    r = run(["-i", str(p), "--dry"])
    assert r.returncode == 1 and "unknown alias" in r.stderr


@pytest.mark.gpu
def test_runner_no10_nv50_ccsd_and_triples(tmp_path):
    from oracle import ccsd, philox, triples
    out = tmp_path / "out.yaml"
    announce("runner No=10 Nv=50 ccsd + (T)")
    r = run(["-i", INPUT, "-o", str(out)])
    assert r.returncode == 0, r.stdout + r.stderr
    vals = {}
    for line in r.stdout.splitlines():
        if "=" in line:
            k, v = line.split("=")
            vals[k.strip()] = float(v)
    assert "failed" in r.stderr and "NoSuchMixer" in r.stderr  # fallible step
    g, ei, ea = philox.system(20261202, 10, 50, 300)
    ref = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_residua=6, max_iterations=30,
                     amplitudes_convergence=1e-12, energy_convergence=1e-14)
    assert abs(vals["CcsdEnergy"] - ref["energy"]) <= 1e-8
    et = triples.triples_energy(ref["T1"], ref["T2"], g, ei, ea)
    assert abs(vals["CcsdPerturbativeTriplesEnergy"] - (ref["energy"] + et)) <= 1e-8
    text = out.read_text()
    assert text.count("iteration:") == 30 and "realtime:" in text


@pytest.mark.gpu
def test_runner_ueg_known_answers(tmp_path):
    """The reference's UEG test end to end: MP2, CCSD and CCSD(T) energies of
    integration-tests/tests/cc4s/ueg/rs1.0-7occ-26virt/cc4s.correct.out.yaml
    (:128, :153, :169)."""
    out = tmp_path / "out.yaml"
    announce("runner ueg rs1 No7 Nv26 known answers")
    r = run(["-i", UEG_INPUT, "-o", str(out)])
    assert r.returncode == 0, r.stdout + r.stderr
    vals = {}
    for line in r.stdout.splitlines():
        if "=" in line:
            k, v = line.split("=")
            vals[k.strip()] = float(v)
    assert abs(vals["CcsdEnergy"] - -0.39269658954585018) <= 1e-12
    assert abs(vals["CcsdPerturbativeTriplesEnergy"]
               - (-0.39269658954585018 - 0.0063019625641725016)) <= 1e-12
    m = re.search(r"iteration: 1, energy: \{value: ([-+0-9.eE]+)\}", out.read_text())
    assert m and abs(float(m.group(1)) - -0.36143028565969504) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("args,ppl,dist", [("ppl: 0", False, False),
                                           ("distinguishable: 1", True, True)])
def test_runner_residual_variants(tmp_path, args, ppl, dist):
    """The ppl and distinguishable arguments of CcsdEnergyFromCoulombIntegrals
    (CcsdEnergyFromCoulombIntegrals.cxx:123-124, 324-338) through the runner,
    against the oracle's driver."""
    from oracle import ccsd, philox
    p = tmp_path / "in.yaml"
    p.write_text("""
- name: GenerateSyntheticSystem
  in: {seed: 7, No: 5, Nv: 20, NG: 80}
  out: {CoulombVertex: $G, HoleEigenEnergies: $I, ParticleEigenEnergies: $A}
- name: CoulombIntegralsFromVertex
  in: {CoulombVertex: $G, HoleEigenEnergies: $I, ParticleEigenEnergies: $A}
  out: {PPHHCoulombIntegrals: $V}
- name: CcsdEnergyFromCoulombIntegrals
  in: {mixer: DiisMixer, maxResidua: 6, maxIterations: 20, integralsSliceSize: 6, %s,
       HoleEigenEnergies: $I, ParticleEigenEnergies: $A, PPHHCoulombIntegrals: $V,
       CoulombVertex: $G}
  out: {CcsdEnergy: $E}
""" % args)
    announce("runner variant %s" % args)
    r = run(["-i", str(p), "-o", str(tmp_path / "out.yaml")])
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"CcsdEnergy = (\S+)", r.stdout)
    g, ei, ea = philox.system(7, 5, 20, 80)
    ref = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_residua=6, max_iterations=20,
                     slice_size=6, ppl=ppl, distinguishable=dist)
    assert m and abs(float(m.group(1)) - ref["energy"]) <= 1e-10
    assert ("distinguishable: 1" in (tmp_path / "out.yaml").read_text()) == dist


@pytest.mark.gpu
def test_runner_full_integral_formulation(tmp_path):
    """CcsdEnergyFromCoulombIntegralsReference fed by CoulombIntegralsFromVertex
    (PPPP/PPPH included) gives CcsdEnergyFromCoulombIntegrals' energy."""
    p = tmp_path / "in.yaml"
    p.write_text("""
- name: GenerateSyntheticSystem
  in: {seed: 3, No: 4, Nv: 14, NG: 40}
  out: {CoulombVertex: $G, HoleEigenEnergies: $I, ParticleEigenEnergies: $A}
- name: CoulombIntegralsFromVertex
  in: {CoulombVertex: $G, HoleEigenEnergies: $I, ParticleEigenEnergies: $A}
  out: {PPHHCoulombIntegrals: $PPHH, PHPHCoulombIntegrals: $PHPH, HHHHCoulombIntegrals: $HHHH,
        HHHPCoulombIntegrals: $HHHP, PPPPCoulombIntegrals: $PPPP, PPPHCoulombIntegrals: $PPPH}
- name: CcsdEnergyFromCoulombIntegralsReference
  in: {mixer: DiisMixer, maxResidua: 6, maxIterations: 20, HoleEigenEnergies: $I,
       ParticleEigenEnergies: $A, PPHHCoulombIntegrals: $PPHH, PHPHCoulombIntegrals: $PHPH,
       HHHHCoulombIntegrals: $HHHH, HHHPCoulombIntegrals: $HHHP, PPPPCoulombIntegrals: $PPPP,
       PPPHCoulombIntegrals: $PPPH}
  out: {CcsdEnergy: $Eref}
- name: CcsdEnergyFromCoulombIntegrals
  in: {mixer: DiisMixer, maxResidua: 6, maxIterations: 20, HoleEigenEnergies: $I,
       ParticleEigenEnergies: $A, PPHHCoulombIntegrals: $PPHH, CoulombVertex: $G}
  out: {CcsdEnergy: $E}
""")
    announce("runner full-integral formulation")
    r = run(["-i", str(p), "-o", str(tmp_path / "out.yaml")])
    assert r.returncode == 0, r.stdout + r.stderr
    e = [float(m) for m in re.findall(r"CcsdEnergy = (\S+)", r.stdout)]
    assert len(e) == 2 and abs(e[0] - e[1]) <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("unrestricted", [0, 1])
def test_runner_drccd(tmp_path, unrestricted):
    """DrccdEnergyFromCoulombIntegrals fed by CoulombIntegralsFromVertex's
    PPHH, PHHP and HHPP: the oracle's direct-ring CCD energy."""
    from oracle import ccsd, drccd, philox
    p = tmp_path / "in.yaml"
    p.write_text("""
- name: GenerateSyntheticSystem
  in: {seed: 5, No: 4, Nv: 12, NG: 36}
  out: {CoulombVertex: $G, HoleEigenEnergies: $I, ParticleEigenEnergies: $A}
- name: CoulombIntegralsFromVertex
  in: {CoulombVertex: $G, HoleEigenEnergies: $I, ParticleEigenEnergies: $A}
  out: {PPHHCoulombIntegrals: $PPHH, PHHPCoulombIntegrals: $PHHP, HHPPCoulombIntegrals: $HHPP}
- name: DrccdEnergyFromCoulombIntegrals
  in: {mixer: DiisMixer, maxResidua: 6, maxIterations: 30, unrestricted: %d,
       adjacentPairsExchange: 1, HoleEigenEnergies: $I, ParticleEigenEnergies: $A,
       PPHHCoulombIntegrals: $PPHH, PHHPCoulombIntegrals: $PHHP, HHPPCoulombIntegrals: $HHPP}
  out: {DrccdEnergy: $E, DrccdDoublesAmplitudes: $T2}
""" % unrestricted)
    announce("runner drccd unrestricted=%d" % unrestricted)
    r = run(["-i", str(p), "-o", str(tmp_path / "out.yaml")])
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"DrccdEnergy = (\S+)", r.stdout)
    g, ei, ea = philox.system(5, 4, 12, 36)
    V = ccsd.integrals_from_vertex(g, 4)["PPHH"]
    ref = drccd.solve(V, ei, ea, spins=1.0 if unrestricted else 2.0, adjacent_pairs_exchange=True,
                      mixer="DiisMixer", max_residua=6, max_iterations=30)
    assert m and abs(float(m.group(1)) - ref["energy"]) <= 1e-10, (r.stdout, ref["energy"])
