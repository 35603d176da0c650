"""Input readers of the C++ step runner: the FTOD binary vertex
(CoulombVertexReader), cc4s tensor headers with text or binary elements
(Read) and DefineHolesAndParticles, checked through TensorNorm against
numpy on files written here in the documented layouts; on the GPU, the
reference's uniform-electron-gas pipeline read from such files."""

import os
import re
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT, announce
from oracle import philox

RUNNER = os.environ.get("SISI_RUNNER") or os.path.join(ROOT, "sisi4s_amd", "sisi4s-mi355x")


def run_steps(tmp_path, text):
    p = tmp_path / "in.yaml"
    p.write_text(text)
    r = subprocess.run([RUNNER, "-i", str(p), "-o", str(tmp_path / "out.yaml")],
                       capture_output=True, text=True, timeout=600)
    vals = {}
    for line in r.stdout.splitlines():
        m = re.match(r"\s+(\w+) = (\S+)", line)
        if m:
            vals.setdefault(m.group(1), []).append(float(m.group(2)))
    return r, vals


def write_ftod(path, gamma, eps, No, magic=b"cc4sFTOD", imag=None, extra_chunk=True,
               truncate=0):
    NG, Np, _ = gamma.shape
    out = bytearray(magic[:8].ljust(8, b"\0"))
    out += struct.pack("<6i", No, Np - No, NG, 1, 1, 0)

    def chunk(m, payload):
        return m + struct.pack("<q", 16 + len(payload)) + payload

    if extra_chunk:  # unknown to the reader: skipped by its size
        out += chunk(b"FTIAreal", np.arange(5, dtype="<f8").tobytes())
    out += chunk(b"FTODreal", np.asarray(gamma, "<f8").ravel(order="F").tobytes())
    out += chunk(b"FTODimag", (np.zeros_like(gamma) if imag is None else imag)
                 .astype("<f8").ravel(order="F").tobytes())
    out += chunk(b"FTODepsi", np.asarray(eps, "<f8").tobytes())
    if truncate:
        out = out[:-truncate]
    path.write_bytes(bytes(out))


NORMS = """
- name: TensorNorm
  in: {Data: $CoulombVertex}
  out: {Norm: $NormG}
- name: TensorNorm
  in: {Data: $HoleEigenEnergies}
  out: {Norm: $NormI}
- name: TensorNorm
  in: {Data: $ParticleEigenEnergies}
  out: {Norm: $NormA}
"""


@pytest.mark.parametrize("magic", [b"cc4sFTOD", b"sisi4sFTOD"])
def test_ftod_reader(tmp_path, magic):
    g, ei, ea = philox.system(5, 3, 5, 7)
    write_ftod(tmp_path / "v.bin", g, np.concatenate([ei, ea]), 3, magic=magic)
    r, vals = run_steps(tmp_path, """
- name: CoulombVertexReader
  in: {file: "%s"}
  out:
    CoulombVertex: $CoulombVertex
    HoleEigenEnergies: $HoleEigenEnergies
    ParticleEigenEnergies: $ParticleEigenEnergies
""" % (tmp_path / "v.bin") + NORMS)
    assert r.returncode == 0, r.stderr
    assert vals["Norm"] == pytest.approx([np.linalg.norm(g), np.linalg.norm(ei),
                                          np.linalg.norm(ea)], rel=1e-15)


@pytest.mark.parametrize("case", ["magic", "truncated"])
def test_ftod_reader_rejects(tmp_path, case):
    g, ei, ea = philox.system(5, 3, 5, 7)
    kw = {"magic": b"FTODxxxx"} if case == "magic" else {"truncate": 9}
    write_ftod(tmp_path / "v.bin", g, np.concatenate([ei, ea]), 3, **kw)
    r, _ = run_steps(tmp_path, """
- name: CoulombVertexReader
  in: {file: "%s"}
  out: {CoulombVertex: $G, HoleEigenEnergies: $I, ParticleEigenEnergies: $A}
""" % (tmp_path / "v.bin"))
    assert r.returncode != 0 and "Invalid file format" in r.stderr


def test_ftod_reader_keeps_a_complex_vertex(tmp_path):
    """A vertex with an imaginary part is read as a complex tensor (norm of
    the complex values); the real path refuses it, naming complex: 1."""
    g, ei, ea = philox.system(5, 3, 5, 7)
    write_ftod(tmp_path / "v.bin", g, np.concatenate([ei, ea]), 3, imag=0.3 * g[::-1])
    r, vals = run_steps(tmp_path, """
- name: CoulombVertexReader
  in: {file: "%s"}
  out: {CoulombVertex: $CoulombVertex, HoleEigenEnergies: $I, ParticleEigenEnergies: $A}
- name: TensorNorm
  in: {Data: $CoulombVertex}
  out: {Norm: $N}
""" % (tmp_path / "v.bin"))
    assert r.returncode == 0, r.stderr
    assert vals["Norm"] == pytest.approx([np.linalg.norm(g + 0.3j * g[::-1])], rel=1e-15)


def write_cc4s_tensor(base, x, scalar="Real64", elements="IeeeBinaryFile", types=None):
    dims = "\n".join("  - length: %d\n    type: %s" % (n, (types or ["State"] * x.ndim)[k])
                     for k, n in enumerate(x.shape))
    base.with_suffix(".yaml").write_text(
        "version: 100\ntype: Tensor\nscalarType: %s\ndimensions:\n%s\nelements:\n  type: %s\nunit: 1.0\n"
        % (scalar, dims, elements))
    flat = x.ravel(order="F")
    if elements == "TextFile":
        base.with_suffix(".elements").write_text("".join("%.17g\n" % v for v in flat))
    elif scalar == "Complex64":
        base.with_suffix(".elements").write_bytes(flat.astype("<c16").tobytes())
    else:
        base.with_suffix(".elements").write_bytes(flat.astype("<f8").tobytes())


@pytest.mark.parametrize("scalar,elements", [("Real64", "TextFile"), ("Real64", "IeeeBinaryFile"),
                                             ("Complex64", "IeeeBinaryFile")])
def test_read_and_define_holes_and_particles(tmp_path, scalar, elements):
    g, ei, ea = philox.system(9, 4, 6, 11)
    x = g.astype(complex) if scalar == "Complex64" else g
    write_cc4s_tensor(tmp_path / "CoulombVertex", x, scalar, elements,
                      ["AuxiliaryField", "State", "State"])
    eps = np.concatenate([ei, ea])
    (tmp_path / "EigenEnergies.yaml").write_text(
        "metaData:\n  fermiEnergy: %.17g\n  energies:\n%s\n"
        % (0.5 * (ei.max() + ea.min()), "\n".join("    - %.17g" % v for v in eps)))
    r, vals = run_steps(tmp_path, """
- name: DefineHolesAndParticles
  in: {fileName: "%s"}
  out: {HoleEigenEnergies: $HoleEigenEnergies, ParticleEigenEnergies: $ParticleEigenEnergies}
- name: Read
  in: {fileName: "%s"}
  out: {destination: $CoulombVertex}
""" % (tmp_path / "EigenEnergies.yaml", tmp_path / "CoulombVertex.yaml") + NORMS)
    assert r.returncode == 0, r.stderr
    assert vals["Norm"] == pytest.approx([np.linalg.norm(g), np.linalg.norm(ei),
                                          np.linalg.norm(ea)], rel=1e-15)
    assert "No: 4" in (tmp_path / "out.yaml").read_text()


def test_read_complex_tensor_and_bad_headers(tmp_path):
    g, _, _ = philox.system(9, 4, 6, 11)
    write_cc4s_tensor(tmp_path / "Z", g + 0.5j * g[::-1], "Complex64", "IeeeBinaryFile")
    r, vals = run_steps(tmp_path, "- name: Read\n  in: {fileName: \"%s\"}\n  out: {destination: $Z}\n"
                        "- name: TensorNorm\n  in: {Data: $Z}\n  out: {Norm: $N}\n"
                        % (tmp_path / "Z.yaml"))
    assert r.returncode == 0, r.stderr
    assert vals["Norm"] == pytest.approx([np.linalg.norm(g + 0.5j * g[::-1])], rel=1e-15)
    (tmp_path / "B.yaml").write_text("version: 200\ntype: Tensor\n")
    r, _ = run_steps(tmp_path, "- name: Read\n  in: {fileName: \"%s\"}\n  out: {destination: $B}\n"
                     % (tmp_path / "B.yaml"))
    assert r.returncode != 0 and "version" in r.stderr


@pytest.mark.gpu
def test_ueg_pipeline_from_files(tmp_path):
    """The reference's UEG test (integration-tests/tests/cc4s/ueg/
    rs1.0-7occ-26virt/in.yaml: DefineHolesAndParticles, Read, integrals,
    CCSD with DIIS(4), (T)) on cc4s-format files holding the oracle's
    real-orbital UEG vertex: the committed MP2, CCSD and (T) energies, and
    the CCSD pair energies against the oracle's."""
    from oracle import ueg
    g, ei, ea, _ = ueg.system(7, 26, 1.0)
    write_cc4s_tensor(tmp_path / "CoulombVertex", g.astype(complex), "Complex64", "IeeeBinaryFile",
                      ["AuxiliaryField", "State", "State"])
    (tmp_path / "EigenEnergies.yaml").write_text(
        "metaData:\n  fermiEnergy: %.17g\n  energies: [%s]\n"
        % (0.5 * (ei.max() + ea.min()), ", ".join("%.17g" % v for v in np.concatenate([ei, ea]))))
    announce("ueg pipeline from cc4s-format files")
    r, vals = run_steps(tmp_path, """
- name: DefineHolesAndParticles
  in: {fileName: "%s"}
  out: {HoleEigenEnergies: $HoleEigenEnergies, ParticleEigenEnergies: $ParticleEigenEnergies}
- name: Read
  in: {fileName: "%s"}
  out: {destination: $CoulombVertex}
- name: CoulombIntegralsFromVertex
  in: {CoulombVertex: $CoulombVertex, HoleEigenEnergies: $HoleEigenEnergies,
       ParticleEigenEnergies: $ParticleEigenEnergies, complex: 0}
  out: {PPHHCoulombIntegrals: $PPHHCoulombIntegrals, HHHPCoulombIntegrals: $HHHPCoulombIntegrals}
- name: CcsdEnergyFromCoulombIntegrals
  in: {mixer: DiisMixer, maxResidua: 4, maxIterations: 10, energyConvergence: 1e-8,
       amplitudesConvergence: 1e-8, HoleEigenEnergies: $HoleEigenEnergies,
       ParticleEigenEnergies: $ParticleEigenEnergies, CoulombVertex: $CoulombVertex,
       PPHHCoulombIntegrals: $PPHHCoulombIntegrals}
  out: {CcsdEnergy: $CcsdEnergy, CcsdSinglesAmplitudes: $T1, CcsdDoublesAmplitudes: $T2,
        PairEnergy: $PairEnergy}
- name: TensorNorm
  in: {Data: $PairEnergy}
  out: {Norm: $NormP}
- name: CcsdPerturbativeTriples
  in: {HoleEigenEnergies: $HoleEigenEnergies, ParticleEigenEnergies: $ParticleEigenEnergies,
       PPHHCoulombIntegrals: $PPHHCoulombIntegrals, HHHPCoulombIntegrals: $HHHPCoulombIntegrals,
       CoulombVertex: $CoulombVertex, CcsdEnergy: $CcsdEnergy, CcsdSinglesAmplitudes: $T1,
       CcsdDoublesAmplitudes: $T2}
  out: {CcsdPerturbativeTriplesEnergy: $CcsdT}
""" % (tmp_path / "EigenEnergies.yaml", tmp_path / "CoulombVertex.yaml"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert abs(vals["CcsdEnergy"][0] - -0.39269658954585018) <= 1e-12
    assert abs(vals["CcsdPerturbativeTriplesEnergy"][0]
               - (-0.39269658954585018 - 0.0063019625641725016)) <= 1e-12
    # PairEnergy of ClusterSinglesDoublesAlgorithm.cxx:218-235 from the same
    # ten DIIS(4) iterations run by the oracle
    from oracle import ccsd
    ref = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_residua=4, max_iterations=10,
                     energy_convergence=1e-8, amplitudes_convergence=1e-8)
    P = ccsd.pair_energy(ref["T1"], ref["T2"], ccsd.integrals_from_vertex(g, 7)["PPHH"])
    assert abs(P.sum() - vals["CcsdEnergy"][0]) <= 1e-12
    assert vals["Norm"][0] == pytest.approx(np.linalg.norm(P), rel=1e-10)


@pytest.mark.gpu
def test_ueg_plane_wave_complex_pipeline_from_files(tmp_path):
    """The reference's UEG test on the complex path: the plane-wave vertex
    (after an orbital gauge and auxiliary rotation, so genuinely complex) in
    a cc4s Complex64 file, CoulombIntegralsFromVertex with complex: 1, CCSD
    with DIIS(4) and CcsdPerturbativeTriplesComplex: the committed MP2, CCSD
    and (T) energies (cc4s.correct.out.yaml:128, 153, 169)."""
    from oracle import ccsd_complex as cz, ueg
    pw = ueg.plane_wave(7, 26, 1.0)
    G, _ = cz.gauge_transform(pw["coeff"].astype(complex), 7, 5)
    ei, ea = pw["eps"][:7], pw["eps"][7:]
    write_cc4s_tensor(tmp_path / "CoulombVertex", G, "Complex64", "IeeeBinaryFile",
                      ["AuxiliaryField", "State", "State"])
    (tmp_path / "EigenEnergies.yaml").write_text(
        "metaData:\n  fermiEnergy: %.17g\n  energies: [%s]\n"
        % (0.5 * (ei.max() + ea.min()), ", ".join("%.17g" % v for v in np.concatenate([ei, ea]))))
    announce("complex ueg pipeline from cc4s-format files")
    r, vals = run_steps(tmp_path, """
- name: DefineHolesAndParticles
  in: {fileName: "%s"}
  out: {HoleEigenEnergies: $HoleEigenEnergies, ParticleEigenEnergies: $ParticleEigenEnergies}
- name: Read
  in: {fileName: "%s"}
  out: {destination: $CoulombVertex}
- name: CoulombIntegralsFromVertex
  in: {CoulombVertex: $CoulombVertex, HoleEigenEnergies: $HoleEigenEnergies,
       ParticleEigenEnergies: $ParticleEigenEnergies, complex: 1}
  out: {PPHHCoulombIntegrals: $PPHH, HHPPCoulombIntegrals: $HHPP, PHHPCoulombIntegrals: $PHHP,
        PHPHCoulombIntegrals: $PHPH, HHHHCoulombIntegrals: $HHHH, HHHPCoulombIntegrals: $HHHP,
        PHHHCoulombIntegrals: $PHHH}
- name: CcsdEnergyFromCoulombIntegrals
  in: {mixer: DiisMixer, maxResidua: 4, maxIterations: 10, energyConvergence: 1e-8,
       amplitudesConvergence: 1e-8, HoleEigenEnergies: $HoleEigenEnergies,
       ParticleEigenEnergies: $ParticleEigenEnergies, CoulombVertex: $CoulombVertex,
       PPHHCoulombIntegrals: $PPHH, HHPPCoulombIntegrals: $HHPP, PHHPCoulombIntegrals: $PHHP,
       PHPHCoulombIntegrals: $PHPH, HHHHCoulombIntegrals: $HHHH, HHHPCoulombIntegrals: $HHHP,
       PHHHCoulombIntegrals: $PHHH}
  out: {CcsdEnergy: $CcsdEnergy, CcsdSinglesAmplitudes: $T1, CcsdDoublesAmplitudes: $T2}
- name: CcsdPerturbativeTriplesComplex
  in: {HoleEigenEnergies: $HoleEigenEnergies, ParticleEigenEnergies: $ParticleEigenEnergies,
       PPHHCoulombIntegrals: $PPHH, PHHHCoulombIntegrals: $PHHH, CoulombVertex: $CoulombVertex,
       CcsdEnergy: $CcsdEnergy, CcsdSinglesAmplitudes: $T1, CcsdDoublesAmplitudes: $T2}
  out: {CcsdPerturbativeTriplesComplexEnergy: $CcsdT}
- name: CcsdPerturbativeTriples
  in: {HoleEigenEnergies: $HoleEigenEnergies, ParticleEigenEnergies: $ParticleEigenEnergies,
       PPHHCoulombIntegrals: $PPHH, HHHPCoulombIntegrals: $HHHP, CoulombVertex: $CoulombVertex,
       CcsdEnergy: $CcsdEnergy, CcsdSinglesAmplitudes: $T1, CcsdDoublesAmplitudes: $T2}
  out: {CcsdPerturbativeTriplesEnergy: $Wrong}
  fallible: 1
""" % (tmp_path / "EigenEnergies.yaml", tmp_path / "CoulombVertex.yaml"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert abs(vals["CcsdEnergy"][0] - -0.39269658954585018) <= 1e-12
    assert abs(vals["CcsdPerturbativeTriplesComplexEnergy"][0]
               - (-0.39269658954585018 - 0.0063019625641725016)) <= 1e-12
    assert "CcsdPerturbativeTriplesComplex" in r.stderr + r.stdout
    assert "CcsdPerturbativeTriplesEnergy" not in vals


@pytest.mark.gpu
def test_runner_factor_ladder_from_files(tmp_path):
    """CcsdEnergyFromCoulombIntegrals with CoulombFactors / FactorOrbitals
    (CcsdEnergyFromCoulombIntegrals.cxx:340-372) read from cc4s files, complex
    factors, factorsSliceFactor: the oracle's energy with the factor-based
    ladder."""
    from oracle import ccsd
    rng = np.random.default_rng(11)
    No, Nv, NG, NR = 4, 12, 16, 30
    Lam = rng.normal(size=(NG, NR)) / np.sqrt(NR * NG)
    Pi = rng.normal(size=(No + Nv, NR)) / 2
    g = np.einsum("GR,pR,qR->Gpq", Lam, Pi, Pi)
    Lz = Lam + 0.3j * rng.normal(size=Lam.shape) / np.sqrt(NR * NG)
    Pz = Pi + 0.3j * rng.normal(size=Pi.shape) / 2
    ei = np.sort(rng.uniform(-2.0, -0.5, No))
    ea = np.sort(rng.uniform(0.2, 3.0, Nv))
    write_cc4s_tensor(tmp_path / "CoulombVertex", g, "Real64", "IeeeBinaryFile")
    write_cc4s_tensor(tmp_path / "Lambda", Lz, "Complex64", "IeeeBinaryFile")
    write_cc4s_tensor(tmp_path / "Pi", Pz, "Complex64", "IeeeBinaryFile")
    (tmp_path / "EigenEnergies.yaml").write_text(
        "metaData:\n  fermiEnergy: %.17g\n  energies: [%s]\n"
        % (0.5 * (ei.max() + ea.min()), ", ".join("%.17g" % v for v in np.concatenate([ei, ea]))))
    announce("runner factor ladder from files")
    r, vals = run_steps(tmp_path, """
- name: DefineHolesAndParticles
  in: {fileName: "%s"}
  out: {HoleEigenEnergies: $I, ParticleEigenEnergies: $A}
- name: Read
  in: {fileName: "%s"}
  out: {destination: $CoulombVertex}
- name: Read
  in: {fileName: "%s"}
  out: {destination: $Lambda}
- name: Read
  in: {fileName: "%s"}
  out: {destination: $Pi}
- name: CoulombIntegralsFromVertex
  in: {CoulombVertex: $CoulombVertex, HoleEigenEnergies: $I, ParticleEigenEnergies: $A}
  out: {PPHHCoulombIntegrals: $V}
- name: CcsdEnergyFromCoulombIntegrals
  in: {mixer: DiisMixer, maxResidua: 6, maxIterations: 15, integralsSliceSize: 5,
       CoulombFactors: $Lambda, FactorOrbitals: $Pi, factorsSliceFactor: 0.4,
       HoleEigenEnergies: $I, ParticleEigenEnergies: $A, PPHHCoulombIntegrals: $V,
       CoulombVertex: $CoulombVertex}
  out: {CcsdEnergy: $E}
""" % tuple(tmp_path / f for f in ("EigenEnergies.yaml", "CoulombVertex.yaml", "Lambda.yaml",
                                    "Pi.yaml")))
    assert r.returncode == 0, r.stdout + r.stderr
    ref = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_residua=6, max_iterations=15, slice_size=5,
                     factors=(Lz, Pz, int(NR * 0.4)))
    assert abs(vals["CcsdEnergy"][0] - ref["energy"]) <= 1e-10


@pytest.mark.gpu
def test_uccsd_from_files(tmp_path):
    """UccsdAmplitudesFromCoulombIntegrals on antisymmetrized spin-orbital
    integral classes read from cc4s-format files (a closed-shell system in
    spin orbitals): the oracle's per-iteration energies, and at convergence
    the restricted CCSD energy."""
    from oracle import ccsd, uccsd
    No, Nv = 2, 5
    g, ei, ea = philox.system(17, No, Nv, 14)
    ints = uccsd.spin_orbital_integrals(g, No, Nv)
    eps = np.concatenate([np.repeat(ei, 2), np.repeat(ea, 2)])
    (tmp_path / "EigenEnergies.yaml").write_text(
        "metaData:\n  fermiEnergy: %.17g\n  energies:\n%s\n"
        % (0.5 * (ei.max() + ea.min()), "\n".join("    - %.17g" % v for v in eps)))
    steps = ["- name: DefineHolesAndParticles\n  in: {fileName: \"%s\"}\n"
             "  out: {HoleEigenEnergies: $I, ParticleEigenEnergies: $A}\n"
             % (tmp_path / "EigenEnergies.yaml")]
    for c in ("PPHH", "HHHH", "PPPP", "HHHP", "PPPH", "PHPH"):
        write_cc4s_tensor(tmp_path / c, ints[c])
        steps.append("- name: Read\n  in: {fileName: \"%s\"}\n  out: {destination: $%s}\n"
                     % (tmp_path / (c + ".yaml"), c))
    steps.append("""- name: UccsdAmplitudesFromCoulombIntegrals
  in: {mixer: DiisMixer, maxResidua: 6, maxIterations: 100, amplitudesConvergence: 1e-10,
       energyConvergence: 1e-12, HoleEigenEnergies: $I, ParticleEigenEnergies: $A,
       PPHHCoulombIntegrals: $PPHH, HHHHCoulombIntegrals: $HHHH, PPPPCoulombIntegrals: $PPPP,
       HHHPCoulombIntegrals: $HHHP, PPPHCoulombIntegrals: $PPPH, PHPHCoulombIntegrals: $PHPH}
  out: {UccsdEnergy: $E, UccsdDoublesAmplitudes: $T2}
""")
    r, vals = run_steps(tmp_path, "".join(steps))
    assert r.returncode == 0, r.stdout + r.stderr
    kw = dict(max_residua=6, max_iterations=100, amplitudes_convergence=1e-10,
              energy_convergence=1e-12)
    ref = uccsd.solve(ints, np.repeat(ei, 2), np.repeat(ea, 2), **kw)
    rr = ccsd.solve(g, ei, ea, **kw)
    E = vals["UccsdEnergy"][-1]
    assert abs(E - ref["energy"]) <= 1e-10 and abs(E - rr["energy"]) <= 1e-9 * abs(rr["energy"])


SPIN_CLASSES = ("PPHH", "HHHH", "PPPP", "HHHP", "PPPH", "PHPH")


def test_unrestricter_and_antisymmetrizer(tmp_path):
    """TensorUnrestricter (spin orbital 2p + s) then TensorAntisymmetrizer
    (V_pqrs - V_pqsr) on spatial integral classes read from files give the
    antisymmetrized spin-orbital classes (compared through their norms and
    the unrestricted eigenenergies)."""
    from oracle import ccsd, uccsd
    No, Nv = 2, 3
    g, ei, ea = philox.system(23, No, Nv, 9)
    sp = ccsd.integrals_from_vertex(g, No)
    sp["PPPP"], sp["PPPH"] = ccsd.vabcd_from_vertex(g, No)
    write_cc4s_tensor(tmp_path / "eps", ei)
    steps = ["- name: Read\n  in: {fileName: \"%s\"}\n  out: {destination: $eps}\n"
             "- name: TensorUnrestricter\n  in: {Data: $eps}\n  out: {Out: $eps}\n"
             "- name: TensorNorm\n  in: {Data: $eps}\n  out: {Norm: $N}\n" % (tmp_path / "eps.yaml")]
    for c in SPIN_CLASSES:
        write_cc4s_tensor(tmp_path / c, sp[c])
        steps.append("- name: Read\n  in: {fileName: \"%s\"}\n  out: {destination: $%s}\n"
                     "- name: TensorUnrestricter\n  in: {Data: $%s}\n  out: {Out: $%s}\n"
                     % (tmp_path / (c + ".yaml"), c, c, c))
    steps.append("- name: TensorAntisymmetrizer\n  in: {%s}\n"
                 % ", ".join("%sCoulombIntegrals: $%s" % (c, c) for c in SPIN_CLASSES))
    for c in SPIN_CLASSES:
        steps.append("- name: TensorNorm\n  in: {Data: $%s}\n  out: {Norm: $N}\n" % c)
    r, vals = run_steps(tmp_path, "".join(steps))
    assert r.returncode == 0, r.stdout + r.stderr
    ref = uccsd.spin_orbital_integrals(g, No, Nv)
    expect = [np.linalg.norm(np.repeat(ei, 2))] + [np.linalg.norm(ref[c]) for c in SPIN_CLASSES]
    assert vals["Norm"] == pytest.approx(expect, rel=1e-14)


@pytest.mark.gpu
def test_uccsd_pipeline_from_vertex_matches_restricted_ccsd(tmp_path):
    """The reference's spin-orbital workflow (integrals, TensorUnrestricter,
    TensorAntisymmetrizer, UccsdAmplitudesFromCoulombIntegrals) on a
    closed-shell system converges to CcsdEnergyFromCoulombIntegrals' energy."""
    conv = "mixer: DiisMixer, maxResidua: 6, maxIterations: 100, amplitudesConvergence: 1e-10, " \
           "energyConvergence: 1e-12"
    steps = ["""- name: GenerateSyntheticSystem
  in: {seed: 31, No: 3, Nv: 9, NG: 30}
  out: {CoulombVertex: $G, HoleEigenEnergies: $I, ParticleEigenEnergies: $A}
- name: CoulombIntegralsFromVertex
  in: {CoulombVertex: $G, HoleEigenEnergies: $I, ParticleEigenEnergies: $A}
  out: {%s}
- name: CcsdEnergyFromCoulombIntegrals
  in: {%s, HoleEigenEnergies: $I, ParticleEigenEnergies: $A, PPHHCoulombIntegrals: $PPHH,
       CoulombVertex: $G}
  out: {CcsdEnergy: $Er}
""" % (", ".join("%sCoulombIntegrals: $%s" % (c, c) for c in SPIN_CLASSES), conv)]
    for x in ["I", "A"] + list(SPIN_CLASSES):
        steps.append("- name: TensorUnrestricter\n  in: {Data: $%s}\n  out: {Out: $%s}\n" % (x, x))
    steps.append("- name: TensorAntisymmetrizer\n  in: {%s}\n"
                 % ", ".join("%sCoulombIntegrals: $%s" % (c, c) for c in SPIN_CLASSES))
    steps.append("- name: UccsdAmplitudesFromCoulombIntegrals\n  in: {%s, HoleEigenEnergies: $I, "
                 "ParticleEigenEnergies: $A, %s}\n  out: {UccsdEnergy: $Eu}\n"
                 % (conv, ", ".join("%sCoulombIntegrals: $%s" % (c, c) for c in SPIN_CLASSES)))
    announce("uccsd pipeline from the vertex")
    r, vals = run_steps(tmp_path, "".join(steps))
    assert r.returncode == 0, r.stdout + r.stderr
    er, eu = vals["CcsdEnergy"][-1], vals["UccsdEnergy"][-1]
    assert abs(eu - er) <= 1e-9 * abs(er), (eu, er)


def test_consumer_steps_reject_what_they_do_not_build(tmp_path):
    """Errors of the direct-ring / spin-orbital steps surface as step failures
    before any device work: an order-3 tensor for TensorUnrestricter, a class
    with no exchange partner for TensorAntisymmetrizer, onlyPPL for UCCSD and
    complex integrals for the direct-ring step."""
    from oracle import ccsd
    g, ei, ea = philox.system(3, 2, 3, 5)
    write_cc4s_tensor(tmp_path / "G", g)
    r, _ = run_steps(tmp_path, "- name: Read\n  in: {fileName: \"%s\"}\n  out: {destination: $G}\n"
                     "- name: TensorUnrestricter\n  in: {Data: $G}\n  out: {Out: $U}\n"
                     % (tmp_path / "G.yaml"))
    assert r.returncode != 0 and "not supported" in r.stderr
    V = ccsd.integrals_from_vertex(g, 2)["PHPH"]
    write_cc4s_tensor(tmp_path / "X", V)
    r, _ = run_steps(tmp_path, "- name: Read\n  in: {fileName: \"%s\"}\n  out: {destination: $X}\n"
                     "- name: TensorAntisymmetrizer\n  in: {PHPHCoulombIntegrals: $X}\n"
                     % (tmp_path / "X.yaml"))
    assert r.returncode != 0 and "could not be antisymmetrized" in r.stderr
    r, _ = run_steps(tmp_path, "- name: UccsdAmplitudesFromCoulombIntegrals\n  in: {onlyPPL: 1}\n")
    assert r.returncode != 0 and "onlyPPL" in r.stderr
    P = ccsd.integrals_from_vertex(g, 2)["PPHH"]
    write_cc4s_tensor(tmp_path / "Z", P + 0.3j * P, "Complex64")  # non-zero imaginary part
    write_cc4s_tensor(tmp_path / "I", ei)
    write_cc4s_tensor(tmp_path / "A", ea)
    r, _ = run_steps(tmp_path, "".join(
        "- name: Read\n  in: {fileName: \"%s\"}\n  out: {destination: $%s}\n" % (tmp_path / (n + ".yaml"), n)
        for n in ("Z", "I", "A")) +
        "- name: DrccdEnergyFromCoulombIntegrals\n  in: {HoleEigenEnergies: $I, ParticleEigenEnergies: $A, "
        "PPHHCoulombIntegrals: $Z, PHHPCoulombIntegrals: $Z, HHPPCoulombIntegrals: $Z}\n"
        "  out: {DrccdEnergy: $E}\n")
    assert r.returncode != 0 and "real integrals" in r.stderr


def test_antisymmetrizer_with_all_sixteen_classes(tmp_path):
    """The reference's full set of sixteen classes (as in its spin-orbital
    workflows), unrestricted and antisymmetrized in the runner, equals the
    blocks of the antisymmetrized spin-orbital tensor <pq||rs> (norms, and
    the class picks among several exchange partners)."""
    from oracle import ccsd
    No, Nv = 2, 3
    g, ei, ea = philox.system(29, No, Nv, 9)
    eri = ccsd.ein("Gpr,Gqs->pqrs", g, g)           # spatial <pq|rs>
    H, P = slice(0, No), slice(No, No + Nv)
    names = ["HHHH", "HHHP", "HHPH", "HHPP", "HPHH", "HPHP", "HPPH", "HPPP",
             "PHHH", "PHHP", "PHPH", "PHPP", "PPHH", "PPHP", "PPPH", "PPPP"]
    blk = lambda c: eri[tuple(H if x == "H" else P for x in c)]
    # spin-orbital antisymmetrized reference, 2p + s inside each block
    n = 2 * (No + Nv)
    so = [(p, s) for p in range(No) for s in (0, 1)] + [(p, s) for p in range(No, No + Nv) for s in (0, 1)]
    sp = np.array([p for p, _ in so]); sg = np.array([s for _, s in so])
    full = eri[np.ix_(sp, sp, sp, sp)] * ((sg[:, None, None, None] == sg[None, None, :, None])
                                          & (sg[None, :, None, None] == sg[None, None, None, :]))
    asym = full - np.swapaxes(full, 2, 3)
    h2, p2 = slice(0, 2 * No), slice(2 * No, n)
    steps = []
    for c in names:
        write_cc4s_tensor(tmp_path / c, blk(c))
        steps.append("- name: Read\n  in: {fileName: \"%s\"}\n  out: {destination: $%s}\n"
                     "- name: TensorUnrestricter\n  in: {Data: $%s}\n  out: {Out: $%s}\n"
                     % (tmp_path / (c + ".yaml"), c, c, c))
    steps.append("- name: TensorAntisymmetrizer\n  in: {%s}\n"
                 % ", ".join("%sCoulombIntegrals: $%s" % (c, c) for c in names))
    for c in names:
        steps.append("- name: TensorNorm\n  in: {Data: $%s}\n  out: {Norm: $N}\n" % c)
    r, vals = run_steps(tmp_path, "".join(steps))
    assert r.returncode == 0, r.stdout + r.stderr
    expect = [np.linalg.norm(asym[tuple(h2 if x == "H" else p2 for x in c)]) for c in names]
    assert vals["Norm"] == pytest.approx(expect, rel=1e-13)


def _ip_eom_system(no, n, seed):
    """Generic real spin-orbital integrals <pq||rs> from a random vertex, `no`
    electrons, and their converged spin-orbital CCSD amplitudes (oracle)."""
    from oracle import uccsd
    rng = np.random.default_rng(seed)
    g = rng.normal(size=(12, n, n)) * 0.25
    g = 0.5 * (g + np.transpose(g, (0, 2, 1)))
    eri = np.einsum("Gpr,Gqs->pqrs", g, g)
    asym = eri - np.swapaxes(eri, 2, 3)
    eps = np.concatenate([np.linspace(-1.3, -0.8, no), np.sort(rng.uniform(0.3, 2.0, n - no))])
    ints = uccsd.classes(asym, no)
    r = uccsd.solve(ints, eps[:no], eps[no:], max_residua=6, max_iterations=300,
                    amplitudes_convergence=1e-12, energy_convergence=1e-14)
    assert r["converged"]
    return asym, eps, ints, r["T1"], r["T2"]


def _ip_eom_steps(tmp_path, asym, eps, no, T1, T2, args, ea=False):
    h, p = slice(0, no), slice(no, asym.shape[0])
    tensors = dict(I=eps[:no], A=eps[no:], T1=T1, T2=T2, HHPP=asym[h, h, p, p], HHHP=asym[h, h, h, p])
    if ea:
        tensors.update(PPPP=asym[p, p, p, p], PPPH=asym[p, p, p, h], PHPH=asym[p, h, p, h])
    else:
        tensors.update(HHHH=asym[h, h, h, h], HPPP=asym[h, p, p, p], HPHP=asym[h, p, h, p])
    classes = [c for c in tensors if len(c) == 4]
    steps = []
    for name, x in tensors.items():
        write_cc4s_tensor(tmp_path / name, np.ascontiguousarray(x))
        steps.append("- name: Read\n  in: {fileName: \"%s\"}\n  out: {destination: $%s}\n"
                     % (tmp_path / (name + ".yaml"), name))
    steps.append("- name: UCcsd%sEquationOfMotionDavidson\n  in: {HoleEigenEnergies: $I, "
                 "ParticleEigenEnergies: $A, SinglesAmplitudes: $T1, DoublesAmplitudes: $T2, %s, %s}\n"
                 % ("EA" if ea else "IP", ", ".join("%sCoulombIntegrals: $%s" % (c, c) for c in classes), args))
    p = tmp_path / "in.yaml"
    p.write_text("".join(steps))
    return subprocess.run([RUNNER, "-i", str(p), "-o", str(tmp_path / "out.yaml")],
                          capture_output=True, text=True, timeout=600, cwd=str(tmp_path))


def test_ip_eom_step_rejects_what_it_does_not_build(tmp_path):
    """UCcsdIPEquationOfMotionDavidson: non-canonical orbitals, refreshIterations
    and the eigenvector / density printing ranges fail the step before any
    device work; amplitudes of the wrong shape too."""
    r, _ = run_steps(tmp_path, "- name: UCcsdIPEquationOfMotionDavidson\n  in: {HHFockMatrix: 1}\n")
    assert r.returncode != 0 and "Fock" in r.stderr
    r, _ = run_steps(tmp_path, "- name: UCcsdIPEquationOfMotionDavidson\n"
                     "  in: {refreshIterations: \"2,4\"}\n")
    assert r.returncode != 0 and "refreshIterations is not supported" in r.stderr
    ei, ea = np.array([-1.0, -0.5]), np.array([0.5, 1.0, 1.5])
    for n, x in dict(I=ei, A=ea, T1=np.zeros((2, 3))).items():  # T1 must be [Nv, No]
        write_cc4s_tensor(tmp_path / n, x)
    r, _ = run_steps(tmp_path, "".join(
        "- name: Read\n  in: {fileName: \"%s\"}\n  out: {destination: $%s}\n" % (tmp_path / (n + ".yaml"), n)
        for n in ("I", "A", "T1")) +
        "- name: UCcsdIPEquationOfMotionDavidson\n  in: {HoleEigenEnergies: $I, "
        "ParticleEigenEnergies: $A, SinglesAmplitudes: $T1, DoublesAmplitudes: $T1}\n")
    assert r.returncode != 0 and "amplitudes must be real" in r.stderr
    r, _ = run_steps(tmp_path, "- name: UCcsdEAEquationOfMotionDavidson\n"
                     "  in: {printEigenvectorsRange: \"1\"}\n")
    assert r.returncode != 0 and "printEigenvectorsRange is not supported" in r.stderr


@pytest.mark.gpu
def test_ip_eom_davidson_step_matches_full_ci_and_the_oracle(tmp_path):
    """The IP-EOM-CCSD Davidson step from files (Read of the reference's
    classes and amplitudes): two electrons reproduce full CI's ionization
    energies; four electrons match the oracle's dense spectrum."""
    from oracle import eom
    asym, eps, ints, T1, T2 = _ip_eom_system(2, 8, 7)
    h = np.diag(eps) - np.einsum("pkqk->pq", asym[:, :2, :, :2])
    dets = [(p, q) for p in range(8) for q in range(p + 1, 8)]
    H = np.array([[h[p, s] * (q == t) - h[p, t] * (q == s) - h[q, s] * (p == t)
                   + h[q, t] * (p == s) + asym[p, q, s, t] for (s, t) in dets] for (p, q) in dets])
    exact = np.sort(np.linalg.eigvalsh(h) - np.linalg.eigvalsh(H)[0])
    announce("ip-eom davidson step, 2 electrons in 8 spin orbitals")
    r = _ip_eom_steps(tmp_path, asym, eps, 2, T1, T2,
                      "eigenstates: 2, energyConvergence: 0.0, amplitudesConvergence: 1e-20, "
                      "maxIterations: 40")
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.loadtxt(tmp_path / "IPEomEnergies.dat", ndmin=2)
    assert np.abs(got[:, 1] - exact[:2]).max() <= 1e-9 and np.abs(got[:, 2]).max() == 0.0
    asym, eps, ints, T1, T2 = _ip_eom_system(4, 10, 8)
    ref = eom.ip_eigenvalues(ints, eps[:4], eps[4:], T1, T2)
    announce("ip-eom davidson step, 4 electrons in 10 spin orbitals")
    r = _ip_eom_steps(tmp_path, asym, eps, 4, T1, T2,
                      "eigenstates: 3, energyConvergence: 0.0, amplitudesConvergence: 1e-20, "
                      "maxIterations: 60")
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.loadtxt(tmp_path / "IPEomEnergies.dat", ndmin=2)
    assert np.abs(got[:, 1] - ref[:3]).max() <= 1e-9


@pytest.mark.gpu
def test_ea_eom_davidson_step_matches_full_ci(tmp_path):
    """The EA-EOM-CCSD Davidson step from files: with two virtual spin
    orbitals its roots equal the oracle's dense EA spectrum, which equals full
    CI there; with four virtuals the lowest roots match the oracle."""
    from oracle import eom
    for no, n, k in ((3, 5, 3), (3, 7, 3)):
        asym, eps, ints, T1, T2 = _ip_eom_system(no, n, 11)
        ref = eom.ea_eigenvalues(ints, eps[:no], eps[no:], T1, T2)
        announce("ea-eom davidson step, %d electrons in %d spin orbitals" % (no, n))
        r = _ip_eom_steps(tmp_path, asym, eps, no, T1, T2,
                          "eigenstates: %d, energyConvergence: 0.0, amplitudesConvergence: 1e-20, "
                          "maxIterations: 60" % k, ea=True)
        assert r.returncode == 0, r.stdout + r.stderr
        got = np.loadtxt(tmp_path / "EAEomEnergies.dat", ndmin=2)
        assert np.abs(got[:, 1] - ref[:k]).max() <= 1e-9
