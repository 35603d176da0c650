"""Lockstep batched CCSD (sisi_ccsd_solve_lockstep): every launch of an
iteration is one grouped launch over all systems.

Checked against the oracle (small systems, DIIS and linear mixers, level
shift, the ppl / distinguishable variants), against the threaded batch
(iteration counts and energies of converged runs), the mixed schedule
against FP64, and the claim that the launch count does not grow with the
number of systems (a batch and the same batch twice launch the same
kernels).  Tolerances: 1e-8 Eh against the oracle, 1e-10 Eh against the
threaded FP64 path (same arithmetic, reductions in another order), 1e-6 Eh
mixed against FP64."""

import time

import numpy as np
import pytest

from conftest import announce
from oracle import ccsd, philox

pytestmark = pytest.mark.gpu


def uneven(n, base, no=(3, 6), nv=(10, 24)):
    rng = np.random.default_rng(base)
    out = []
    for s in range(n):
        No = int(rng.integers(no[0], no[1] + 1))
        Nv = int(rng.integers(nv[0], nv[1] + 1))
        out.append((base + s, No, Nv, 4 * Nv))
    return out


@pytest.mark.parametrize("mixer,shift,ratio", [("DiisMixer", 0.0, 1.0), ("DiisMixer", 2.0, 1.0),
                                               ("LinearMixer", 4.0, 0.5)])
def test_lockstep_matches_oracle(mixer, shift, ratio):
    """Ten iterations per system (no convergence test) equal the oracle's."""
    from sisi4s_amd import engine
    specs = uneven(6, 20261202)
    systems = [philox.system(*sp) for sp in specs]
    kw = dict(mixer=mixer, max_residua=6, max_iterations=10, level_shift=shift,
              mixing_ratio=ratio, amplitudes_convergence=0.0, energy_convergence=0.0)
    announce("lockstep 6 systems %s shift %g" % (mixer, shift))
    E, reps, st = engine.ccsd_solve_lockstep(systems, engine.default_opts(**kw))
    assert st["iterations"] == 10
    for sp, (g, ei, ea), e, r in zip(specs, systems, E, reps):
        assert r["iterations"] == 10 and r["status"] == 0
        ref = ccsd.solve(g, ei, ea, **kw)
        assert abs(e - ref["energy"]) <= 1e-8, (sp, e, ref["energy"])


@pytest.mark.parametrize("ppl,dist", [(0, 0), (1, 1)])
def test_lockstep_residual_variants_match_oracle(ppl, dist):
    from sisi4s_amd import engine
    specs = uneven(3, 77)
    systems = [philox.system(*sp) for sp in specs]
    kw = dict(mixer="DiisMixer", max_residua=4, max_iterations=6,
              amplitudes_convergence=0.0, energy_convergence=0.0)
    announce("lockstep variants ppl=%d distinguishable=%d" % (ppl, dist))
    E, reps, _ = engine.ccsd_solve_lockstep(systems, engine.default_opts(ppl=ppl,
                                                                         distinguishable=dist,
                                                                         **kw))
    for (g, ei, ea), e in zip(systems, E):
        ref = ccsd.solve(g, ei, ea, ppl=bool(ppl), distinguishable=bool(dist), **kw)
        assert abs(e - ref["energy"]) <= 1e-8


def test_lockstep_converged_runs_match_threaded_batch():
    """Converging FP64 runs stop at the same iteration with the same energy
    as the threaded batch; systems leave the lockstep batch one by one."""
    from sisi4s_amd import engine
    specs = uneven(8, 4242, no=(3, 8), nv=(10, 40))
    systems = [philox.system(*sp) for sp in specs]
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=300, level_shift=2.0,
              amplitudes_convergence=1e-8, energy_convergence=1e-10)
    announce("lockstep vs threaded, 8 converging systems")
    Et, rt = engine.ccsd_solve_batched(systems, engine.default_opts(**kw), workers=4)
    El, rl, st = engine.ccsd_solve_lockstep(systems, engine.default_opts(**kw))
    print("threaded iterations", [r["iterations"] for r in rt], "lockstep",
          [r["iterations"] for r in rl], flush=True)
    # (one of these systems does not converge in 300 iterations on either path)
    assert sum(r["converged"] for r in rt) >= 6
    for a, b, ea_, eb in zip(rt, rl, Et, El):
        assert a["converged"] == b["converged"] and a["iterations"] == b["iterations"]
        if a["converged"]:
            assert abs(ea_ - eb) <= 1e-10
    assert st["iterations"] == max(r["iterations"] for r in rl)
    assert st["plans"] >= 2  # the schedule shrank as systems converged


def test_lockstep_mixed_within_1uEh_of_fp64():
    from sisi4s_amd import engine
    specs = uneven(6, 99)
    systems = [philox.system(*sp) for sp in specs]
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=300, level_shift=2.0,
              amplitudes_convergence=1e-8, energy_convergence=1e-10)
    announce("lockstep mixed vs fp64, 6 systems")
    e64, r64, _ = engine.ccsd_solve_lockstep(systems, engine.default_opts(**kw))
    emx, rmx, st = engine.ccsd_solve_lockstep(
        systems, engine.default_opts(precision="bf16x3", switch_threshold=1e-5, **kw))
    assert all(r["converged"] for r in r64) and all(r["converged"] for r in rmx)
    assert all(r["fp64_iterations"] < r["iterations"] for r in rmx)
    assert all(r["precision"] == engine.FP64 for r in rmx)
    assert float(np.max(np.abs(e64 - emx))) <= 1e-6


def test_lockstep_launches_do_not_grow_with_systems():
    """The same batch twice runs the same grouped launches as the batch once
    (twice the records), with identical energies."""
    from sisi4s_amd import engine
    specs = uneven(4, 5)
    systems = [philox.system(*sp) for sp in specs]
    kw = dict(mixer="DiisMixer", max_residua=4, max_iterations=5,
              amplitudes_convergence=0.0, energy_convergence=0.0)
    announce("lockstep launch counts, 4 and 8 systems")
    E1, _, s1 = engine.ccsd_solve_lockstep(systems, engine.default_opts(**kw))
    E2, _, s2 = engine.ccsd_solve_lockstep(systems + systems, engine.default_opts(**kw))
    assert s1["launches"] == s2["launches"]
    assert s2["records"] == 2 * s1["records"]
    assert np.array_equal(np.concatenate([E1, E1]), E2)


def test_lockstep_64_systems_config():
    """The 64-system batch (No in [4,16], Nv in [30,120], NG = 4 Nv, DIIS(6),
    levelShift 2, to |dT|/|T| <= 1e-8) in lockstep: FP32 bulk then FP64,
    every energy within 1 uEh of the threaded FP64 batch."""
    from sisi4s_amd import engine
    rng = np.random.default_rng(20261202)
    specs = []
    for k in range(64):
        no = int(rng.integers(4, 17))
        nv = int(rng.integers(30, 121))
        specs.append((20261202 + k, no, nv, 4 * nv))
    systems = [engine.generate_system(*sp) for sp in specs]
    kw = dict(mixer="DiisMixer", max_residua=6, max_iterations=300, level_shift=2.0,
              amplitudes_convergence=1e-8, energy_convergence=1e-10)
    announce("64 systems, threaded fp64")
    e64, r64 = engine.ccsd_solve_batched(systems, engine.default_opts(**kw), workers=8)
    announce("64 systems, lockstep mixed")
    t0 = time.perf_counter()
    emx, rmx, st = engine.ccsd_solve_lockstep(
        systems, engine.default_opts(precision="fp32", switch_threshold=1e-5, **kw))
    dt = time.perf_counter() - t0
    print("lockstep 64 systems: %.2f s, %d iterations, %.0f launches, %.0f records, "
          "%d plans, max |dE| %.2e" % (dt, st["iterations"], st["launches"], st["records"],
                                      st["plans"], float(np.max(np.abs(emx - e64)))), flush=True)
    assert all(r["converged"] for r in rmx)
    assert float(np.max(np.abs(emx - e64))) <= 1e-6
