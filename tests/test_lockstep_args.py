"""Argument checks of the lockstep batched solver (sisi_ccsd_solve_lockstep),
all raised before any device work, so they run without a GPU."""

import numpy as np
import pytest

from oracle import philox


def test_lockstep_rejects_non_finite_input_and_names_the_system():
    from sisi4s_amd import engine
    g, ei, ea = philox.system(1, 3, 8, 32)
    bad = ea.copy()
    bad[0] = np.nan
    with pytest.raises(engine.EngineError) as ex:
        engine.ccsd_solve_lockstep([(g, ei, ea), (g, ei, bad)],
                                   engine.default_opts(max_iterations=2))
    assert "system 1" in str(ex.value) and "ParticleEigenEnergies" in str(ex.value)


@pytest.mark.parametrize("kw,msg", [
    (dict(precision="bf16x3", refine_interval=4), "refine_interval"),
    (dict(precision="bf16x3", bulk_refine_interval=4), "bulk_refine_interval"),
    (dict(mixer="DiisMixer", max_residua=31), "maxResidua"),
    (dict(mixer="DiisMixer", max_residua=0), "maxResidua"),
])
def test_lockstep_rejects_unsupported_options(kw, msg):
    from sisi4s_amd import engine
    g, ei, ea = philox.system(1, 3, 8, 32)
    with pytest.raises(engine.EngineError) as ex:
        engine.ccsd_solve_lockstep([(g, ei, ea)], engine.default_opts(**kw))
    assert msg in str(ex.value)


def test_lockstep_empty_batch():
    from sisi4s_amd import engine
    E, reps, st = engine.ccsd_solve_lockstep([], engine.default_opts())
    assert len(E) == 0 and reps == [] and st["launches"] == 0
