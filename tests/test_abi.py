"""C-ABI library checks that need no GPU: it loads, exports every symbol
include/sisi_engine.h declares, its host generator matches the oracle's
restatement, and errors come back as status codes."""

import ctypes
import os
import re

import numpy as np

from conftest import ROOT
from oracle import philox


def declared_functions():
    src = open(os.path.join(ROOT, "include", "sisi_engine.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sisi_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_header():
    from sisi4s_amd import engine
    bound = {n for n, _, _ in engine.SIGNATURES}
    assert bound == set(declared_functions())


def test_host_generator_matches_oracle(lib):
    from sisi4s_amd import engine
    for seed, No, Nv, NG in [(20261202, 10, 50, 300), (7, 3, 5, 9)]:
        g, ei, ea = engine.generate_system(seed, No, Nv, NG)
        g2, ei2, ea2 = philox.system(seed, No, Nv, NG)
        np.testing.assert_array_equal(ei, ei2)
        np.testing.assert_array_equal(ea, ea2)
        assert np.abs(g - g2).max() <= 4e-16 * np.abs(g2).max()


def test_invalid_arguments_return_status(lib):
    from sisi4s_amd import engine
    d = ctypes.c_double()
    assert lib.sisi_generate_system(1, 0, 5, 5, None, None, None) == engine.SISI_ERR_INVALID
    assert b"positive" in lib.sisi_last_error()
    assert lib.sisi_gemm_host(0, 0, 0, -1, 1, 1, 1.0, None, 1, None, 1, 0.0, None, 1) == \
        engine.SISI_ERR_INVALID
    assert lib.sisi_gemm_host(7, 0, 0, 1, 1, 1, 1.0, None, 1, None, 1, 0.0, None, 1) == \
        engine.SISI_ERR_INVALID
    assert lib.sisi_mp2_energy(None, ctypes.byref(d)) == engine.SISI_ERR_INVALID
    assert b"null engine" in lib.sisi_last_error()
    assert lib.sisi_engine_create(None, None) == engine.SISI_ERR_INVALID


def test_opts_defaults_follow_reference(lib):
    from sisi4s_amd import engine
    o = engine.default_opts()
    # ClusterSinglesDoublesAlgorithm.hpp:42-46, DiisMixer.cxx:55, LinearMixer.cxx:17
    assert o.max_iterations == 16 and o.amplitudes_convergence == 1e-5
    assert o.energy_convergence == 1e-6 and o.level_shift == 0.0
    assert o.mixer == engine.MIXER_LINEAR and o.max_residua == 4 and o.mixing_ratio == 1.0
