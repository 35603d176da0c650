"""C-ABI library checks that need no GPU: it loads, exports every symbol
include/sisi_engine.h declares, its host generator matches the oracle's
restatement, and errors come back as status codes."""

import ctypes
import os
import re

import numpy as np
import pytest

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


def test_host_ueg_generator_matches_oracle(lib):
    from oracle import ueg
    from sisi4s_amd import engine
    for No, Nv, rs in [(7, 26, 1.0), (7, 12, 2.0), (1, 18, 0.5)]:
        g, ei, ea = engine.generate_ueg(No, Nv, rs)
        g2, ei2, ea2, _ = ueg.system(No, Nv, rs)
        assert g.shape == g2.shape
        np.testing.assert_allclose(ei, ei2, rtol=0, atol=4e-15)
        np.testing.assert_allclose(ea, ea2, rtol=0, atol=4e-15)
        assert np.abs(g - g2).max() <= 4e-15
    # open shells cannot be written in real orbitals; a wrong buffer size is refused
    with pytest.raises(engine.EngineError, match="closed shells"):
        engine.generate_ueg(5, 28, 1.0)
    ng = ctypes.c_int(3)
    buf = np.zeros(3 * 33 * 33)
    rc = lib.sisi_generate_ueg(7, 26, 1.0, 1, -1.0, ctypes.byref(ng),
                               buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None, None)
    assert rc == engine.SISI_ERR_INVALID and b"NG = 3" in lib.sisi_last_error()


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
    assert o.refine_interval == 0 and o.bulk_refine_interval == 0
