"""(T) parity: the HIP path through the C ABI against the oracle's literal
restatement of CcsdPerturbativeTriples on the same inputs."""

import numpy as np
import pytest

from conftest import announce, random_amplitudes
from oracle import philox, triples

pytestmark = pytest.mark.gpu


def make(seed, No, Nv, NG):
    from sisi4s_amd import engine
    g, ei, ea = philox.system(seed, No, Nv, NG)
    e = engine.Engine(No, Nv, NG)
    e.set_system(g, ei, ea)
    return e, g, ei, ea


@pytest.mark.parametrize("No,Nv,NG", [(3, 5, 9), (4, 10, 20), (5, 17, 30)])
def test_triples_matches_oracle(No, Nv, NG):
    e, g, ei, ea = make(40 + No, No, Nv, NG)
    T1, T2 = random_amplitudes(Nv, No, seed=No * Nv)
    announce("triples No=%d Nv=%d NG=%d" % (No, Nv, NG))
    et = e.triples(T1, T2)
    ref = triples.triples_energy(T1, T2, g, ei, ea)
    assert abs(et - ref) <= 1e-11 * abs(ref)


def test_triples_blocks_partition_the_sum():
    No, Nv, NG = 4, 10, 20
    e, g, ei, ea = make(50, No, Nv, NG)
    T1, T2 = random_amplitudes(Nv, No, seed=3)
    blocks = triples.sorted_blocks(No)
    announce("triples blocks")
    total = e.triples(T1, T2)
    parts = [e.triples(T1, T2, blocks=[b]) for b in blocks[:6]]
    rest = e.triples(T1, T2, blocks=blocks[6:])
    assert abs(sum(parts) + rest - total) <= 1e-12 * abs(total)
    inp = triples.TriplesInputs(T1, T2, g, ei, ea)
    for b, p in zip(blocks[:6], parts):
        ref = triples.block_energy(inp, b)
        assert abs(p - ref) <= 1e-12 * abs(total) + 1e-11 * abs(ref)


def test_triples_reduced_precision():
    No, Nv, NG = 4, 24, 40
    e, g, ei, ea = make(51, No, Nv, NG)
    T1, T2 = random_amplitudes(Nv, No, seed=4, scale=0.02)
    announce("triples fp32 / bf16x3")
    e64 = e.triples(T1, T2, precision="fp64")
    e32 = e.triples(T1, T2, precision="fp32")
    eb3 = e.triples(T1, T2, precision="bf16x3")
    assert abs(e32 - e64) <= 1e-5 * abs(e64)
    assert abs(eb3 - e64) <= 1e-4 * abs(e64)


def test_triples_after_no10_nv50_ccsd():
    """(T) on No=10, Nv=50, NG=300 amplitudes from the engine's own CCSD run."""
    from sisi4s_amd import engine
    No, Nv, NG = 10, 50, 300
    e, g, ei, ea = make(20261202, No, Nv, NG)
    announce("No=10 Nv=50 ccsd 10 iterations + (T)")
    e.ccsd_solve(engine.default_opts(mixer="DiisMixer", max_residua=6, max_iterations=10))
    T1, T2 = e.amplitudes()
    et = e.triples()
    ref = triples.triples_energy(T1, T2, g, ei, ea)
    assert abs(et - ref) <= 1e-10 * abs(ref)
