"""Writes tests/golden/tiny_system.npz from the oracle (a regression fixture:
inputs, one residual, the (T) energy and 8 DIIS(6) iteration energies of a
tiny seeded system).  The oracle itself is pinned by tests/test_oracle.py."""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ccsd, philox, triples  # noqa: E402


def main():
    seed, No, Nv, NG, s = 424242, 3, 8, 24, 3
    g, ei, ea = philox.system(seed, No, Nv, NG)
    rng = np.random.default_rng(7)
    T1 = 0.05 * rng.standard_normal((Nv, No))
    T2 = 0.05 * rng.standard_normal((Nv, Nv, No, No))
    T2 = T2 + T2.transpose(1, 0, 3, 2)
    ints = ccsd.integrals_from_vertex(g, No)
    R1, R2 = ccsd.residual(1, T1, T2, g, ints, slice_size=s)
    et = triples.triples_energy(T1, T2, g, ei, ea)
    r = ccsd.solve(g, ei, ea, mixer="DiisMixer", max_residua=6, max_iterations=8,
                   amplitudes_convergence=0.0, energy_convergence=0.0)
    np.savez_compressed(os.path.join(HERE, "tiny_system.npz"), seed=seed,
                        dims=np.array([No, Nv, NG]), slice=s, gamma=g, eps_i=ei, eps_a=ea,
                        T1=T1, T2=T2, R1=R1, R2=R2, E_T=et, energies=np.array(r["energies"]))


if __name__ == "__main__":
    main()
