"""CPU FP64 oracle for the CCSD / (T) hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``sisi4s_amd``) imports,
links or runs this package.  It is used by ``tests/`` as the checker, by
``__graft_entry__.smoke()`` to check one small GPU invocation, and by the
``cpu_baseline`` leg of ``bench.py`` as the timed CPU path.

It is a numpy restatement of the reference's real (FP64) code path:

* ``philox``       -- the seeded Philox-4x64-10 input generator (the product
                      has its own C++ implementation; this one checks it).
* ``ccsd``         -- CoulombIntegralsFromVertex (real), the vertex-sliced CCSD
                      residual, the amplitude update, DIIS / linear mixers,
                      the energy and the iteration driver.
* ``triples``      -- CcsdPerturbativeTriples (i<=j<=k slices, Helgaker form).
* ``formulations`` -- independent cross-formulations from the reference tree
                      used only to pin the restatement: full-Vabcd CCSD
                      residual, the unsliced 6-index (T), the MP2 closed form.

Every function cites the reference file:line it follows.  Tensors use the
reference's logical index order (T2[a,b,i,j], Gamma[G,p,q]); the memory layout
of the C-ABI buffers is first-index-fastest (Fortran order), converted by
``to_abi`` / ``from_abi``.
"""

import numpy as np


def to_abi(x):
    """Flatten a logical tensor into the first-index-fastest buffer layout."""
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel(order="F"))


def from_abi(buf, shape):
    """Inverse of ``to_abi``."""
    return np.asarray(buf, dtype=np.float64).reshape(shape, order="F")
