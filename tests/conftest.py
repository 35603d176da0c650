import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity through the C ABI")


def announce(msg):
    """Name the GPU case before it runs (a fault may end the process)."""
    print("[case] " + msg, flush=True)


def random_amplitudes(Nv, No, seed=1, scale=0.1):
    """Random T1 and a T2 with the closed-shell symmetry T_abij = T_baji."""
    rng = np.random.default_rng(seed)
    T1 = scale * rng.standard_normal((Nv, No))
    T2 = scale * rng.standard_normal((Nv, Nv, No, No))
    T2 = T2 + T2.transpose(1, 0, 3, 2)
    return T1, T2


@pytest.fixture(scope="session")
def lib():
    from sisi4s_amd import engine
    return engine.load()
