"""MI355X-native CCSD / (T) contraction engine with the sisi4s plugin surface.

The compute path is the HIP library ``libsisi_engine.so`` (built in-tree for
gfx950 by ``sisi4s_amd.build``) behind the C ABI declared in
``include/sisi_engine.h``; ``sisi4s_amd.engine`` binds it with ctypes and
``sisi4s_amd.algorithms`` mirrors the reference's algorithm plugins.
"""

from . import engine  # noqa: F401

__all__ = ["engine"]
