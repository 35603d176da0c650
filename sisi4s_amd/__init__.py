"""MI355X-native CCSD / (T) contraction engine with the sisi4s plugin surface.

The compute path is the HIP library ``libsisi_engine.so`` (built in-tree for
gfx950 by ``sisi4s_amd.build``) behind the C ABI declared in
``include/sisi_engine.h``; ``sisi4s_amd.engine`` binds it with ctypes, and
the C++ step runner under ``sisi4s_amd/host`` (``sisi4s-mi355x``) mirrors the
reference's algorithm plugins over the same ABI.
"""

from . import engine  # noqa: F401

__all__ = ["engine"]
