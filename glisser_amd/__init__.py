"""MI355X (gfx950) time-block Wisdom-Holman test-particle integrator.

The device path lives in ``libglisser_amd.so`` (hand-written HIP kernels
behind the C ABI of ``include/glisser_amd.h``); :mod:`glisser_amd.native`
binds it with ctypes.  Importing the package does not touch the GPU.
"""
from .native import (Context, GlxError, Planets, from_elements, load, load_h0,  # noqa: F401
                     to_elements)

__all__ = ["Context", "GlxError", "Planets", "from_elements", "load", "load_h0", "to_elements"]
