"""xyst_amd -- MI355X-native edge-based flow path of Xyst.

Device code lives in libxg.so (hand-written HIP for gfx950, built from
xyst_amd/csrc); this package is its Python mirror over the C-ABI in
include/xg.h.  Importing it does not touch the GPU; creating a Context does.
"""
from . import meshgen  # noqa: F401
from ._xg import LIBPATH, XgError, lib  # noqa: F401

__all__ = ["meshgen", "lib", "LIBPATH", "XgError"]
