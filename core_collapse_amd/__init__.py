"""core_collapse_amd -- MI355X (gfx950) solver step for the core_collapse collapse codes.

Host-side mirror of the reference programs' step interface (Params defaults, function names and
argument meaning) over the C ABI of libccamd.so (include/ccamd.h).  See mhd2d.Mhd2D.
"""
from . import _lib  # noqa: F401
from ._lib import CC_FP64, CC_MIXED, CCError, device_info  # noqa: F401
