"""MI355X-native ensemble integration path of xSPDE (peterddrummond/xspde_matlab).

Hand-written HIP kernels for gfx950 behind the C ABI in include/xspde_hip.h;
this package holds the kernel sources (csrc/), their build recipe (build.py),
the ctypes binding (_lib.py) and the host mirror of xsim/xensemble/xpath
(xsim.py).
"""
from .xsim import GpuFormError, Problem, preferences, make_desc, statistics, xsim  # noqa: F401

__all__ = ["xsim", "Problem", "preferences", "make_desc", "statistics", "GpuFormError"]
