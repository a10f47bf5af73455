"""platepcg: host side of the MI355X plate-bending PCG library.

The compute runs in the HIP library ``lib/libplatepcg.so`` (gfx950); this
package binds its C ABI and builds problem inputs.
"""
from . import _lib, problem
from ._lib import PlatePCGError, load
from .solver import Batch, Context, Net, Plan, lognormal_thickness, solve_gpu, stats_dict

__all__ = ["Batch", "Context", "Net", "Plan", "lognormal_thickness", "solve_gpu", "stats_dict", "PlatePCGError", "load", "problem",
           "_lib"]
