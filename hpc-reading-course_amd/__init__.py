"""MI355X-native (gfx950) SBP-SAT hot path of the HPC-reading-course solver.

Python host binding over the C-ABI in include/sbp_hip.h.  Importing the
package does not touch the GPU; ``capi.load_library()`` loads libsbp_hip.so
(raising if it is not built) on first use.
"""
from . import problems  # noqa: F401

__all__ = ["problems", "capi"]


def __getattr__(name):
    if name == "capi":
        from . import capi
        return capi
    raise AttributeError(name)
