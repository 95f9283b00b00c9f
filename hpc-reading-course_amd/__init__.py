"""MI355X-native (gfx950) SBP-SAT hot path of the HPC-reading-course solver.

Python host binding over the C-ABI in include/sbp_hip.h.  Importing the
package does not touch the GPU; ``capi.load_library()`` loads libsbp_hip.so
(raising if it is not built) on first use.
"""
import importlib

from . import problems  # noqa: F401

__all__ = ["problems", "capi", "corrector"]


def __getattr__(name):
    if name in ("capi", "corrector"):
        return importlib.import_module(__name__ + "." + name)
    raise AttributeError(name)
