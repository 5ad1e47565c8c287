"""htx -- MI355X-native Hermite-TENO finite-volume residual / time-step path.

The numerics live in libhtx.so (HIP for gfx950, C ABI in include/htx.h); this
package is the Python-side mirror of the reference's solver interface.
"""
from ._lib import HtxError, load  # noqa: F401
from .solver import System2D  # noqa: F401
from .solver1d import System1D  # noqa: F401
from . import ic  # noqa: F401

__all__ = ["System2D", "System1D", "HtxError", "load", "ic"]
