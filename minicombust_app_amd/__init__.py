"""MI355X-native flow-solver SIMPLE step of miniCOMBUST.

The product is the HIP library ``_lib/libmcflow.so`` (C-ABI in
``include/mcflow.h``) and the C++ drop-in ``host/FlowSolver.hpp``; this
package holds their sources, the build script and the ctypes binding.
"""
from . import mcflow  # noqa: F401

__all__ = ["mcflow"]
