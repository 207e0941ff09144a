"""MI355X-native Metropolis / Wolff sampling hot path of the header-only C++23
lattice Monte Carlo library (Jumqyc/Universal_Classical_MonteCarlo).

The product is the HIP library libucmc.so behind the C ABI in include/ucmc.h
and the C++23 host layer in include/ucmc/gpu.hpp.  This package gives Python
callers the ctypes binding (``abi``) and host-side statistics (``stats``).
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
