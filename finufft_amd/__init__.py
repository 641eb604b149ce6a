"""MI355X-native nonuniform FFT (types 1, 2, 3 in 1-3 dimensions).

The product is the C-ABI library ``libcufinufft.so`` (include/cufinufft.h,
a drop-in for the cuFINUFFT guru and simple interfaces). This package is its
Python face: ``Plan`` mirrors ``cufinufft.Plan``.
"""

from ._plan import FinufftError, Plan  # noqa: F401

__all__ = ["Plan", "FinufftError"]
