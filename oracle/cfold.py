"""ctypes access to the oracle's C restatement (oracle/_build/liboracle.so).

Test infrastructure only (see oracle/__init__.py).
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
_lib = None


def build() -> Path:
    """Runs the oracle Makefile (cheap when up to date)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        vp = ctypes.c_void_p
        for name in ("oracle_fold_rescale_f", "oracle_fold_rescale_d"):
            f = getattr(_lib, name)
            f.argtypes = [vp, ctypes.c_int64, ctypes.c_int, vp]
            f.restype = None
    return _lib


def fold_rescale(x: np.ndarray, n: int) -> np.ndarray:
    """Device fold of x into [0, n) with round-down subtract and multiply,
    in x's precision (float32 or float64)."""
    x = np.ascontiguousarray(x)
    out = np.empty_like(x)
    if x.dtype == np.float32:
        fn = lib().oracle_fold_rescale_f
    elif x.dtype == np.float64:
        fn = lib().oracle_fold_rescale_d
    else:
        raise TypeError("float32 or float64 expected")
    fn(x.ctypes.data, x.size, int(n), out.ctypes.data)
    return out
