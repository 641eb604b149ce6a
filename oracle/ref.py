"""ctypes access to the reference's own kernel-selection code, compiled from
its sources where they lie (src/common/pswf.cpp, kernel.cpp, utils.cpp) into
oracle/_ref/libref_kernel.so by `make -C oracle ref` (oracle/Makefile; the
shim oracle/csrc/ref_kernel_shim.cpp only calls it).

Functions: the PSWF psi_0^c (finufft::common::PSWF0), the width law
(theoretical_kernel_ns, clamp_kernel_ns), the shape parameter
(set_kernel_shape_given_ns), the reference CPU library's per-interval
Horner fit (poly_fit), next235 and gaussquad.

Test infrastructure only (see oracle/__init__.py). `available()` is False
where the library was not built (no reference tree when it was made).
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_ref" / "libref_kernel.so"
REF_TREE = Path("/root/reference")
_lib = None


def build() -> bool:
    """Builds the library when the reference tree is present."""
    if not (REF_TREE / "src" / "common" / "pswf.cpp").exists():
        return LIB.exists()
    subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)
    return True


def available() -> bool:
    return LIB.exists()


def lib():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(str(LIB))
        vp, ci, cd, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
        _lib.ref_pswf.argtypes = [cd, vp, ci, vp]
        _lib.ref_pswf.restype = None
        _lib.ref_theoretical_ns.argtypes = [cd, ci, ci, cd]
        _lib.ref_theoretical_ns.restype = ci
        _lib.ref_clamp_ns.argtypes = [ci, cd, ci]
        _lib.ref_clamp_ns.restype = ci
        _lib.ref_beta.argtypes = [ci, cd, ci]
        _lib.ref_beta.restype = cd
        _lib.ref_pswf_panel_coeffs.argtypes = [cd, ci, ci, ci, vp]
        _lib.ref_pswf_panel_coeffs.restype = None
        _lib.ref_next235.argtypes = [i64, i64]
        _lib.ref_next235.restype = i64
        _lib.ref_gaussquad.argtypes = [ci, vp, vp]
        _lib.ref_gaussquad.restype = None
    return _lib


def pswf(c: float, z) -> np.ndarray:
    z = np.ascontiguousarray(z, np.float64)
    out = np.empty_like(z)
    lib().ref_pswf(float(c), z.ctypes.data, z.size, out.ctypes.data)
    return out


def kernel_ns(tol: float, dim: int, ttype: int, sigma: float, is_float: bool) -> int:
    """Width the reference CPU plan uses (theoretical then clamped)."""
    ns = lib().ref_theoretical_ns(float(tol), dim, ttype, float(sigma))
    return lib().ref_clamp_ns(ns, float(sigma), int(is_float))


def beta(ns: int, sigma: float, kerformula: int = 8) -> float:
    return lib().ref_beta(ns, float(sigma), kerformula)


def pswf_panel_coeffs(beta_: float, ns: int, j: int, n: int) -> np.ndarray:
    """Coefficients (highest degree first) of interval j's polynomial in the
    local variable x in [-1, 1]."""
    out = np.empty(n)
    lib().ref_pswf_panel_coeffs(float(beta_), ns, j, n, out.ctypes.data)
    return out


def next235(n: int, f: int) -> int:
    return lib().ref_next235(n, f)


def gaussquad(n: int):
    x = np.empty(n)
    w = np.empty(n)
    lib().ref_gaussquad(n, x.ctypes.data, w.ctypes.data)
    return x, w
