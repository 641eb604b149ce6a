"""ctypes binding of ``libcufinufft.so`` (the C ABI of include/cufinufft.h and
include/finufft_amd.h).

Mirrors the reference's ctypes layer (python/cufinufft/cufinufft/_cufinufft.py:
90-201): the same opts struct layout and the same function signatures, with
device pointers passed as ``c_void_p``.

torch is imported before the library is loaded: both then share the one HIP
runtime that torch brings into the process.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_int64, c_size_t, c_void_p
from pathlib import Path

try:  # noqa: SIM105 - import order matters, see the module docstring
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is part of the supported setup
    torch = None

LIB_PATH = Path(os.environ.get("FINUFFT_AMD_LIB", Path(__file__).with_name("libcufinufft.so")))


class NufftOpts(ctypes.Structure):
    """Byte-compatible mirror of ``cufinufft_opts`` (include/cufinufft_opts.h)."""

    _fields_ = [
        ("upsampfac", c_double),
        ("gpu_method", c_int),
        ("gpu_sort", c_int),
        ("gpu_binsizex", c_int),
        ("gpu_binsizey", c_int),
        ("gpu_binsizez", c_int),
        ("gpu_obinsizex", c_int),
        ("gpu_obinsizey", c_int),
        ("gpu_obinsizez", c_int),
        ("gpu_maxsubprobsize", c_int),
        ("gpu_kerevalmeth", c_int),
        ("gpu_spreadinterponly", c_int),
        ("gpu_maxbatchsize", c_int),
        ("gpu_device_id", c_int),
        ("gpu_stream", c_void_p),
        ("modeord", c_int),
        ("gpu_np", c_int),
        ("debug", c_int),
    ]


_lib = None


def _declare(lib):
    vp, i64p = c_void_p, POINTER(c_int64)
    optp = POINTER(NufftOpts)
    lib.cufinufft_default_opts.argtypes = [optp]
    lib.cufinufft_default_opts.restype = None
    for prefix, real in (("cufinufft", c_double), ("cufinufftf", c_float)):
        mk = getattr(lib, f"{prefix}_makeplan")
        mk.argtypes = [c_int, c_int, i64p, c_int, c_int, real, POINTER(c_void_p), optp]
        mk.restype = c_int
        sp = getattr(lib, f"{prefix}_setpts")
        sp.argtypes = [vp, c_int64, vp, vp, vp, c_int, vp, vp, vp]
        sp.restype = c_int
        for name in ("execute", "execute_adjoint"):
            f = getattr(lib, f"{prefix}_{name}")
            f.argtypes = [vp, vp, vp]
            f.restype = c_int
        g = getattr(lib, f"{prefix}_interp_grad")
        g.argtypes = [vp, vp, vp, vp]
        g.restype = c_int
        d = getattr(lib, f"{prefix}_destroy")
        d.argtypes = [vp]
        d.restype = c_int
    lib.finufft_amd_batch_makeplan.argtypes = [c_int, c_int, POINTER(c_void_p), c_size_t,
                                               POINTER(c_void_p)]
    lib.finufft_amd_batch_makeplan.restype = c_int
    lib.finufft_amd_batch_execute.argtypes = [vp, POINTER(c_void_p), POINTER(c_void_p)]
    lib.finufft_amd_batch_execute.restype = c_int
    lib.finufft_amd_batch_destroy.argtypes = [vp]
    lib.finufft_amd_batch_destroy.restype = c_int
    lib.finufft_amd_batch_info.argtypes = [vp, i64p]
    lib.finufft_amd_batch_info.restype = c_int
    lib.finufft_amd_batch_set_graph.argtypes = [vp, c_int]
    lib.finufft_amd_batch_set_graph.restype = c_int
    lib.finufft_amd_mlp_forward.argtypes = [vp, c_int, vp, vp, vp]
    lib.finufft_amd_mlp_forward.restype = c_int
    lib.finufft_amd_mlp_grad.argtypes = [vp, c_int, vp, vp, vp, vp, vp, vp]
    lib.finufft_amd_mlp_grad.restype = c_int
    lib.finufft_amd_mlp_adam.argtypes = [vp, vp, vp, vp, c_int, c_float, vp]
    lib.finufft_amd_mlp_adam.restype = c_int
    lib.finufft_amd_plan_info.argtypes = [vp, c_int, i64p, POINTER(c_double)]
    lib.finufft_amd_plan_info.restype = c_int
    lib.finufft_amd_copy_sort.argtypes = [vp, c_int, vp, vp]
    lib.finufft_amd_copy_sort.restype = c_int
    lib.finufft_amd_set_stage_timing.argtypes = [vp, c_int, c_int]
    lib.finufft_amd_set_stage_timing.restype = c_int
    lib.finufft_amd_get_stage_times.argtypes = [vp, c_int, POINTER(c_double)]
    lib.finufft_amd_get_stage_times.restype = c_int
    lib.finufft_amd_set_fft_path.argtypes = [c_int]
    lib.finufft_amd_set_fft_path.restype = c_int
    lib.finufft_amd_density_stats.argtypes = [c_int, c_int64, vp, vp, vp, c_int, c_int,
                                              POINTER(c_double), vp]
    lib.finufft_amd_density_stats.restype = c_int
    lib.finufft_amd_density_stats_noalloc.argtypes = [c_int, c_int64, vp, vp, vp, c_int, c_int,
                                                      vp, vp]
    lib.finufft_amd_density_stats_noalloc.restype = c_int
    lib.finufft_amd_lds_probe.argtypes = [c_int, c_int, c_int, POINTER(c_double),
                                          POINTER(c_double)]
    lib.finufft_amd_lds_probe.restype = c_int
    lib.finufft_amd_next235.argtypes = [c_int64, c_int64]
    lib.finufft_amd_next235.restype = c_int64
    lib.finufft_amd_gaussquad.argtypes = [c_int, POINTER(c_double), POINTER(c_double)]
    lib.finufft_amd_gaussquad.restype = None
    lib.finufft_amd_kernel_shape.argtypes = [
        c_double, c_int, c_double, c_int, c_int, c_int, POINTER(c_int), POINTER(c_double),
        POINTER(c_int)]
    lib.finufft_amd_kernel_shape.restype = c_int
    lib.finufft_amd_horner_coeffs.argtypes = [c_int, c_double, POINTER(c_double), c_int]
    lib.finufft_amd_horner_coeffs.restype = c_int
    lib.finufft_amd_horner_coeffs_pswf.argtypes = [c_int, c_double, POINTER(c_double), c_int]
    lib.finufft_amd_horner_coeffs_pswf.restype = c_int
    lib.finufft_amd_pswf.argtypes = [c_double, vp, c_int64, vp]
    lib.finufft_amd_pswf.restype = None
    lib.finufft_amd_fold_rescale_f.argtypes = [vp, c_int64, c_int, vp]
    lib.finufft_amd_fold_rescale_f.restype = None
    lib.finufft_amd_fold_rescale_d.argtypes = [vp, c_int64, c_int, vp]
    lib.finufft_amd_fold_rescale_d.restype = None
    lib.finufft_amd_fine_grid_len.argtypes = [c_int64, c_double, c_int]
    lib.finufft_amd_fine_grid_len.restype = c_int64
    return lib


def lib():
    """The loaded native library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m finufft_amd.build` "
                "(there is no CPU fallback)"
            )
        _lib = _declare(ctypes.CDLL(str(LIB_PATH)))
    return _lib


def default_opts() -> NufftOpts:
    o = NufftOpts()
    lib().cufinufft_default_opts(ctypes.byref(o))
    return o
