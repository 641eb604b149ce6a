"""The C ABI on the CPU: every declared symbol is exported, the opts struct and
its defaults match the reference, and the argument checks that return before
any device call give the reference's error codes
(test/cuda/cufinufft_error_handling.cu:20-99, src/cuda/c_interface.cpp).
"""

import ctypes
import re
from ctypes import byref, c_int64, c_void_p
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        for m in re.finditer(r"FINUFFT_EXPORT\s+[\w\s\*]+?\b(\w+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_header_symbols_are_exported():
    from finufft_amd import _lib

    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = _declared_symbols()
    assert len(names) >= 45 + 4
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_simple_interface_count():
    """36 one-shot entry points (3 dims x 3 types x {single, many} x 2 precisions)."""
    names = _declared_symbols()
    simple = [n for n in names if re.fullmatch(r"cufinufftf?[123]d[123](many)?", n)]
    assert len(simple) == 36


def test_opts_layout_and_defaults():
    """Field order and defaults of src/cuda/c_interface.cpp:153-172."""
    from finufft_amd import _lib

    o = _lib.default_opts()
    assert ctypes.sizeof(o) == 8 + 13 * 4 + 4 + 8 + 4 * 3 + 4  # double, 13 ints, pad, ptr, 3 ints, pad
    expect = dict(upsampfac=0.0, gpu_method=0, gpu_sort=1, gpu_binsizex=0, gpu_binsizey=0,
                  gpu_binsizez=0, gpu_obinsizex=0, gpu_obinsizey=0, gpu_obinsizez=0,
                  gpu_maxsubprobsize=1024, gpu_kerevalmeth=1, gpu_spreadinterponly=0,
                  gpu_maxbatchsize=0, gpu_device_id=0, modeord=0, gpu_np=0, debug=0)
    for k, v in expect.items():
        assert getattr(o, k) == v, k
    assert not o.gpu_stream


def _makeplanf(typ, dim, modes, ntr=1, tol=1e-5, opts=None):
    from finufft_amd import _lib

    lib = _lib.lib()
    m = (c_int64 * 3)(*modes)
    h = c_void_p(0)
    return lib.cufinufftf_makeplan(typ, dim, m, 1, ntr, tol, byref(h),
                                   byref(opts) if opts is not None else None)


def test_error_codes_before_device():
    from finufft_amd import _lib

    N = (16, 16, 16)
    assert _makeplanf(1, 0, N) == 12  # FINUFFT_ERR_DIM_NOTVALID
    assert _makeplanf(1, 4, N) == 12
    assert _makeplanf(4, 1, N) == 10  # FINUFFT_ERR_TYPE_NOTVALID
    assert _makeplanf(1, 1, N, ntr=0) == 9  # FINUFFT_ERR_NTRANS_NOTVALID
    assert _makeplanf(1, 1, (0, 1, 1)) == 14  # FINUFFT_ERR_NDATA_NOTVALID
    assert _makeplanf(1, 2, (2 ** 31, 1, 1)) == 14
    assert _makeplanf(1, 3, (2 ** 11, 2 ** 11, 2 ** 10)) == 14  # product above INT32_MAX
    o = _lib.default_opts()
    o.upsampfac = 0.9
    o.gpu_kerevalmeth = 0
    assert _makeplanf(1, 2, N, opts=o) == 7  # FINUFFT_ERR_UPSAMPFAC_TOO_SMALL
    o.gpu_kerevalmeth = 1
    assert _makeplanf(1, 2, N, opts=o) == 8  # FINUFFT_ERR_HORNER_WRONG_BETA
    o = _lib.default_opts()
    o.gpu_method = 3
    assert _makeplanf(2, 2, N, opts=o) == 17  # FINUFFT_ERR_METHOD_NOTVALID: method 3 only spreads
    o.gpu_method = 5
    assert _makeplanf(1, 2, N, opts=o) == 19  # FINUFFT_ERR_INSUFFICIENT_SHMEM (heuristics.cu:440)
    lib = _lib.lib()
    assert lib.cufinufft_destroy(None) == 16  # FINUFFT_ERR_PLAN_NOTVALID
    assert lib.cufinufftf_destroy(None) == 16
    # the process-wide transform path: 0 or 1, anything else refused
    assert lib.finufft_amd_set_fft_path(2) == 21
    assert lib.finufft_amd_set_fft_path(-1) == 21
    assert lib.finufft_amd_set_fft_path(0) == 0
    assert lib.cufinufftf_execute(None, None, None) == 16
    assert lib.cufinufftf_setpts(None, 1, None, None, None, 0, None, None, None) == 16
    # a pointer that is not a live plan (never made, destroyed, or of the other
    # precision) is refused without being dereferenced
    stray = c_void_p(0x1000)
    assert lib.cufinufftf_execute(stray, None, None) == 16
    assert lib.cufinufft_execute(stray, None, None) == 16
    assert lib.cufinufftf_setpts(stray, 1, None, None, None, 0, None, None, None) == 16
    assert lib.cufinufftf_destroy(stray) == 16
    assert lib.cufinufft_destroy(stray) == 16
    iout = (c_int64 * 16)()
    assert lib.finufft_amd_plan_info(stray, 0, iout, None) == 16
    assert lib.finufft_amd_batch_execute(stray, None, None) == 16
    assert lib.finufft_amd_batch_destroy(stray) == 16


def test_simple_interface_argument_errors():
    from finufft_amd import _lib

    lib = _lib.lib()
    f = lib.cufinufftf1d1
    f.argtypes = [c_int64, c_void_p, c_void_p, ctypes.c_int, ctypes.c_float, c_int64, c_void_p,
                  c_void_p]
    f.restype = ctypes.c_int
    assert f(2 ** 31, None, None, 1, 1e-5, 10, None, None) == 14  # M > INT32_MAX
    assert f(10, None, None, 1, 1e-5, 0, None, None) == 14  # zero modes


@pytest.mark.parametrize("shape", [(40,), (12, 16), (10, 12, 8)])
def test_cpu_restatement_vs_direct(shape):
    """The CPU baseline restatement is itself accurate at tol 1e-6 (fp32)."""
    from oracle import cpu, nudft

    r = np.random.default_rng(8)
    M = 2000
    pts = [r.uniform(-np.pi, np.pi, M).astype(np.float32) for _ in shape]
    p64 = [p.astype(np.float64) for p in pts]
    c = (r.standard_normal(M) + 1j * r.standard_normal(M)).astype(np.complex64)
    f = (r.standard_normal(shape) + 1j * r.standard_normal(shape)).astype(np.complex64)
    p1 = cpu.CpuPlan(1, shape, 1e-6, nthreads=2)
    p1.setpts(*pts)
    assert nudft.relerr(p1.execute(c), nudft.type1(p64, c, 1, shape)) < 5e-6
    p2 = cpu.CpuPlan(2, shape, 1e-6, nthreads=2)
    p2.setpts(*pts)
    assert nudft.relerr(p2.execute(f), nudft.type2(p64, f, -1)) < 5e-6
