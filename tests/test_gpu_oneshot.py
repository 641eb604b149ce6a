"""The 36 simple (one-shot) entry points cufinufft{,f}{1,2,3}d{1,2,3}[many]
(reference declarations include/cufinufft.h:45-186, implementation
src/cuda/c_interface.cpp:177-470) called through the C ABI, each against the
FP64 direct sum (the oracle) with the reference's pass criteria
(python/cufinufft/tests/utils.py:134,149,168: 25 tol for types 1/2, 100 tol
for type 3); and the error codes of the reference's
test/cuda/cufinufft_error_handling.cu:20-99 on the guru interface."""

import ctypes
import pathlib
import re

import numpy as np
import pytest

from oracle import nudft
from tests.helpers import cnormal, rng, to_dev, to_host

pytestmark = pytest.mark.gpu

SHAPES = {1: (40,), 2: (14, 18), 3: (8, 10, 12)}
TOL = {"f": 1e-5, "": 1e-11}
REAL = {"f": np.float32, "": np.float64}
CPLX = {"f": np.complex64, "": np.complex128}


def _fn(prec, dim, ttype, many):
    from finufft_amd import _lib

    lib = _lib.lib()
    f = getattr(lib, f"cufinufft{prec}{dim}d{ttype}{'many' if many else ''}")
    cr = ctypes.c_float if prec == "f" else ctypes.c_double
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    args = ([ctypes.c_int] if many else []) + [i64] + [vp] * dim + [vp, ctypes.c_int, cr]
    args += [i64] * (dim if ttype != 3 else 1)
    args += ([vp] * dim if ttype == 3 else []) + [vp, vp]
    f.argtypes = args
    f.restype = ctypes.c_int
    return f


@pytest.mark.parametrize("many", [False, True])
@pytest.mark.parametrize("ttype", [1, 2, 3])
@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("prec", ["f", ""])
def test_oneshot_vs_direct(torch_gpu, prec, dim, ttype, many):
    torch = torch_gpu
    r = rng(31 + dim + 3 * ttype)
    ntr = 3 if many else 1
    M = 700
    tol = TOL[prec]
    real, cplx = REAL[prec], CPLX[prec]
    shape = SHAPES[dim]
    pts = [r.uniform(-np.pi, np.pi, M).astype(real) for _ in range(dim)]
    p64 = [p.astype(np.float64) for p in pts]
    # C order (x fastest): the C interface's x is the last array axis
    dx = [to_dev(torch, p) for p in reversed(pts)]
    iflag = -1 if ttype == 2 else 1
    f = _fn(prec, dim, ttype, many)
    head = [ntr] if many else []
    if ttype == 1:
        c = cnormal(r, (ntr, M), cplx)
        dc = to_dev(torch, c)
        out = torch.zeros((ntr, *shape), dtype=dc.dtype, device="cuda")
        ier = f(*head, M, *[p.data_ptr() for p in dx], dc.data_ptr(), iflag, tol,
                *reversed(shape), out.data_ptr(), None)
        assert ier == 0
        got = to_host(out)
        refs = [nudft.type1(p64, c[t], iflag, shape) for t in range(ntr)]
        crit = 25 * tol
    elif ttype == 2:
        fk = cnormal(r, (ntr, *shape), cplx)
        dfk = to_dev(torch, fk)
        out = torch.zeros((ntr, M), dtype=dfk.dtype, device="cuda")
        ier = f(*head, M, *[p.data_ptr() for p in dx], out.data_ptr(), iflag, tol,
                *reversed(shape), dfk.data_ptr(), None)
        assert ier == 0
        got = to_host(out)
        refs = [nudft.type2(p64, fk[t], iflag) for t in range(ntr)]
        crit = 25 * tol
    else:
        K = 300
        trg = [r.uniform(-12, 12, K).astype(real) for _ in range(dim)]
        c = cnormal(r, (ntr, M), cplx)
        dc = to_dev(torch, c)
        # the device arrays are held for the call: a temporary's block could be
        # handed to the next allocation while its pointer is still in use
        dt = [to_dev(torch, t_) for t_ in reversed(trg)]
        out = torch.zeros((ntr, K), dtype=dc.dtype, device="cuda")
        ier = f(*head, M, *[p.data_ptr() for p in dx], dc.data_ptr(), iflag, tol, K,
                *[t_.data_ptr() for t_ in dt], out.data_ptr(), None)
        assert ier == 0
        got = to_host(out)
        refs = [nudft.type3(p64, c[t], iflag, [t_.astype(np.float64) for t_ in trg])
                for t in range(ntr)]
        crit = 100 * tol
    for t in range(ntr):
        err = nudft.relerr(got[t].reshape(refs[t].shape), refs[t])
        print(f"cufinufft{prec}{dim}d{ttype}{'many' if many else ''} vector {t}: {err:.2e}")
        assert err < crit


# ------------------------------------------------ error codes (guru interface)


def _codes():
    """The error enum of include/finufft_errors.h, by name."""
    text = (pathlib.Path(__file__).resolve().parents[1] / "include" / "finufft_errors.h").read_text()
    return {k: int(v) for k, v in re.findall(r"(FINUFFT_\w+)\s*=\s*(\d+)", text)}


def _makeplanf(ttype, dim, nmodes, ntr=1, opts=None):
    from finufft_amd import _lib

    lib = _lib.lib()
    n = (ctypes.c_int64 * 3)(*nmodes)
    plan = ctypes.c_void_p()
    ier = lib.cufinufftf_makeplan(ttype, dim, n, 1, ntr, 1e-5, ctypes.byref(plan),
                                  ctypes.byref(opts) if opts is not None else None)
    return ier, plan


def test_error_codes_of_the_reference(torch_gpu):
    """test/cuda/cufinufft_error_handling.cu:20-99, case by case."""
    from finufft_amd import _lib

    lib = _lib.lib()
    E = _codes()
    N = (16, 16, 16)
    assert _makeplanf(1, 0, N)[0] == E["FINUFFT_ERR_DIM_NOTVALID"]
    assert _makeplanf(4, 1, N)[0] == E["FINUFFT_ERR_TYPE_NOTVALID"]
    assert _makeplanf(1, 1, N, ntr=0)[0] == E["FINUFFT_ERR_NTRANS_NOTVALID"]
    assert _makeplanf(1, 1, (0, 1, 1))[0] == E["FINUFFT_ERR_NDATA_NOTVALID"]
    o = _lib.default_opts()
    o.upsampfac = 0.9
    o.gpu_kerevalmeth = 0
    assert _makeplanf(1, 2, N, opts=o)[0] == E["FINUFFT_ERR_UPSAMPFAC_TOO_SMALL"]
    o = _lib.default_opts()
    o.upsampfac = 0.9
    o.gpu_kerevalmeth = 1
    assert _makeplanf(1, 2, N, opts=o)[0] == E["FINUFFT_ERR_HORNER_WRONG_BETA"]
    # huge bins for the LDS method in 3-D
    o = _lib.default_opts()
    o.gpu_method = 2
    o.gpu_binsizex = o.gpu_binsizey = o.gpu_binsizez = 1024
    ier, plan = _makeplanf(1, 3, N, opts=o)
    assert ier == E["FINUFFT_ERR_INSUFFICIENT_SHMEM"]
    assert lib.cufinufft_destroy(None) == E["FINUFFT_ERR_PLAN_NOTVALID"]
    # type 3 with null targets
    ier, plan = _makeplanf(3, 1, N)
    assert ier == 0
    assert (lib.cufinufftf_setpts(plan, 1, None, None, None, 1, None, None, None)
            == E["FINUFFT_ERR_INVALID_ARGUMENT"])
    assert lib.cufinufftf_destroy(plan) == 0
    # type 3 with a negative number of targets
    ier, plan = _makeplanf(3, 1, N)
    assert ier == 0
    assert (lib.cufinufftf_setpts(plan, 1, None, None, None, -1, None, None, None)
            == E["FINUFFT_ERR_NUM_NU_PTS_INVALID"])
    assert lib.cufinufftf_destroy(plan) == 0
