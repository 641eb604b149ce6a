"""GPU tests of the pruned 3-D line transforms between the fine grid and the
modes (finufft_amd/csrc/fft3.hip), the default for 3-D plans whose fine-grid
sides are powers of two.

* against the FP64 direct sum (the oracle) at the reference's pass criterion
  (25 tol, python/cufinufft/tests/utils.py:134,149), both types, both
  precisions, both mode orders, several vectors in batches;
* against the same plan run with rocFFT and separate deconvolution /
  amplification (finufft_amd_set_fft_path(1)) at the headline grid
  (256^3 modes, nf = 512^3) and at every line length the kernels take from
  16 to 1024, where the two paths must agree to rounding.
The kernels' index arithmetic is checked on the CPU under AddressSanitizer by
tests/test_emu_fft3.py.
"""

import numpy as np
import pytest

from oracle import nudft
from tests.helpers import REAL, cnormal, rng, to_dev, to_host, uniform_pts

pytestmark = pytest.mark.gpu

TOL = {"complex64": 1e-5, "complex128": 1e-11}


def _set_path(path):
    from finufft_amd import _lib

    assert _lib.lib().finufft_amd_set_fft_path(path) == 0


def _plan(ttype, shape, dtype, tol, ntr=1, **kw):
    from finufft_amd import Plan

    return Plan(ttype, shape, n_trans=ntr, eps=tol, dtype=dtype, **kw)


def _exec(torch, plan, ttype, data):
    return to_host(plan.execute(to_dev(torch, data)))


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("ttype", [1, 2])
@pytest.mark.parametrize("modeord", [0, 1])
def test_pruned_vs_direct(torch_gpu, dtype, ttype, modeord):
    torch = torch_gpu
    shape = (16, 32, 16)  # array order (z, y, x): nf (32, 64, 32)
    tol = TOL[dtype]
    r = rng(31 + ttype + 2 * modeord)
    M, ntr = 3000, 3
    pts = uniform_pts(r, 3, M, REAL[dtype])
    plan = _plan(ttype, shape, dtype, tol, ntr=ntr, modeord=modeord, gpu_maxbatchsize=2)
    assert plan.info()["fft_path"] == 1
    plan.setpts(*[to_dev(torch, p) for p in pts])
    p64 = [p.astype(np.float64) for p in pts]
    if ttype == 1:
        c = cnormal(r, (ntr, M), np.dtype(dtype))
        got = _exec(torch, plan, 1, c)
        ref = np.stack([nudft.type1(p64, c[t], plan._isign, shape, modeord=modeord)
                        for t in range(ntr)])
    else:
        f = cnormal(r, (ntr,) + shape, np.dtype(dtype))
        got = _exec(torch, plan, 2, f)
        ref = np.stack([nudft.type2(p64, f[t], plan._isign, modeord=modeord)
                        for t in range(ntr)])
    err = nudft.relerr(got, ref)
    print(f"type {ttype} {dtype} modeord {modeord}: relerr {err:.2e}")
    assert err < 25 * tol


def _both_paths(torch, ttype, shape, dtype, tol, M, seed, **kw):
    """Outputs of the pruned path and of rocFFT + separate (de)convolution on
    the same points and data."""
    r = rng(seed)
    pts = uniform_pts(r, 3, M, REAL[dtype])
    data = cnormal(r, (M,) if ttype == 1 else shape, np.dtype(dtype))
    out = []
    try:
        for path in (0, 1):
            _set_path(path)
            plan = _plan(ttype, shape, dtype, tol, **kw)
            assert plan.info()["fft_path"] == (1 - path)
            plan.setpts(*[to_dev(torch, p) for p in pts])
            out.append(_exec(torch, plan, ttype, data))
            del plan
    finally:
        _set_path(0)
    return out


@pytest.mark.parametrize("ttype", [1, 2])
def test_pruned_vs_rocfft_headline_grid(torch_gpu, ttype):
    """256^3 modes on a 512^3 grid (the 3-D workload's transform), fp32."""
    a, b = _both_paths(torch_gpu, ttype, (256, 256, 256), "complex64", 1e-6, 200_000, 5)
    err = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"type {ttype}: pruned vs rocFFT relative difference {err:.2e}")
    assert err < 2e-6


LINE_CASES = [
    # array order (z, y, x); nf follows as next235(2 m), a power of two here
    ("complex64", (8, 64, 512)),     # nf (1024, 128, 16)
    ("complex64", (512, 8, 16)),     # nf (32, 16, 1024)
    ("complex64", (127, 31, 255)),   # nf (512, 64, 256): odd mode counts
    ("complex128", (16, 64, 512)),   # nf (1024, 128, 32)
    ("complex128", (512, 16, 16)),   # nf (32, 32, 1024)
    ("complex128", (127, 31, 255)),
]


@pytest.mark.parametrize("dtype,shape", LINE_CASES, ids=lambda c: str(c))
def test_pruned_vs_rocfft_line_lengths(torch_gpu, dtype, shape):
    """Every line length 16..1024 on some axis, odd and even mode counts."""
    tol = TOL[dtype]
    for ttype in (1, 2):
        a, b = _both_paths(torch_gpu, ttype, shape, dtype, tol, 5000, 7 + ttype)
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"{shape} {dtype} type {ttype}: relative difference {err:.2e}")
        assert err < (2e-6 if dtype == "complex64" else 1e-13)


def test_non_power_of_two_grid_keeps_rocfft(torch_gpu):
    plan = _plan(1, (10, 13, 16), "complex64", 1e-5)
    assert plan.info()["fft_path"] == 0
