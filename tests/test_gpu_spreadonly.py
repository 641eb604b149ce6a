"""Spread-only and interpolate-only plans (gpu_spreadinterponly = 1): the
plan spreads onto (or interpolates from) a grid of the mode counts, with no
FFT or deconvolution (src/cuda/execute.cu:105-204 with the plan's
spreadinterponly flag).

* The reference's known-answer test, test/cuda/cufinufft1dspreadinterponly_
  test.cu:68-152, run by test/cuda/CMakeLists.txt:231-236 (N1 = 1e3, M = 1e6;
  fp64 tol 1e-6 at sigma 2 checked to 1e-5, fp32 tol 1e-5 at sigma 1.25 to
  1e-4): the kernel's sum from one unit point at 0 (kersum); the mass of a
  spread of M random points, which must equal kersum times the sum of the
  strengths; the interpolation of a grid of ones, which must give kersum at
  every point.
* The same identities in 2-D and 3-D (the 3-D fp32 plan runs the column
  spreader on a spread-only grid geometry).
* A direct FP64 spread with the PSWF kernel at the plan's shape parameter
  (oracle.kernel.pswf) on small 2-D and 3-D grids, every grid value.
"""

import math

import numpy as np
import pytest

from oracle import kernel
from tests.helpers import rng, to_dev, to_host

pytestmark = pytest.mark.gpu

REAL = {"complex64": np.float32, "complex128": np.float64}


def _plan(ttype, shape, dtype, tol, sigma, **kw):
    from finufft_amd import Plan

    return Plan(ttype, shape, eps=tol, dtype=dtype, upsampfac=sigma, gpu_spreadinterponly=1,
                gpu_maxbatchsize=1, **kw)


def _kersum(torch, shape, dtype, tol, sigma):
    """Sum of the grid after spreading one unit strength at the origin."""
    p = _plan(1, shape, dtype, tol, sigma)
    x = [to_dev(torch, np.zeros(1, REAL[dtype])) for _ in shape]
    p.setpts(*x)
    fk = to_host(p.execute(to_dev(torch, np.ones(1, np.dtype(dtype)))))
    return fk.sum()


@pytest.mark.parametrize("dim,shape,M", [(1, (1000,), 1_000_000), (2, (200, 160), 400_000),
                                         (3, (48, 40, 44), 300_000)])
@pytest.mark.parametrize("dtype,tol,sigma,check", [("complex128", 1e-6, 2.0, 1e-5),
                                                   ("complex64", 1e-5, 1.25, 1e-4)])
def test_mass_and_interp_identities(torch_gpu, dim, shape, M, dtype, tol, sigma, check):
    torch = torch_gpu
    kersum = _kersum(torch, shape, dtype, tol, sigma)
    r = rng(1)
    real = REAL[dtype]
    x = [r.uniform(-np.pi, np.pi, M).astype(real) for _ in range(dim)]
    c = (r.uniform(-1, 1, M) + 1j * r.uniform(-1, 1, M)).astype(dtype)
    p1 = _plan(1, shape, dtype, tol, sigma)
    p1.setpts(*[to_dev(torch, a) for a in x])
    info = p1.info()
    assert info["nf"][:dim] == tuple(shape[::-1])  # the grid is the modes, x fastest
    fk = to_host(p1.execute(to_dev(torch, c)))
    mass = fk.astype(np.complex128).sum()
    mass_err = abs(mass - kersum * c.astype(np.complex128).sum()) / abs(mass)
    p2 = _plan(2, shape, dtype, tol, sigma)
    p2.setpts(*[to_dev(torch, a) for a in x])
    cout = to_host(p2.execute(to_dev(torch, np.ones(shape, np.dtype(dtype)))))
    sup_err = np.abs(cout - kersum).max() / abs(kersum)
    print(dim, dtype, info, f"kersum {kersum:.6g} mass err {mass_err:.2e} sup err {sup_err:.2e}")
    assert mass_err < check and sup_err < check


@pytest.mark.parametrize("shape", [(40, 36), (20, 24, 28)])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_spread_vs_direct(torch_gpu, shape, dtype):
    """Every value of a spread-only grid against a direct FP64 spread with
    the PSWF kernel at the plan's width and shape parameter."""
    torch = torch_gpu
    dim = len(shape)
    tol = 1e-5 if dtype == "complex64" else 1e-9
    r = rng(2)
    M = 600
    real = REAL[dtype]
    x = [r.uniform(-np.pi, np.pi, M).astype(real) for _ in range(dim)]
    c = (r.standard_normal(M) + 1j * r.standard_normal(M)).astype(dtype)
    p = _plan(1, shape, dtype, tol, 2.0)
    p.setpts(*[to_dev(torch, a) for a in x])
    info = p.info()
    ns, beta = info["ns"], info["beta"]
    got = to_host(p.execute(to_dev(torch, c))).astype(np.complex128)
    ref = np.zeros(shape, np.complex128)
    # array axis d <-> coordinate d; grid node i of n sits at -pi + 2 pi i / n
    idx, w = [], []
    for d in range(dim):
        n = shape[d]
        xr = ((x[d].astype(np.float64) + np.pi) % (2 * np.pi)) / (2 * np.pi) * n
        s = np.ceil(xr - ns / 2).astype(int)
        offs = s[:, None] + np.arange(ns)[None, :]
        w.append(kernel.pswf(beta, 2 * (offs - xr[:, None]) / ns))
        idx.append(offs % n)
    for j in range(M):
        blk = c[j]
        for d in range(dim):
            blk = np.multiply.outer(blk, w[d][j]) if d else blk * w[d][j]
        ref[np.ix_(*[idx[d][j] for d in range(dim)])] += blk
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(shape, dtype, info, f"relerr {err:.2e}")
    assert err < (2e-6 if dtype == "complex64" else 1e-8)


def test_spread_only_type3_inner_grid_unchanged(torch_gpu):
    """A spread-only plan of the 3-D column geometry spreads the same grid
    with one vector and with several."""
    torch = torch_gpu
    shape = (48, 40, 44)
    r = rng(3)
    M = 20_000
    x = [r.uniform(-np.pi, np.pi, M).astype(np.float32) for _ in range(3)]
    c = (r.standard_normal((3, M)) + 1j * r.standard_normal((3, M))).astype(np.complex64)
    from finufft_amd import Plan

    many = Plan(1, shape, n_trans=3, eps=1e-6, dtype="complex64", gpu_spreadinterponly=1)
    many.setpts(*[to_dev(torch, a) for a in x])
    got = to_host(many.execute(to_dev(torch, c)))
    one = _plan(1, shape, "complex64", 1e-6, 2.0)
    one.setpts(*[to_dev(torch, a) for a in x])
    for t in range(3):
        ref = to_host(one.execute(to_dev(torch, c[t])))
        assert np.abs(got[t] - ref).max() <= 1e-5 * np.abs(ref).max()
    assert math.isfinite(float(np.abs(got).sum()))
