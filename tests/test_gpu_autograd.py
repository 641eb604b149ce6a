"""Gradients of the differentiable NUFFTs (finufft_amd.autograd).

Checked against the FP64 direct-sum gradient (closed form: the adjoint for the
data gradient, sum_k f_k (i s k_d) e^{i s k.x} for the coordinate gradient)
and against central finite differences of the GPU transform itself at step
1e-4 on 64 random entries (complex128 plans, where the transform error is far
below the difference error). Loss: 0.5 ||T(data, x) - y||^2 with y complex
normal, as the differentiation workload (3-D type 2 on radial spokes plus
jitter) prescribes.
"""

import numpy as np
import pytest

from oracle import nudft
from tests.helpers import cnormal, rng, to_dev, to_host

pytestmark = pytest.mark.gpu

TOL = {"complex64": 1e-5, "complex128": 1e-11}
REAL = {"complex64": np.float32, "complex128": np.float64}


def _spokes(r, M, dim, dtype):
    """Radial spokes through the origin plus uniform jitter of 1e-3 (3-D) or
    uniform points (other dims)."""
    if dim != 3:
        return [r.uniform(-np.pi, np.pi, M).astype(dtype) for _ in range(dim)]
    nsp = 20
    u = r.standard_normal((nsp, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    rad = np.linspace(0, np.pi, M // nsp, endpoint=False)
    p = (u[:, None, :] * rad[None, :, None]).reshape(-1, 3)
    p = p + r.uniform(-1e-3, 1e-3, p.shape)
    return [p[:, d].astype(dtype) for d in range(3)]


def _grads_type2(torch, dim, dtype, shape, M, seed, ntr=1):
    from finufft_amd.autograd import NufftType2

    r = rng(seed)
    pts = _spokes(r, M, dim, REAL[dtype])
    M = len(pts[0])
    fshape = ((ntr,) if ntr > 1 else ()) + shape
    f = cnormal(r, fshape, np.dtype(dtype))
    y = cnormal(r, ((ntr,) if ntr > 1 else ()) + (M,), np.dtype(dtype))
    xt = [to_dev(torch, p).requires_grad_(True) for p in pts]
    ft = to_dev(torch, f).requires_grad_(True)
    op = NufftType2(shape, eps=TOL[dtype], n_trans=ntr, dtype=dtype)
    c = op(ft, *xt)
    loss = 0.5 * (c - to_dev(torch, y)).abs().pow(2).sum()
    loss.backward()
    return pts, f, y, to_host(c.detach()), to_host(ft.grad), [to_host(x.grad) for x in xt], op


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_type2_grads_vs_direct(torch_gpu, dim, dtype):
    shape = {1: (40,), 2: (20, 24), 3: (12, 14, 10)}[dim]
    pts, f, y, c, gf, gx, op = _grads_type2(torch_gpu, dim, dtype, shape, 1200, 40 + dim)
    p64 = [p.astype(np.float64) for p in pts]
    s = op.plan._isign
    c_ref = nudft.type2(p64, f, s)
    res = c_ref - y
    gf_ref = nudft.type1(p64, res, -s, shape)
    dc = nudft.type2_grad(p64, f, s)
    gx_ref = [np.real(np.conj(res) * dc[d]) for d in range(dim)]
    tol = TOL[dtype]
    e_f = nudft.relerr(gf, gf_ref)
    e_x = [nudft.relerr(gx[d], gx_ref[d]) for d in range(dim)]
    print(dim, dtype, e_f, e_x)
    assert e_f < 25 * tol
    assert max(e_x) < 100 * tol


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_type1_grads_vs_direct(torch_gpu, dtype):
    from finufft_amd.autograd import NufftType1

    torch = torch_gpu
    r = rng(50)
    shape = (12, 10, 14)
    M = 1000
    pts = [r.uniform(-np.pi, np.pi, M).astype(REAL[dtype]) for _ in range(3)]
    c = cnormal(r, (M,), np.dtype(dtype))
    y = cnormal(r, shape, np.dtype(dtype))
    xt = [to_dev(torch, p).requires_grad_(True) for p in pts]
    ct = to_dev(torch, c).requires_grad_(True)
    op = NufftType1(shape, eps=TOL[dtype], dtype=dtype)
    f = op(ct, *xt)
    loss = 0.5 * (f - to_dev(torch, y)).abs().pow(2).sum()
    loss.backward()
    p64 = [p.astype(np.float64) for p in pts]
    s = op.plan._isign
    res = nudft.type1(p64, c, s, shape) - y
    gc_ref = nudft.type2(p64, res, -s)
    # d f_k / d x_j = c_j (i s k) e^{i s k x_j}: grad_x = Re(c_j conj(d/dx_j A(res)))
    dA = nudft.type2_grad(p64, res, -s)
    gx_ref = [np.real(c * np.conj(dA[d])) for d in range(3)]
    tol = TOL[dtype]
    e_c = nudft.relerr(to_host(ct.grad), gc_ref)
    e_x = [nudft.relerr(to_host(xt[d].grad), gx_ref[d]) for d in range(3)]
    print(dtype, e_c, e_x)
    assert e_c < 25 * tol
    assert max(e_x) < 100 * tol


def test_grads_batched(torch_gpu):
    """n_trans vectors sharing the points: the coordinate gradient sums over
    the vectors."""
    pts, f, y, c, gf, gx, op = _grads_type2(torch_gpu, 3, "complex128", (10, 8, 12), 800, 60,
                                             ntr=3)
    p64 = [p.astype(np.float64) for p in pts]
    s = op.plan._isign
    res = nudft.type2(p64, f, s) - y
    dc = nudft.type2_grad(p64, f, s)
    for d in range(3):
        ref = np.real(np.conj(res) * dc[d]).sum(axis=0)
        assert nudft.relerr(gx[d], ref) < 1e-9
    assert nudft.relerr(gf, nudft.type1(p64, res, -s, (10, 8, 12))) < 1e-9


def test_finite_differences(torch_gpu):
    """Central differences (h = 1e-4) of the loss through the GPU transform at
    64 random entries of f (real and imaginary parts) and of x, y, z."""
    import torch as _t

    from finufft_amd.autograd import NufftType2

    torch = torch_gpu
    r = rng(70)
    shape = (16, 16, 16)
    pts = _spokes(r, 2000, 3, np.float64)
    M = len(pts[0])
    f = cnormal(r, shape, np.dtype("complex128"))
    y = cnormal(r, (M,), np.dtype("complex128"))
    op = NufftType2(shape, eps=1e-12, dtype="complex128")
    yt = to_dev(torch, y)

    def loss_of(ft, xs):
        with _t.no_grad():
            return float(0.5 * (op(ft, *xs) - yt).abs().pow(2).sum())

    xt = [to_dev(torch, p).requires_grad_(True) for p in pts]
    ft = to_dev(torch, f).requires_grad_(True)
    loss = 0.5 * (op(ft, *xt) - yt).abs().pow(2).sum()
    loss.backward()
    gf = to_host(ft.grad).reshape(-1)
    gx = [to_host(x.grad) for x in xt]
    h = 1e-4
    idx_f = r.choice(f.size, 32, replace=False)
    errs = []
    for i in idx_f:
        for part, unit in ((0, 1.0), (1, 1j)):
            fp = f.reshape(-1).copy()
            fm = f.reshape(-1).copy()
            fp[i] += h * unit
            fm[i] -= h * unit
            fd = (loss_of(to_dev(torch, fp.reshape(shape)), [x.detach() for x in xt])
                  - loss_of(to_dev(torch, fm.reshape(shape)), [x.detach() for x in xt])) / (2 * h)
            ad = gf[i].real if part == 0 else gf[i].imag
            errs.append(abs(fd - ad) / max(abs(ad), 1e-3))
    idx_x = r.choice(M, 32, replace=False)
    for j in idx_x:
        for d in range(3):
            xp = [x.detach().clone() for x in xt]
            xm = [x.detach().clone() for x in xt]
            xp[d][j] += h
            xm[d][j] -= h
            fd = (loss_of(ft.detach(), xp) - loss_of(ft.detach(), xm)) / (2 * h)
            ad = gx[d][j]
            errs.append(abs(fd - ad) / max(abs(ad), 1e-3))
    print("max relative FD error", max(errs))
    assert max(errs) < 1e-4
