"""Output-driven spreading (gpu_method 3) against the FP64 direct sum (the
oracle): types 1 and 3 in 1-3 dimensions and both precisions, user np and bin
sizes, clustered points cut into many subproblems, several vectors, and the
adjoint of a method-3 plan. The reference runs the same matrix for method 3
through test/cuda/CMakeLists.txt:26-250 (the method argument of its
cufinufft*_test drivers) with the pass criteria of
python/cufinufft/tests/utils.py:121-168. The kernel (spread_od_kernel) is
also run from its source on the CPU (tests/test_emu_spreadinterp.py)."""

import numpy as np
import pytest

from oracle import nudft
from tests.helpers import REAL, cnormal, rng, to_dev, to_host, uniform_pts

pytestmark = pytest.mark.gpu

SHAPES = {1: (97,), 2: (24, 31), 3: (10, 13, 16)}
TOL = {"complex64": 1e-5, "complex128": 1e-11}


def _type1(torch, shape, dtype, pts, ntr=1, seed=0, **kw):
    from finufft_amd import Plan

    r = rng(seed)
    M = len(pts[0])
    tol = TOL[dtype]
    plan = Plan(1, shape, n_trans=ntr, eps=tol, dtype=dtype, gpu_method=3, **kw)
    plan.setpts(*[to_dev(torch, p) for p in pts])
    assert plan.info()["gpu_method"] == 3
    c = cnormal(r, (ntr, M) if ntr > 1 else (M,), np.dtype(dtype))
    f = to_host(plan.execute(to_dev(torch, c)))
    ref = nudft.type1([p.astype(np.float64) for p in pts], c, plan._isign, shape)
    return nudft.relerr(f, ref), tol


def _clustered(r, dim, M, real):
    """Half the points in one small corner cube (dense bins, many
    subproblems per bin), half uniform."""
    pts = []
    for _ in range(dim):
        a = r.uniform(-np.pi, np.pi, M)
        a[: M // 2] = r.uniform(-np.pi, -np.pi + 0.25, M // 2)
        pts.append(a.astype(real))
    return pts


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_type1_vs_direct(torch_gpu, dim, dtype):
    pts = uniform_pts(rng(0), dim, 1500, REAL[dtype])
    err, tol = _type1(torch_gpu, SHAPES[dim], dtype, pts)
    print(f"method 3 dim={dim} {dtype}: relerr={err:.3e}")
    assert err < 25 * tol


@pytest.mark.parametrize("shape,kw", [
    ((97,), dict(gpu_np=16, gpu_maxsubprobsize=50)),
    ((24, 31), dict(gpu_np=5, gpu_binsizex=5, gpu_binsizey=9)),
    ((10, 13, 16), dict(gpu_np=7, gpu_binsizex=3, gpu_binsizey=5, gpu_binsizez=4,
                        gpu_maxsubprobsize=40)),
    ((24, 31), dict(gpu_np=512)),
    ((40, 48, 36), dict()),
])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_options_and_clustered_points(torch_gpu, shape, kw, dtype):
    """User np (from a handful of points to a few hundred), user bins that
    neither divide the grid nor exceed the kernel width, small subproblems,
    and half the points packed into a few bins."""
    pts = _clustered(rng(9), len(shape), 4000, REAL[dtype])
    err, tol = _type1(torch_gpu, shape, dtype, pts, seed=9, **kw)
    print(shape, kw, dtype, f"relerr={err:.3e}")
    assert err < 25 * tol


@pytest.mark.parametrize("dim", [2, 3])
def test_several_vectors(torch_gpu, dim):
    """ntransf = 5 with a batch of 2: the vectors run in the launch's second
    grid dimension and in a last, partial batch."""
    pts = uniform_pts(rng(4), dim, 2000, np.float32)
    err, tol = _type1(torch_gpu, SHAPES[dim], "complex64", pts, ntr=5, seed=4,
                      gpu_maxbatchsize=2)
    assert err < 25 * tol


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_type3_vs_direct(torch_gpu, dim, dtype):
    from finufft_amd import Plan

    torch = torch_gpu
    r = rng(3)
    real = REAL[dtype]
    M, N = 1200, 900
    tol = TOL[dtype]
    x = [r.uniform(-2.0, 3.0, M).astype(real) for _ in range(dim)]
    s = [r.uniform(-40.0, 25.0, N).astype(real) for _ in range(dim)]
    c = cnormal(r, (M,), np.dtype(dtype))
    plan = Plan(3, dim, eps=tol, dtype=dtype, gpu_method=3)
    dx = [to_dev(torch, a) for a in x] + [None] * (3 - dim)
    ds = [to_dev(torch, a) for a in s] + [None] * (3 - dim)
    plan.setpts(*dx, s=ds[0], t=ds[1], u=ds[2])
    f = to_host(plan.execute(to_dev(torch, c)))
    ref = nudft.type3([a.astype(np.float64) for a in x], c, 1, [a.astype(np.float64) for a in s])
    err = nudft.relerr(f, ref)
    print(f"method 3 type3 dim={dim} {dtype}: relerr={err:.3e}")
    assert err < 100 * tol


@pytest.mark.parametrize("dtype,shape", [("complex64", (40, 48, 36)), ("complex128", (60, 50))])
def test_adjointness(torch_gpu, dtype, shape):
    """<type1(c), f> = <c, type1^H(f)> on a method-3 plan: its adjoint
    interpolates (LDS method) with the bins chosen for output-driven
    spreading."""
    from finufft_amd import Plan

    torch = torch_gpu
    r = rng(5)
    M = 3000
    pts = uniform_pts(r, len(shape), M, REAL[dtype])
    tol = 1e-12 if dtype == "complex128" else 1e-6
    plan = Plan(1, shape, eps=tol, dtype=dtype, gpu_method=3)
    plan.setpts(*[to_dev(torch, p) for p in pts])
    c = cnormal(r, (M,), np.dtype(dtype))
    f = cnormal(r, shape, np.dtype(dtype))
    F = to_host(plan.execute(to_dev(torch, c)))
    C = to_host(plan.execute_adjoint(to_dev(torch, f)))
    lhs = np.vdot(f.ravel().astype(np.complex128), F.ravel().astype(np.complex128))
    rhs = np.vdot(C.astype(np.complex128), c.astype(np.complex128))
    denom = np.linalg.norm(f) * np.linalg.norm(F) / np.sqrt(np.prod(shape))
    assert abs(lhs - rhs) / denom < (1e-10 if dtype == "complex128" else 2e-5)


def test_matches_method2(torch_gpu):
    """Methods 2 and 3 spread the same points to the same grid (spread-only
    plans, so the comparison is of the spreading alone); both sum in fp64
    tiles, in different orders."""
    from finufft_amd import Plan

    torch = torch_gpu
    r = rng(12)
    shape = (48, 40, 44)
    M = 20000
    pts = _clustered(r, 3, M, np.float32)
    c = cnormal(r, (M,), np.complex64)
    outs = []
    for method in (2, 3):
        plan = Plan(1, shape, eps=1e-6, dtype="complex64", gpu_method=method,
                    gpu_spreadinterponly=1)
        plan.setpts(*[to_dev(torch, p) for p in pts])
        outs.append(to_host(plan.execute(to_dev(torch, c))).astype(np.complex128))
    rel = np.linalg.norm(outs[0] - outs[1]) / np.linalg.norm(outs[0])
    print("method 2 vs 3 spread-only relative difference", rel)
    assert rel < 1e-5


def test_np_too_large(torch_gpu):
    """A batch of np points whose staging exceeds the LDS is refused at plan
    time, as the reference's validate_fit (heuristics.cu:343-351)."""
    from finufft_amd import FinufftError, Plan

    with pytest.raises(FinufftError) as e:
        Plan(1, (16, 16, 16), eps=1e-11, dtype="complex128", gpu_method=3, gpu_np=2048)
    assert e.value.code == 19  # FINUFFT_ERR_INSUFFICIENT_SHMEM
