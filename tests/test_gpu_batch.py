"""Fused execution of many uneven type-2 (and type-1) problems
(finufft_amd.batch).

Every problem of the batch must give what its own plan gives (same kernels;
the FFT runs batched over more vectors, so agreement is to rounding), and the
direct sum to the reference's 25 tol criterion. The recorded graph is
replayed on new data and compared with direct launches.
"""

import numpy as np
import pytest

from oracle import nudft
from tests.helpers import cnormal, rng, to_dev, to_host

pytestmark = pytest.mark.gpu


def _problems(torch, nprob, ntr, dtype, seed, dim=2):
    from finufft_amd import Plan

    r = rng(seed)
    real = np.float32 if dtype == "complex64" else np.float64
    tol = 1e-5 if dtype == "complex64" else 1e-11
    plans, pts, fks = [], [], []
    for _ in range(nprob):
        n = int(np.exp(r.uniform(np.log(8), np.log(48))))
        M = int(np.exp(r.uniform(np.log(100), np.log(3000))))
        shape = (n,) * dim
        x = [r.uniform(-np.pi, np.pi, M).astype(real) for _ in range(dim)]
        p = Plan(2, shape, n_trans=ntr, eps=tol, dtype=dtype)
        xd = [to_dev(torch, a) for a in x]
        p.setpts(*xd)
        plans.append(p)
        pts.append((x, xd))
        fks.append(cnormal(r, (ntr,) + shape if ntr > 1 else shape, np.dtype(dtype)))
    return plans, pts, fks, tol


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
@pytest.mark.parametrize("dim", [2, 3])
def test_batch_matches_plans_and_direct(torch_gpu, dtype, dim):
    from finufft_amd.batch import BatchPlan

    torch = torch_gpu
    ntr = 4
    plans, pts, fks, tol = _problems(torch, 12, ntr, dtype, 80 + dim, dim)
    fk_d = [to_dev(torch, f) for f in fks]
    outs = [torch.empty((ntr, len(x[0])), dtype=fk.dtype, device="cuda")
            for (x, _), fk in zip(pts, fk_d)]
    b = BatchPlan(plans, arena_bytes=1 << 26)  # small arena: several rounds
    info = b.info()
    assert info["nprob"] == 12 and info["vectors_per_round"] <= ntr
    b.execute(fk_d, outs)
    torch.cuda.synchronize()
    for p, (x, _), fk, fkh, o in zip(plans, pts, fk_d, fks, outs):
        ref_plan = to_host(p.execute(fk))
        got = to_host(o)
        assert nudft.relerr(got, ref_plan) < 50 * np.finfo(got.real.dtype).eps * 10
        ref = nudft.type2([a.astype(np.float64) for a in x], fkh, p._isign)
        assert nudft.relerr(got, ref) < 25 * tol
    print(info)


def test_batch_graph_replay(torch_gpu):
    """Replays with new inputs and outputs, graph against direct launches."""
    from finufft_amd.batch import BatchPlan

    torch = torch_gpu
    ntr = 2
    plans, pts, fks, tol = _problems(torch, 20, ntr, "complex64", 90)
    b = BatchPlan(plans)
    r = rng(91)
    results = []
    for rep in range(3):
        fk = [to_dev(torch, cnormal(r, f.shape, np.dtype("complex64"))) for f in fks]
        out_g = [torch.empty((ntr, len(x[0])), dtype=torch.complex64, device="cuda")
                 for x, _ in pts]
        out_e = [torch.empty_like(o) for o in out_g]
        b.set_graph(True)
        b.execute(fk, out_g)
        torch.cuda.synchronize()
        b.set_graph(False)
        b.execute(fk, out_e)
        torch.cuda.synchronize()
        for a, e in zip(out_g, out_e):
            assert torch.equal(a, e)
        results.append([to_host(o) for o in out_g])
    info = b.info()
    assert info["graph"] == 1, info
    # different inputs gave different outputs (the pointer table is refreshed)
    assert not np.allclose(results[0][0], results[1][0])


def test_batch_rejects_mixed_plans(torch_gpu):
    from finufft_amd import FinufftError, Plan
    from finufft_amd.batch import BatchPlan

    torch = torch_gpu
    r = rng(92)
    x = [to_dev(torch, r.uniform(-np.pi, np.pi, 100).astype(np.float32)) for _ in range(2)]
    p2 = Plan(2, (16, 16), eps=1e-5, dtype="complex64")
    p2.setpts(*x)
    p1 = Plan(1, (16, 16), eps=1e-5, dtype="complex64")
    p1.setpts(*x)
    p3 = Plan(2, (16, 16), eps=1e-5, dtype="complex64", n_trans=2)
    p3.setpts(*x)
    with pytest.raises(FinufftError):
        BatchPlan([p2, p1])
    with pytest.raises(FinufftError):
        BatchPlan([p2, p3])


def test_batch_chunks_with_remainder(torch_gpu):
    """Ten problems of one grid shape, a budget of three per chunk: chunks of
    3, 3, 3 and 1 reuse one grid arena in turn."""
    from finufft_amd import Plan
    from finufft_amd.batch import BatchPlan

    torch = torch_gpu
    r = rng(93)
    ntr, shape = 2, (16, 16)
    plans, pts, fks = [], [], []
    for _ in range(10):
        M = int(r.integers(50, 900))
        x = [r.uniform(-np.pi, np.pi, M).astype(np.float32) for _ in range(2)]
        p = Plan(2, shape, n_trans=ntr, eps=1e-5, dtype="complex64")
        p.setpts(*[to_dev(torch, a) for a in x])
        plans.append(p)
        pts.append(x)
        fks.append(to_dev(torch, cnormal(r, (ntr,) + shape, np.dtype("complex64"))))
    nftot = int(np.prod(plans[0].info()["nf"][:2]))
    b = BatchPlan(plans, arena_bytes=3 * ntr * nftot * 8)
    info = b.info()
    assert info["grid_shapes"] == 1 and info["vectors_per_round"] == ntr
    assert info["chunks"] == 4, info
    outs = [torch.empty((ntr, len(x[0])), dtype=torch.complex64, device="cuda") for x in pts]
    b.execute(fks, outs)
    torch.cuda.synchronize()
    for p, fk, o in zip(plans, fks, outs):
        ref = to_host(p.execute(fk))
        assert nudft.relerr(to_host(o), ref) < 5e-6


@pytest.mark.parametrize("dtype", ["complex64", "complex128"])
def test_batch_fused_transform_shapes(torch_gpu, dtype):
    """2-D batches take the fused column / row transforms: rectangular and
    large grids (one column or row per workgroup), 2,3,5-smooth lengths,
    both mode orders and signs, against each problem's own plan (rocFFT)."""
    from finufft_amd import Plan
    from finufft_amd.batch import BatchPlan

    torch = torch_gpu
    r = rng(94)
    real = np.float32 if dtype == "complex64" else np.float64
    tol = 1e-5 if dtype == "complex64" else 1e-10
    ntr = 2
    for modeord, isign in ((0, 1), (1, -1)):
        plans, fks, outs = [], [], []
        for shape in ((37, 250), (512, 9), (1000, 1000), (37, 250)):
            M = 20000
            x = [to_dev(torch, r.uniform(-np.pi, np.pi, M).astype(real)) for _ in range(2)]
            p = Plan(2, shape, n_trans=ntr, eps=tol, dtype=dtype, isign=isign, modeord=modeord)
            p.setpts(*x)
            plans.append(p)
            fks.append(to_dev(torch, cnormal(r, (ntr,) + shape, np.dtype(dtype))))
            outs.append(torch.empty((ntr, M), dtype=fks[-1].dtype, device="cuda"))
        b = BatchPlan(plans)
        assert b.info()["grid_shapes"] == 3
        b.execute(fks, outs)
        torch.cuda.synchronize()
        bound = 2e-5 if dtype == "complex64" else 1e-12
        for p, fk, o in zip(plans, fks, outs):
            ref = to_host(p.execute(fk))
            assert nudft.relerr(to_host(o), ref) < bound, (p.info()["nf"], modeord, isign)


@pytest.mark.parametrize("dtype,dim", [("complex64", 2), ("complex128", 2), ("complex128", 3),
                                       ("complex64", 1)])
def test_batch_type1_matches_plans_and_direct(torch_gpu, dtype, dim):
    """A batch of type-1 plans: every problem's modes equal its own plan's
    (the same spreading and deconvolution, the FFT batched) and the direct
    sum to 25 tol; several rounds of vectors (small arena) and both mode
    orders."""
    from finufft_amd import Plan
    from finufft_amd.batch import BatchPlan

    torch = torch_gpu
    r = rng(120 + dim)
    real = np.float32 if dtype == "complex64" else np.float64
    tol = 1e-5 if dtype == "complex64" else 1e-11
    ntr = 3
    for modeord in (0, 1):
        plans, pts, cs = [], [], []
        for _ in range(9):
            n = int(np.exp(r.uniform(np.log(8), np.log(40 if dim < 3 else 16))))
            M = int(np.exp(r.uniform(np.log(100), np.log(3000))))
            shape = (n * (8 if dim == 1 else 1),) * dim
            x = [r.uniform(-np.pi, np.pi, M).astype(real) for _ in range(dim)]
            p = Plan(1, shape, n_trans=ntr, eps=tol, dtype=dtype, modeord=modeord)
            p.setpts(*[to_dev(torch, a) for a in x])
            plans.append(p)
            pts.append(x)
            cs.append(to_dev(torch, cnormal(r, (ntr, M), np.dtype(dtype))))
        outs = [torch.full((ntr,) + tuple(p.n_modes), float("nan"), dtype=c.dtype, device="cuda")
                for p, c in zip(plans, cs)]
        b = BatchPlan(plans, arena_bytes=1 << 22)  # small arena: several rounds and chunks
        assert b.info()["nprob"] == 9
        b.execute(cs, outs)
        b.execute(cs, outs)  # the recorded graph, replayed
        torch.cuda.synchronize()
        for p, x, c, o in zip(plans, pts, cs, outs):
            got = to_host(o)
            assert np.isfinite(got).all()
            ref_plan = to_host(p.execute(c))
            assert nudft.relerr(got, ref_plan) < 1e-5
            ch = to_host(c)
            ref = np.stack([nudft.type1([a.astype(np.float64) for a in x], ch[t], p._isign,
                                        tuple(p.n_modes), modeord=modeord) for t in range(ntr)])
            err = nudft.relerr(got, ref)
            assert err < 25 * tol, err


def test_batch_type1_refuses_column_plans(torch_gpu):
    """3-D fp32 type-1 plans spread with the column-window method, whose
    subproblem table covers only its dense bins: the batch refuses them."""
    from finufft_amd import FinufftError, Plan
    from finufft_amd.batch import BatchPlan

    torch = torch_gpu
    r = rng(125)
    x = [to_dev(torch, r.uniform(-np.pi, np.pi, 500).astype(np.float32)) for _ in range(3)]
    p = Plan(1, (16, 16, 16), eps=1e-5, dtype="complex64")
    p.setpts(*x)
    assert p.info()["col_side"] > 0
    with pytest.raises(FinufftError):
        BatchPlan([p])
