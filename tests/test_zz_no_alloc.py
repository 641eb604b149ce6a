"""The allocation-free density entry point (finufft_amd_density_stats_noalloc)
against the allocating one (finufft_amd_density_stats), through the C ABI:
two calls in a row on one stream, the second with a smaller input, each
result equal to the allocating entry point's on the same points; a grid
beyond the library's scratch array is refused with an error."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(torch, lib, pts, dim, g, is_double):
    M = int(pts[0].numel())
    ptrs = [int(p.data_ptr()) for p in reversed(pts)] + [None] * (3 - dim)
    ref = (ctypes.c_double * 3)()
    assert lib.finufft_amd_density_stats(dim, M, ptrs[0], ptrs[1], ptrs[2], is_double, g, ref,
                                         None) == 0
    d_stats = torch.full((3,), -1.0, dtype=torch.float64, device="cuda")
    assert lib.finufft_amd_density_stats_noalloc(dim, M, ptrs[0], ptrs[1], ptrs[2], is_double, g,
                                                 d_stats.data_ptr(), None) == 0
    torch.cuda.synchronize()
    return list(ref), d_stats.cpu().tolist()


@pytest.mark.parametrize("dim,g", [(1, 4096), (2, 128), (3, 32), (3, 64)])
@pytest.mark.parametrize("real", ["float32", "float64"])
def test_noalloc_matches_allocating(torch_gpu, dim, g, real):
    from finufft_amd import _lib, costmodel

    torch = torch_gpu
    lib = _lib.lib()
    is_double = 1 if real == "float64" else 0
    for kind, M in (("clustered", 300_000), ("uniform", 40_000)):  # second call smaller
        pts = costmodel.make_points(kind, dim, M, seed=9, real=real)
        ref, got = _both(torch, lib, pts, dim, g, is_double)
        print(dim, g, real, kind, ref, got)
        assert got == ref
    # no points: zeros
    empty = [torch.empty(0, dtype=getattr(torch, real), device="cuda")] * dim
    d_stats = torch.full((3,), -1.0, dtype=torch.float64, device="cuda")
    assert lib.finufft_amd_density_stats_noalloc(dim, 0, None, None, None, is_double, g,
                                                 d_stats.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert d_stats.cpu().tolist() == [0.0, 0.0, 0.0]
    del empty


def test_noalloc_refuses_oversized_grid(torch_gpu):
    from finufft_amd import _lib

    torch = torch_gpu
    lib = _lib.lib()
    x = torch.zeros(16, dtype=torch.float32, device="cuda")
    d_stats = torch.zeros(3, dtype=torch.float64, device="cuda")
    # 128^3 = 2^21 cells > 2^18
    ier = lib.finufft_amd_density_stats_noalloc(3, 16, x.data_ptr(), x.data_ptr(), x.data_ptr(),
                                                0, 128, d_stats.data_ptr(), None)
    assert ier == 21  # FINUFFT_ERR_INVALID_ARGUMENT
