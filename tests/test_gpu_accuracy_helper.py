"""The FP64 GPU direct sums used by bench.py agree with the CPU oracle."""

import numpy as np
import pytest

from oracle import nudft
from tests.helpers import cnormal, rng, to_dev, to_host, uniform_pts

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(9,), (7, 10), (6, 8, 5)])
def test_direct_sums_match_oracle(torch_gpu, shape):
    from finufft_amd import accuracy as acc

    torch = torch_gpu
    r = rng(21)
    M = 700
    pts = uniform_pts(r, len(shape), M, np.float64)
    c = cnormal(r, (M,))
    f = cnormal(r, shape)
    dp = [to_dev(torch, p) for p in pts]
    midx = r.choice(int(np.prod(shape)), min(25, int(np.prod(shape))), replace=False)
    got1 = to_host(acc.type1_at(dp, to_dev(torch, c), 1, shape, midx, chunk=128))
    ref1 = nudft.type1_at(pts, c, 1, shape, midx)
    assert np.allclose(got1, ref1, rtol=1e-12, atol=1e-9)
    pidx = r.choice(M, 30, replace=False)
    got2 = to_host(acc.type2_at(dp, to_dev(torch, f), -1, pidx))
    ref2 = nudft.type2(pts, f, -1, idx=pidx)
    assert np.allclose(got2, ref2, rtol=1e-12, atol=1e-9)
