"""GPU arrays of other frameworks: the Python Plan takes any array exposing
__cuda_array_interface__ (cupy, numba, pycuda) or DLPack, without a copy,
as the reference's wrapper takes cupy, numba, pycuda and torch arrays
(python/cufinufft/cufinufft/_compat.py:6-111; its tests run every case for
each framework, python/cufinufft/tests/test_array_ordering.py). Neither
cupy nor numba is installed here, so the arrays are minimal stand-ins that
expose only the interface, over torch-allocated memory."""

import numpy as np
import pytest

from oracle import nudft
from tests.helpers import cnormal, rng, to_dev, to_host

pytestmark = pytest.mark.gpu


TYPESTR = {"torch.float32": "<f4", "torch.float64": "<f8", "torch.complex64": "<c8",
           "torch.complex128": "<c16"}


class CudaArray:
    """An array known only by its __cuda_array_interface__ (version 3)."""

    def __init__(self, t):
        self._t = t  # owns the memory
        self.__cuda_array_interface__ = {
            "shape": tuple(t.shape), "typestr": TYPESTR[str(t.dtype)],
            "data": (t.data_ptr(), False), "version": 3, "strides": None}


class DlpackArray:
    """An array known only by __dlpack__ / __dlpack_device__."""

    def __init__(self, t):
        self._t = t

    def __dlpack__(self, stream=None):
        return self._t.__dlpack__(stream=stream) if stream is not None else self._t.__dlpack__()

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()


@pytest.mark.parametrize("wrap", [CudaArray, DlpackArray])
def test_type1_and_type2_with_foreign_arrays(torch_gpu, wrap):
    from finufft_amd import Plan

    torch = torch_gpu
    r = rng(40)
    shape, M, tol = (20, 24), 900, 1e-5
    pts = [r.uniform(-np.pi, np.pi, M).astype(np.float32) for _ in range(2)]
    p64 = [p.astype(np.float64) for p in pts]
    dpts = [wrap(to_dev(torch, p)) for p in pts]
    c = cnormal(r, (M,), np.complex64)
    p1 = Plan(1, shape, eps=tol, dtype="complex64")
    p1.setpts(*dpts)
    out = torch.zeros(shape, dtype=torch.complex64, device="cuda")
    res = p1.execute(wrap(to_dev(torch, c)), out=wrap(out))
    assert isinstance(res, wrap)  # the caller's own output array comes back
    torch.cuda.synchronize()
    assert nudft.relerr(to_host(out), nudft.type1(p64, c, 1, shape)) < 25 * tol
    fk = cnormal(r, shape, np.complex64)
    p2 = Plan(2, shape, eps=tol, dtype="complex64")
    p2.setpts(*dpts)
    got = p2.execute(wrap(to_dev(torch, fk)))  # no out: a torch tensor
    assert nudft.relerr(to_host(got), nudft.type2(p64, fk, -1)) < 25 * tol


def test_non_array_refused(torch_gpu):
    from finufft_amd import Plan

    p = Plan(1, (16,), eps=1e-5, dtype="complex64")
    with pytest.raises(TypeError, match="GPU array"):
        p.setpts([0.0, 1.0])
