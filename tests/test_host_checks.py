"""Host-side argument checks (no GPU) of the Python wrappers whose library
calls take bare device pointers: BatchPlan.execute (the library writes
through the pointers it is given without knowing their sizes, so every array
is checked against its plan before the call) and the cost model's network
kernels."""

import pytest
import torch

from finufft_amd.batch import BatchPlan


class FakeTensor:
    """Stands in for a GPU tensor: the attributes the checks read."""

    def __init__(self, n, dtype=torch.complex64, cuda=True, contiguous=True):
        self.n, self.dtype, self.is_cuda, self._contiguous = n, dtype, cuda, contiguous

    def numel(self):
        return self.n

    def is_contiguous(self):
        return self._contiguous

    def data_ptr(self):
        raise AssertionError("the checks must fail before any pointer is taken")


def _bare_batch(fk_sizes, out_sizes):
    """A BatchPlan with its per-problem sizes set and no library handle."""
    b = object.__new__(BatchPlan)
    b._plans = [None] * len(fk_sizes)
    b._fk_sizes, b._out_sizes = list(fk_sizes), list(out_sizes)
    b._dtype = torch.complex64
    b._h = None
    return b


def test_one_array_per_problem():
    b = _bare_batch([8, 8], [5, 6])
    with pytest.raises(ValueError, match="one input and one output"):
        b.execute([FakeTensor(8)], [FakeTensor(5), FakeTensor(6)])


@pytest.mark.parametrize("bad", [dict(cuda=False), dict(dtype=torch.complex128),
                                 dict(contiguous=False)])
def test_device_dtype_contiguity(bad):
    b = _bare_batch([8, 8], [5, 6])
    with pytest.raises(TypeError, match="problem 1"):
        b.execute([FakeTensor(8), FakeTensor(8)], [FakeTensor(5), FakeTensor(6, **bad)])


@pytest.mark.parametrize("fk,out", [((8, 7), (5, 6)), ((8, 8), (5, 7))])
def test_sizes_against_the_plans(fk, out):
    b = _bare_batch([8, 8], [5, 6])
    with pytest.raises(ValueError, match="problem 1: .* its plan needs 8 and 6"):
        b.execute([FakeTensor(n) for n in fk], [FakeTensor(n) for n in out])


def test_costmodel_pointer_checks():
    from finufft_amd import costmodel

    with pytest.raises(TypeError, match="float32 tensor on the GPU"):
        costmodel._ptr(torch.zeros(costmodel.NPARAM), costmodel.NPARAM, "params")
    with pytest.raises(TypeError):
        costmodel._ptr(torch.zeros(costmodel.NPARAM, dtype=torch.float64))
    with pytest.raises(TypeError):
        costmodel._ptr(None)
