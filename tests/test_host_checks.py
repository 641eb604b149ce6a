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


def _bare_batch(fk_sizes, out_sizes, ttype=2):
    """A BatchPlan (of type-2 plans unless ttype says otherwise) with its
    per-problem sizes set and no library handle."""
    b = object.__new__(BatchPlan)
    b._plans = [None] * len(fk_sizes)
    b._fk_sizes, b._out_sizes = list(fk_sizes), list(out_sizes)
    b._dtype = torch.complex64
    b._type = ttype
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


class _FakePlan:
    """Records which points are set when the adjoint runs (no GPU)."""

    def __init__(self):
        self._pts_gen = 0
        self.current = None
        self.seen = []

    def setpts(self, *pts):
        self._pts_gen += 1
        self.current = pts[0].clone()

    def execute(self, data):
        return data.clone()

    def execute_adjoint(self, g):
        self.seen.append(self.current.clone())
        return g.clone()


@pytest.mark.parametrize("kind", ["type1", "type2"])
def test_autograd_two_forwards_one_backward(kind):
    """Two forward calls on one plan before a backward: each backward runs
    with the points of its own forward call."""
    from finufft_amd import autograd

    fn = autograd._Type1 if kind == "type1" else autograd._Type2
    plan = _FakePlan()
    x1, x2 = torch.linspace(0, 1, 5), torch.linspace(2, 3, 5)
    d1 = torch.ones(5, dtype=torch.complex64, requires_grad=True)
    d2 = torch.ones(5, dtype=torch.complex64, requires_grad=True)
    plan.setpts(x1)
    y1 = fn.apply(d1, plan, x1)
    plan.setpts(x2)
    y2 = fn.apply(d2, plan, x2)
    (y1.abs().sum() + 2 * y2.abs().sum()).backward()
    assert len(plan.seen) == 2
    # the graph runs the second call's backward first (its points are set),
    # then the first call's, which sets x1 again
    got = sorted(plan.seen, key=lambda t: float(t[0]))
    assert torch.equal(got[0], x1) and torch.equal(got[1], x2)
    assert torch.allclose(d1.grad.real, torch.ones(5)) and torch.allclose(d2.grad.real,
                                                                         2 * torch.ones(5))


def test_type1_batch_takes_strengths_in_and_modes_out():
    """A type-1 batch's execute takes the strengths as input and the modes as
    output: the sizes are checked that way round before the library runs."""
    b = _bare_batch([8, 8], [5, 6], ttype=1)
    c = [torch.empty(5, dtype=torch.complex64), torch.empty(6, dtype=torch.complex64)]
    fk = [torch.empty(8, dtype=torch.complex64), torch.empty(8, dtype=torch.complex64)]
    # the right sizes get as far as the device check (CPU tensors here)
    with pytest.raises(TypeError, match="problem 0"):
        b.execute(c, fk)
    # swapped (type-2 order) fails the size check
    with pytest.raises((TypeError, ValueError)):
        b.execute(fk, c)
