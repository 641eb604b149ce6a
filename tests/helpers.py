"""Shared test helpers: seeded inputs, device transfer, the plan under test."""

from __future__ import annotations

import numpy as np


def rng(seed=0):
    return np.random.default_rng(seed)


def uniform_pts(r, dim, M, dtype=np.float64):
    return [r.uniform(-np.pi, np.pi, M).astype(dtype) for _ in range(dim)]


def cnormal(r, shape, dtype=np.complex128):
    return (r.standard_normal(shape) + 1j * r.standard_normal(shape)).astype(dtype)


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def to_host(t):
    return t.detach().cpu().numpy()


REAL = {"complex64": np.float32, "complex128": np.float64}
