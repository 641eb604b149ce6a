"""Oracle of the point bin sort.

Bin of a point (include/cufinufft/spreadinterp.hpp:214-229): per dimension
``b = floor(fold_rescale(x) * (1/binsize))`` in the plan precision, then
``b = b - 1 if b >= nbins``, ``b = max(b, 0)``; the flat bin is x-fastest.
The permutation is the stable order by bin (ties by increasing original
index), i.e. the order of the reference's CPU bin_sort_singlethread_impl
(include/finufft/spread.hpp:459-584) for this bin geometry.

Test infrastructure only (see oracle/__init__.py).
"""

from __future__ import annotations

import numpy as np

from . import cfold


def nbins(nf, binsize):
    return [(n + b - 1) // b for n, b in zip(nf, binsize)]


def bin_keys(pts_c, nf, binsize):
    """pts_c: coordinates in C order (x fastest first), each float32/float64."""
    dt = pts_c[0].dtype.type
    nb = nbins(nf, binsize)
    key = np.zeros(len(pts_c[0]), np.int64)
    stride = 1
    for d, x in enumerate(pts_c):
        r = cfold.fold_rescale(x, nf[d])
        inv = dt(1) / dt(binsize[d])
        b = np.floor(r * inv).astype(np.int64)  # product rounded in the plan precision
        b = np.where(b >= nb[d], b - 1, b)
        b = np.maximum(b, 0)
        key += b * stride
        stride *= nb[d]
    return key


def bin_sort(pts_c, nf, binsize):
    """(permutation, binstart) of the stable bin sort."""
    key = bin_keys(pts_c, nf, binsize)
    perm = np.argsort(key, kind="stable").astype(np.int32)
    total = int(np.prod(nbins(nf, binsize)))
    counts = np.bincount(key, minlength=total)
    binstart = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return perm, binstart


def bin_sort_bruteforce(pts_c, nf, binsize):
    """Same permutation via lexsort on (index, key); for tiny inputs."""
    key = bin_keys(pts_c, nf, binsize)
    idx = np.arange(len(key))
    return np.lexsort((idx, key)).astype(np.int32)
