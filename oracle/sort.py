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


def start_bin_keys(pts_c, nf, binsize, ns, fold64=True):
    """Bins of the first support node s = ceil(r - ns/2), wrapped into [0, nf),
    with r folded in fp64 (fp32 plans) or in the points' own precision."""
    nb = nbins(nf, binsize)
    key = np.zeros(len(pts_c[0]), np.int64)
    stride = 1
    for d, x in enumerate(pts_c):
        xx = x.astype(np.float64) if fold64 else x
        r = cfold.fold_rescale(xx, nf[d])
        r = np.where((r >= 0) & (r < nf[d]), r, 0)
        s = np.ceil(r - np.asarray(ns * 0.5, r.dtype)).astype(np.int64)
        s = np.where(s < 0, s + nf[d], s)
        b = np.clip(s // binsize[d], 0, nb[d] - 1)
        key += b * stride
        stride *= nb[d]
    return key


def bin_sort(pts_c, nf, binsize, start_ns=None, fold64=True):
    """(permutation, binstart) of the stable bin sort; start_ns selects bins of
    the first support node for kernel width start_ns."""
    if start_ns is None:
        key = bin_keys(pts_c, nf, binsize)
    else:
        key = start_bin_keys(pts_c, nf, binsize, start_ns, fold64)
    perm = np.argsort(key, kind="stable").astype(np.int32)
    total = int(np.prod(nbins(nf, binsize)))
    counts = np.bincount(key, minlength=total)
    binstart = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return perm, binstart


def bin_sort_bruteforce(pts_c, nf, binsize):
    """Same permutation by a plain counting loop over bins; for tiny inputs."""
    key = bin_keys(pts_c, nf, binsize)
    out = []
    for b in range(int(np.prod(nbins(nf, binsize)))):
        for i in range(len(key)):
            if key[i] == b:
                out.append(i)
    return np.array(out, np.int32)
