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


def start_nodes(pts_c, nf, ns, fold64=True):
    """First support node s = ceil(r - ns/2) per dimension, wrapped into
    [0, nf), with r folded in fp64 (fp32 plans) or in the points' own
    precision; non-finite r is pinned to 0."""
    out = []
    for d, x in enumerate(pts_c):
        xx = x.astype(np.float64) if fold64 else x
        r = cfold.fold_rescale(xx, nf[d])
        r = np.where((r >= 0) & (r < nf[d]), r, 0)
        s = np.ceil(r - np.asarray(ns * 0.5, r.dtype)).astype(np.int64)
        out.append(np.where(s < 0, s + nf[d], s))
    return out


def start_bin_keys(pts_c, nf, binsize, ns, fold64=True, col_side=0, bank=False):
    """Bins of the first support node. With col_side (3-D), the key is refined
    inside each bin to (sub-column, local z): sub-columns are col_side x
    col_side first-node squares in (x, y), x fastest, and points of a
    sub-column are ordered by the local z of their first node. With bank, the
    key is refined by (lx + P0 ly + P0 P1 lz) mod 32, P = binsize + ns - 1 (the
    LDS bank class of the point's first tile cell)."""
    nb = nbins(nf, binsize)
    key = np.zeros(len(pts_c[0]), np.int64)
    loc = []
    stride = 1
    for d, s in enumerate(start_nodes(pts_c, nf, ns, fold64)):
        b = np.clip(s // binsize[d], 0, nb[d] - 1)
        loc.append(np.clip(s - b * binsize[d], 0, binsize[d] - 1))
        key += b * stride
        stride *= nb[d]
    if col_side:
        nsx, nsy = binsize[0] // col_side, binsize[1] // col_side
        sub = np.minimum(loc[0] // col_side, nsx - 1) + nsx * np.minimum(loc[1] // col_side, nsy - 1)
        key = (key * (nsx * nsy) + sub) * binsize[2] + loc[2]
    elif bank:
        P = [b + ns - 1 for b in binsize]
        cell = loc[0].copy()
        if len(loc) > 1:
            inner = loc[1] + (P[1] * loc[2] if len(loc) > 2 else 0)
            cell = cell + P[0] * inner
        key = key * 32 + (cell & 31)
    return key


def bin_sort(pts_c, nf, binsize, start_ns=None, fold64=True, col_side=0, bank=False):
    """(permutation, binstart) of the stable bin sort; start_ns selects bins of
    the first support node for kernel width start_ns, col_side the sub-column
    refinement of those bins (3-D), bank the bank-class refinement."""
    if start_ns is None:
        key = bin_keys(pts_c, nf, binsize)
        bin_key = key
    else:
        key = start_bin_keys(pts_c, nf, binsize, start_ns, fold64, col_side, bank)
        refined = col_side or bank
        bin_key = start_bin_keys(pts_c, nf, binsize, start_ns, fold64) if refined else key
    perm = np.argsort(key, kind="stable").astype(np.int32)
    total = int(np.prod(nbins(nf, binsize)))
    counts = np.bincount(bin_key, minlength=total)
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
