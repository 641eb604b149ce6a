"""Fused execution of many independent NUFFT problems of one type (1 or 2).

Wraps finufft_amd_batch_* (include/finufft_amd.h): the plans of many problems,
each with its own points, modes and sizes, run in cache-sized chunks of
problems of equal grid shape. A type-2 chunk is one amplification (or fused
column / row transform) stage and one interpolation launch; a type-1 chunk is
one clearing, one spreading launch over all its problems, one batched FFT and
one deconvolution launch. The schedule is recorded once into a HIP graph and
replayed. The reference has no counterpart; its way is a loop of per-plan
executes.
"""

from __future__ import annotations

from ctypes import byref, c_int64, c_void_p

import torch

from . import _lib
from ._plan import FinufftError, Plan


class BatchPlan:
    """plans: finufft_amd.Plan objects (all type 1 or all type 2, points set)
    sharing dimension, n_trans, isign and tolerance. They must stay alive with
    their points."""

    def __init__(self, plans, arena_bytes: int = 0):
        self._plans = list(plans)
        if not self._plans:
            raise ValueError("no plans")
        if any(not isinstance(p, Plan) for p in self._plans):
            raise TypeError("plans must be finufft_amd.Plan objects")
        self._double = self._plans[0]._double
        if any(p._double != self._double for p in self._plans):
            raise TypeError("plans must share the precision")
        self._dtype = getattr(torch, self._plans[0].dtype)
        self._type = self._plans[0].type  # the library checks that all share it
        # element counts of every problem's modes and values (n_trans vectors)
        self._fk_sizes, self._out_sizes = [], []
        for p in self._plans:
            modes = 1
            for m in p.n_modes:
                modes *= int(m)
            self._fk_sizes.append(p.n_trans * modes)
            self._out_sizes.append(p.n_trans * p._nj)
        handles = (c_void_p * len(self._plans))(*[p.handle for p in self._plans])
        h = c_void_p(0)
        ier = _lib.lib().finufft_amd_batch_makeplan(int(self._double), len(self._plans), handles,
                                                    int(arena_bytes), byref(h))
        if ier != 0:
            raise FinufftError("finufft_amd_batch_makeplan", ier)
        self._h = h

    def execute(self, in_list, out_list):
        """in_list[p]: plan p's execute input (type 2: its modes; type 1: its
        strengths), out_list[p]: its output (values; modes), all contiguous on
        the GPU."""
        n = len(self._plans)
        if self._type == 1:  # strengths in, modes out
            fk_list, c_list = out_list, in_list
        else:
            fk_list, c_list = in_list, out_list
        if len(fk_list) != n or len(c_list) != n:
            raise ValueError("one input and one output per problem")
        # The library reads and writes through these pointers without knowing
        # the sizes: every array is checked against its plan (kept short, it
        # runs per problem on every execute).
        want = self._dtype
        for i, (f, o, nf, no) in enumerate(zip(fk_list, c_list, self._fk_sizes,
                                              self._out_sizes)):
            if not (f.dtype == want and o.dtype == want and f.is_cuda and o.is_cuda
                    and f.is_contiguous() and o.is_contiguous()):
                raise TypeError(f"problem {i}: inputs and outputs must be contiguous "
                                f"{want} tensors on the GPU")
            if f.numel() != nf or o.numel() != no:
                raise ValueError(f"problem {i}: {f.numel()} modes and {o.numel()} values "
                                 f"given, its plan needs {nf} and {no}")
        c = (c_void_p * n)(*[o.data_ptr() for o in c_list])
        fk = (c_void_p * n)(*[f.data_ptr() for f in fk_list])
        ier = _lib.lib().finufft_amd_batch_execute(self._h, c, fk)
        if ier != 0:
            raise FinufftError("finufft_amd_batch_execute", ier)
        return out_list

    def set_graph(self, on: bool):
        ier = _lib.lib().finufft_amd_batch_set_graph(self._h, int(on))
        if ier != 0:
            raise FinufftError("finufft_amd_batch_set_graph", ier)

    def info(self) -> dict:
        v = (c_int64 * 7)()
        ier = _lib.lib().finufft_amd_batch_info(self._h, v)
        if ier != 0:
            raise FinufftError("finufft_amd_batch_info", ier)
        return {"nprob": v[0], "vectors_per_round": v[1], "grid_shapes": v[2], "graph": v[3],
                "capture_stage": v[4], "capture_error": v[5], "chunks": v[6]}

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                _lib.lib().finufft_amd_batch_destroy(self._h)
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
            self._h = None
