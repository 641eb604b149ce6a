"""Fused execution of many independent type-2 NUFFT problems.

Wraps finufft_amd_batch_* (include/finufft_amd.h): the type-2 plans of many
problems, each with its own points, modes and sizes, run in cache-sized chunks
of problems of equal grid shape, each chunk as one amplification launch, one
batched FFT and one interpolation launch, recorded once into a HIP graph and
replayed. The reference has no counterpart; its way is a loop of per-plan
executes.
"""

from __future__ import annotations

from ctypes import byref, c_int64, c_void_p

from . import _lib
from ._plan import FinufftError, Plan


class BatchPlan:
    """plans: finufft_amd.Plan objects (type 2, points set) sharing dimension,
    n_trans, isign and tolerance. They must stay alive with their points."""

    def __init__(self, plans, arena_bytes: int = 0):
        self._plans = list(plans)
        if not self._plans:
            raise ValueError("no plans")
        if any(not isinstance(p, Plan) for p in self._plans):
            raise TypeError("plans must be finufft_amd.Plan objects")
        self._double = self._plans[0]._double
        handles = (c_void_p * len(self._plans))(*[p.handle for p in self._plans])
        h = c_void_p(0)
        ier = _lib.lib().finufft_amd_batch_makeplan(int(self._double), len(self._plans), handles,
                                                    int(arena_bytes), byref(h))
        if ier != 0:
            raise FinufftError("finufft_amd_batch_makeplan", ier)
        self._h = h

    def execute(self, fk_list, out_list):
        """fk_list[p]: modes of problem p (plan p's execute input); out_list[p]:
        its values (plan p's execute output), both contiguous on the GPU."""
        n = len(self._plans)
        if len(fk_list) != n or len(out_list) != n:
            raise ValueError("one input and one output per problem")
        c = (c_void_p * n)(*[o.data_ptr() for o in out_list])
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
