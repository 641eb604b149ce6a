"""Python ``Plan`` over the C ABI, with the interface of the reference's
``cufinufft.Plan`` (python/cufinufft/cufinufft/_plan.py:39-400).

Arrays are torch tensors on the GPU, or any GPU array that exposes
``__cuda_array_interface__`` (cupy, numba, pycuda) or DLPack
(``__dlpack__``), taken without a copy, as the reference's wrapper takes
cupy, numba, pycuda and torch arrays (python/cufinufft/cufinufft/_compat.py:
6-111). Results come back as torch tensors, or in the caller's ``out``
array. As in the reference wrapper, ``n_modes`` is given in array-shape
order (slowest axis first) and the points as ``(x, y, z)`` along those
axes; both are reversed on the way into the x-fastest C layout.
"""

from __future__ import annotations

import ctypes
import numbers
from ctypes import byref, c_double, c_int64, c_void_p

from . import _lib

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

_DTYPES = {
    "complex64": ("complex64", "float32", False),
    "complex128": ("complex128", "float64", True),
}


class FinufftError(RuntimeError):
    """A non-zero return code of the library (codes of finufft_errors.h)."""

    def __init__(self, where: str, code: int):
        super().__init__(f"{where} failed with error code {code}")
        self.code = code


def _dtype_name(dtype) -> str:
    if torch is not None and isinstance(dtype, torch.dtype):
        return str(dtype).replace("torch.", "")
    return str(getattr(dtype, "name", dtype))


def _ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def _as_tensor(t):
    """A torch view of a GPU array of another framework (no copy), or t."""
    if torch is None or isinstance(t, torch.Tensor):
        return t
    if hasattr(t, "__cuda_array_interface__"):
        return torch.as_tensor(t, device="cuda")
    if hasattr(t, "__dlpack__"):
        return torch.from_dlpack(t)
    return t


class Plan:
    """A NUFFT plan: make once, set the points, execute many times.

    Args mirror ``cufinufft.Plan``: ``nufft_type`` (1, 2 or 3), ``n_modes``
    (tuple of mode counts; the dimension for type 3), ``n_trans``, ``eps``,
    ``isign`` (defaults: -1 for type 2, +1 otherwise), ``dtype``
    ('complex64' or 'complex128') and option keywords of ``cufinufft_opts``
    (``upsampfac``, ``gpu_method``, ``gpu_sort``, ``gpu_kerevalmeth``,
    ``gpu_maxsubprobsize``, ``gpu_binsize{x,y,z}``, ``gpu_maxbatchsize``,
    ``gpu_spreadinterponly``, ``gpu_device_id``, ``gpu_stream``, ``modeord``,
    ``debug``). ``gpu_stream`` defaults to torch's current stream.
    """

    def __init__(self, nufft_type, n_modes, n_trans=1, eps=1e-6, isign=None,
                 dtype="complex64", **kwargs):
        self._plan = None
        name = _dtype_name(dtype)
        if name not in _DTYPES:
            raise TypeError("Expected complex64 or complex128.")
        self._cname, self._rname, self._double = _DTYPES[name]
        if isign is None:
            isign = -1 if nufft_type == 2 else 1
        if nufft_type == 3:
            if not isinstance(n_modes, numbers.Integral):
                raise ValueError("For a type 3 plan, n_modes must be the dimension")
            dim = int(n_modes)
            n_modes = (1,) * dim
        else:
            n_modes = (int(n_modes),) if isinstance(n_modes, numbers.Integral) else tuple(
                int(m) for m in n_modes)
            dim = len(n_modes)
        if dim not in (1, 2, 3):
            raise ValueError("Only dimensions 1, 2, and 3 supported")
        self._type, self._dim, self._n_modes = nufft_type, dim, n_modes
        self._n_trans, self._isign, self._eps = int(n_trans), int(isign), float(eps)
        self._refs = []
        self._nj = self._nk = 0
        self._pts_gen = 0  # counts setpts calls (autograd checks which points are set)

        opts = _lib.default_opts()
        stream_given = False
        for k, v in kwargs.items():
            if k not in dict(_lib.NufftOpts._fields_):
                raise TypeError(f"Invalid option '{k}'")
            if k == "gpu_stream":
                stream_given = True
                v = getattr(v, "cuda_stream", v)
                v = c_void_p(int(v) if v else 0)
            setattr(opts, k, v)
        if not stream_given and torch is not None and torch.cuda.is_available():
            dev = torch.device("cuda", opts.gpu_device_id)
            opts.gpu_stream = c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        self._opts = opts

        lib = _lib.lib()
        modes_c = (c_int64 * 3)(*(list(self._n_modes[::-1]) + [1] * (3 - dim)))
        handle = c_void_p(0)
        mk = lib.cufinufft_makeplan if self._double else lib.cufinufftf_makeplan
        ier = mk(nufft_type, dim, modes_c, self._isign, self._n_trans, self._eps,
                 byref(handle), byref(opts))
        if ier != 0:
            raise FinufftError("makeplan", ier)
        self._plan = handle.value
        self._pref = "cufinufft" if self._double else "cufinufftf"

    # -- properties as in the reference wrapper
    @property
    def type(self):
        return self._type

    @property
    def dtype(self):
        return self._cname

    @property
    def dim(self):
        return self._dim

    @property
    def n_modes(self):
        return self._n_modes

    @property
    def n_trans(self):
        return self._n_trans

    @property
    def handle(self) -> int:
        """Raw plan pointer for the C ABI."""
        return self._plan

    def info(self) -> dict:
        """Kernel width, grid, method, bins and batch the plan settled on."""
        iout = (c_int64 * 16)()
        dout = (c_double * 4)()
        ier = _lib.lib().finufft_amd_plan_info(self._plan, int(self._double), iout, dout)
        if ier != 0:
            raise FinufftError("plan_info", ier)
        v = list(iout)
        return {
            "type": v[0], "dim": v[1], "ns": v[2], "nf": tuple(v[3:6]), "gpu_method": v[6],
            "binsize": tuple(v[7:10]), "batchsize": v[10], "M": v[11], "nbins": v[12],
            "col_side": v[13], "colored": v[14], "dense_subproblems": v[15],
            "beta": dout[0], "upsampfac": dout[1], "np": int(dout[2]),
            "fft_path": int(dout[3]),
        }

    STAGES = ("zero", "spread", "fft", "deconvolve", "interp", "type3")

    def set_stage_timing(self, on: bool = True):
        """Record HIP events around every execute stage (see stage_times)."""
        ier = _lib.lib().finufft_amd_set_stage_timing(self._plan, int(self._double), int(on))
        if ier != 0:
            raise FinufftError("set_stage_timing", ier)

    def stage_times(self) -> dict:
        """Milliseconds per stage since the last call (waits for the GPU)."""
        ms = (c_double * 8)()
        ier = _lib.lib().finufft_amd_get_stage_times(self._plan, int(self._double), ms)
        if ier != 0:
            raise FinufftError("get_stage_times", ier)
        return {k: ms[i] for i, k in enumerate(self.STAGES)}

    def _check(self, t, name, real, output=False):
        if t is None:
            return None
        t = _as_tensor(t)
        if torch is None or not isinstance(t, torch.Tensor):
            raise TypeError(f"Argument `{name}` must be a GPU array (a torch tensor, or an "
                            "array with __cuda_array_interface__ or __dlpack__)")
        want = getattr(torch, self._rname if real else self._cname)
        if t.dtype != want:
            raise TypeError(f"Argument `{name}` does not have the correct dtype: "
                            f"{t.dtype} was given, but {want} was expected.")
        if not t.is_cuda:
            raise TypeError(f"Argument `{name}` must be on the GPU")
        if not t.is_contiguous():
            if output:
                raise ValueError(f"Argument `{name}` must be contiguous")
            t = t.contiguous()
            self._hold_for_plan_stream(t)
        return t

    def _hold_for_plan_stream(self, t):
        """A tensor allocated here (a contiguous copy of an input, or an
        output) that the library reads or writes asynchronously on the plan's
        stream: when that is not torch's current stream, torch's allocator is
        told not to reuse its memory before the plan's stream has reached
        this point."""
        ptr = self._opts.gpu_stream or 0
        if ptr and ptr != torch.cuda.current_stream(t.device).cuda_stream:
            t.record_stream(torch.cuda.ExternalStream(ptr, device=t.device))

    def setpts(self, x, y=None, z=None, s=None, t=None, u=None):
        """Set the nonuniform points (and the type-3 targets s, t, u)."""
        pts = [self._check(a, n, True) for a, n in ((x, "x"), (y, "y"), (z, "z"))]
        if any(p is None for p in pts[: self._dim]):
            raise TypeError("missing point coordinates for the plan's dimension")
        M = pts[0].numel()
        for p in pts[: self._dim]:
            if p.numel() != M:
                raise TypeError("Arrays `x`, `y`, `z` must be of the same length")
        trg = [None, None, None]
        N = 0
        if self._type == 3:
            trg = [self._check(a, n, True) for a, n in ((s, "s"), (t, "t"), (u, "u"))]
            if any(p is None for p in trg[: self._dim]):
                raise TypeError("missing target coordinates for the plan's dimension")
            N = trg[0].numel()
        # reverse into the x-fastest C order
        cp = [_ptr(p) for p in pts[: self._dim][::-1]] + [None] * (3 - self._dim)
        ct = ([_ptr(p) for p in trg[: self._dim][::-1]] + [None] * (3 - self._dim)
              if self._type == 3 else [None] * 3)
        self._refs = [p for p in pts + trg if p is not None]
        ier = getattr(_lib.lib(), f"{self._pref}_setpts")(self._plan, M, *cp, N, *ct)
        self._nj, self._nk = M, N
        self._pts_gen += 1
        if ier != 0:
            raise FinufftError("setpts", ier)

    def _shapes(self, adjoint):
        kind = self._type
        if adjoint:
            kind = 2 if kind == 1 else 1
        if kind == 1:
            return (self._nj,), tuple(self._n_modes)
        if kind == 2:
            return tuple(self._n_modes), (self._nj,)
        return (self._nj,), (self._nk,)

    def _run(self, data, out, adjoint):
        if adjoint and self._type == 3:
            raise ValueError("execute_adjoint is provided for types 1 and 2")
        data = self._check(data, "data", False)
        in_shape, out_shape = self._shapes(adjoint)
        n = self._n_trans
        if data.shape == in_shape and n == 1:
            batch = ()
        elif data.shape == (n, *in_shape) or (n == 1 and data.shape == (1, *in_shape)):
            batch = (n,)
        else:
            raise ValueError(f"Invalid data shape {tuple(data.shape)}, expected "
                             f"{(n, *in_shape) if n > 1 else in_shape}")
        given = out
        if out is None:
            out = torch.empty(batch + out_shape, dtype=data.dtype, device=data.device)
            self._hold_for_plan_stream(out)
        else:
            out = self._check(out, "out", False, output=True)
            if out.numel() != n * int(torch.tensor(out_shape).prod()):
                raise ValueError("Invalid out shape")
        kind = self._type if not adjoint else (2 if self._type == 1 else 1)
        # c is the NU-point side, fk the mode side (types 1/2); type 3: c -> fk
        if kind == 2:
            c_ptr, fk_ptr = out.data_ptr(), data.data_ptr()
        else:
            c_ptr, fk_ptr = data.data_ptr(), out.data_ptr()
        fn = f"{self._pref}_execute_adjoint" if adjoint else f"{self._pref}_execute"
        ier = getattr(_lib.lib(), fn)(self._plan, c_ptr, fk_ptr)
        if ier != 0:
            raise FinufftError(fn, ier)
        # the caller's own output array (of any framework) when one was given
        return out if given is None or isinstance(given, torch.Tensor) else given

    def execute(self, data, out=None):
        """Type 1: strengths -> modes; type 2: modes -> values; type 3: strengths -> targets."""
        return self._run(data, out, adjoint=False)

    def execute_adjoint(self, data, out=None):
        """The adjoint transform on the same points (types 1 and 2)."""
        return self._run(data, out, adjoint=True)

    def interp_grad(self, fk, value=True):
        """The type-2 direction of the plan (type 2: the transform, type 1: the
        adjoint) applied to the modes fk, and the derivatives of the result
        with respect to the point coordinates.

        Returns (c, dc): c has the shape of the values (None when value is
        False); dc[k] is the derivative with respect to the k-th coordinate
        array passed to setpts (x, y, z), each with the shape of c.
        """
        if self._type == 3:
            raise ValueError("interp_grad is provided for types 1 and 2")
        fk = self._check(fk, "fk", False)
        modes = tuple(self._n_modes)
        n = self._n_trans
        if fk.shape == modes and n == 1:
            batch = ()
        elif fk.shape == (n, *modes) or (n == 1 and fk.shape == (1, *modes)):
            batch = (n,)
        else:
            raise ValueError(f"Invalid fk shape {tuple(fk.shape)}")
        c = torch.empty(batch + (self._nj,), dtype=fk.dtype, device=fk.device) if value else None
        dc = torch.empty((self._dim,) + batch + (self._nj,), dtype=fk.dtype, device=fk.device)
        for t in (c, dc):
            if t is not None:
                self._hold_for_plan_stream(t)
        fn = f"{self._pref}_interp_grad"
        ier = getattr(_lib.lib(), fn)(self._plan, c.data_ptr() if value else None,
                                      dc.data_ptr(), fk.data_ptr())
        if ier != 0:
            raise FinufftError(fn, ier)
        # the library's dimension 0 is x, the last array axis
        return c, dc.flip(0)

    def __del__(self):
        if getattr(self, "_plan", None):
            try:
                getattr(_lib.lib(), f"{self._pref}_destroy")(self._plan)
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
            self._plan = None
