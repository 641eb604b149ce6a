"""CPU NUFFT restatement (types 1 and 2, fp32) used as the CPU baseline.

Spread / interpolate / deconvolve run in C with OpenMP
(oracle/csrc/cpu_nufft.c); the fine-grid FFT is scipy.fft with the same
thread count. Kernel width, shape, Horner coefficients (PSWF) and the fine grid
size come from oracle.kernel (the reference's rules). The kernel's Fourier series
uses numpy's Gauss-Legendre rule with the node count of src/cuda/fft.cu:185-202.

Pinned by tests/test_abi_cpu.py::test_cpu_restatement_vs_direct against the FP64
direct sum. Test
infrastructure and CPU baseline only (see oracle/__init__.py).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from . import cfold, kernel


def _cfuncs():
    lib = cfold.lib()
    if not hasattr(lib, "_cpu_declared"):
        vp, i64, ci = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        lib.cpu_spread_f.argtypes = [ci, i64, vp, vp, vp, vp, vp, ci, ci, vp, vp, ci]
        lib.cpu_spread_f.restype = None
        lib.cpu_interp_f.argtypes = [ci, i64, vp, vp, vp, vp, vp, ci, ci, vp, vp, ci]
        lib.cpu_interp_f.restype = None
        lib.cpu_deconv_f.argtypes = [ci, vp, vp, vp, vp, vp, vp, vp, ci, ci]
        lib.cpu_deconv_f.restype = None
        lib._cpu_declared = True
    return lib


def fourier_series_gl(ns, beta, nf, formula="pswf"):
    """(-1)^k sum_n f_n 2 cos(2 pi k z_n / nf), k = 0..nf/2 (Gauss-Legendre),
    of the PSWF (default) or ES kernel."""
    q = int(2 + 1.5 * ns)
    x, w = np.polynomial.legendre.leggauss(2 * q)
    z = (ns / 2.0) * x[:q]
    phi = kernel.pswf(beta, 2.0 * z / ns) if formula == "pswf" else kernel.es(z, ns, beta)
    f = (ns / 2.0) * w[:q] * phi
    k = np.arange(nf // 2 + 1)
    out = (f[None, :] * 2 * np.cos(2 * math.pi * np.outer(k, z) / nf)).sum(axis=1)
    return out * np.where(k % 2 == 1, -1.0, 1.0)


class CpuPlan:
    """Type 1 or 2, fp32, upsampfac 2, CMCL mode order; python axis order."""

    def __init__(self, nufft_type, n_modes, eps=1e-6, isign=None, nthreads=1):
        import scipy.fft  # noqa: F401 - fail early if missing

        self.type = nufft_type
        self.isign = isign if isign is not None else (-1 if nufft_type == 2 else 1)
        self.dim = len(n_modes)
        self.ms_c = list(n_modes)[::-1]
        # the library's default kernel: PSWF with the width law for this
        # dimension and type
        self.ns, self.beta = kernel.kernel_shape_pswf(eps, 2.0, len(n_modes), nufft_type, False)
        nc = self.ns + 2 + (self.ns <= 7) - (self.ns == 2)
        coef = kernel.horner_coeffs_pswf(self.ns, self.beta, nc)
        self.nc = coef.shape[0]
        self.coef = np.ascontiguousarray(coef, np.float64)
        self.nf_c = [kernel.fine_grid_len(m, 2.0, self.ns) for m in self.ms_c]
        self.ker = [np.ascontiguousarray(fourier_series_gl(self.ns, self.beta, n))
                    for n in self.nf_c]
        self.nthreads = nthreads
        self.pts_c = None

    def setpts(self, *pts):
        self.pts_c = [np.ascontiguousarray(p, np.float32) for p in pts[::-1]]
        self.M = len(self.pts_c[0])

    def _ptrs(self):
        p = self.pts_c + [self.pts_c[0]] * (3 - self.dim)
        return [a.ctypes.data for a in p]

    def execute(self, data):
        import scipy.fft

        lib = _cfuncs()
        nf = (ctypes.c_int * 3)(*(self.nf_c + [1] * (3 - self.dim)))
        ms = (ctypes.c_int * 3)(*(self.ms_c + [1] * (3 - self.dim)))
        grid_shape = tuple(self.nf_c[::-1])
        kp = [k.ctypes.data for k in self.ker] + [self.ker[0].ctypes.data] * (3 - self.dim)
        fw = np.zeros(grid_shape, np.complex64)
        if self.type == 1:
            c = np.ascontiguousarray(data, np.complex64)
            lib.cpu_spread_f(self.dim, self.M, *self._ptrs(), c.ctypes.data, nf, self.ns,
                             self.nc, self.coef.ctypes.data, fw.ctypes.data, self.nthreads)
            if self.isign > 0:
                fw = scipy.fft.ifftn(fw, norm="forward", workers=self.nthreads, overwrite_x=True)
            else:
                fw = scipy.fft.fftn(fw, workers=self.nthreads, overwrite_x=True)
            fw = np.ascontiguousarray(fw, np.complex64)
            fk = np.empty(tuple(self.ms_c[::-1]), np.complex64)
            lib.cpu_deconv_f(self.dim, ms, nf, *kp, fw.ctypes.data, fk.ctypes.data, 0,
                             self.nthreads)
            return fk
        fk = np.ascontiguousarray(data, np.complex64)
        lib.cpu_deconv_f(self.dim, ms, nf, *kp, fw.ctypes.data, fk.ctypes.data, 1, self.nthreads)
        if self.isign > 0:
            fw = scipy.fft.ifftn(fw, norm="forward", workers=self.nthreads, overwrite_x=True)
        else:
            fw = scipy.fft.fftn(fw, workers=self.nthreads, overwrite_x=True)
        fw = np.ascontiguousarray(fw, np.complex64)
        c = np.empty(self.M, np.complex64)
        lib.cpu_interp_f(self.dim, self.M, *self._ptrs(), c.ctypes.data, nf, self.ns, self.nc,
                         self.coef.ctypes.data, fw.ctypes.data, self.nthreads)
        return c
