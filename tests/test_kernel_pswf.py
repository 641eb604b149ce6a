"""The PSWF spreading kernel (the reference CPU library's default kernel,
include/finufft_common/pswf.h, src/common/kernel.cpp:41-150) and its Horner
polynomials, on the CPU.

Pins, for the library's psi_0^c (finufft_amd_pswf) and the oracle's:
* the reference's own PSWF0, compiled from src/common/pswf.cpp into
  oracle/_ref (skipped where that build is absent);
* the defining property of psi_0^c: it is an eigenfunction of the finite
  Fourier transform, int_{-1}^{1} exp(i c x t) psi(t) dt = lambda psi(x);
* scipy.special.pro_ang1 (accurate for moderate c) and psi(0) = 1, evenness.
The width law and shape parameter are pinned against the reference's
theoretical_kernel_ns / clamp_kernel_ns / set_kernel_shape_given_ns built the
same way, and the Horner polynomials against the PSWF itself and against the
reference CPU library's own per-interval fit (poly_fit).
"""

import ctypes
import math

import numpy as np
import pytest

from oracle import kernel, ref

C_VALUES = [3.72, 10.0, 16.443361431346414, 23.5, 30.0, 37.65]
needs_ref = pytest.mark.skipif(not ref.available(), reason="reference kernel build absent")


def _lib_pswf(c, z):
    from finufft_amd import _lib

    z = np.ascontiguousarray(z, np.float64)
    out = np.empty_like(z)
    _lib.lib().finufft_amd_pswf(c, z.ctypes.data, z.size, out.ctypes.data)
    return out


@needs_ref
@pytest.mark.parametrize("c", C_VALUES)
def test_pswf_matches_reference_build(c):
    z = np.linspace(-1.0, 1.0, 801)
    r = ref.pswf(c, z)
    assert np.abs(_lib_pswf(c, z) - r).max() < 1e-14
    assert np.abs(kernel.pswf(c, z) - r).max() < 1e-14


@pytest.mark.parametrize("c", C_VALUES)
def test_pswf_is_finite_fourier_eigenfunction(c):
    """int exp(i c x t) psi(t) dt / psi(x) is the same constant for every x
    (Gauss-Legendre quadrature of the smooth integrand to machine
    precision); a wrong matrix entry or basis normalisation breaks it."""
    t, w = np.polynomial.legendre.leggauss(200)
    x = np.linspace(-0.95, 0.95, 39)
    for fn in (_lib_pswf, kernel.pswf):
        F = np.exp(1j * c * np.outer(x, t)) @ (w * fn(c, t))
        lam = F[19] / fn(c, np.array([0.0]))[0]  # x = 0
        assert abs(lam.imag) < 1e-14 * abs(lam)
        assert np.abs(F - lam * fn(c, x)).max() < 1e-13 * abs(lam)


@pytest.mark.parametrize("c", [3.72, 10.0, 16.443361431346414])
def test_pswf_matches_scipy(c):
    import scipy.special as sp

    z = np.linspace(-0.99, 0.99, 45)
    s = np.array([sp.pro_ang1(0, 0, c, v)[0] for v in z]) / sp.pro_ang1(0, 0, c, 0.0)[0]
    got = _lib_pswf(c, z)
    assert np.abs(got - s).max() < 1e-12
    assert got[22] == pytest.approx(1.0, abs=1e-15)  # z = 0
    assert np.array_equal(_lib_pswf(c, -z), got)  # even: P_2k(-z) = P_2k(z) exactly
    assert _lib_pswf(c, np.array([1.0000001, -2.0])).tolist() == [0.0, 0.0]


@needs_ref
@pytest.mark.parametrize("is_float", [True, False])
@pytest.mark.parametrize("sigma", [2.0, 1.25])
def test_width_law_matches_reference_build(is_float, sigma):
    for dim in (1, 2, 3):
        for ttype in (1, 2, 3):
            for tol in [1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-11, 1e-13]:
                eps = float(np.float32(tol)) if is_float else tol
                eps = max(eps, float(np.finfo(np.float32 if is_float else np.float64).eps))
                ns, beta = kernel.kernel_shape_pswf(tol, sigma, dim, ttype, not is_float)
                assert ns == ref.kernel_ns(eps, dim, ttype, sigma, is_float), (dim, ttype, tol)
                assert beta == pytest.approx(ref.beta(ns, sigma, 8), abs=1e-13)


def _lib_horner_pswf(ns, beta):
    from finufft_amd import _lib

    out = np.zeros(18 * ns)
    nc = _lib.lib().finufft_amd_horner_coeffs_pswf(
        ns, beta, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 18)
    return out[: nc * ns].reshape(nc, ns)


@pytest.mark.parametrize("ns", [2, 3, 4, 5, 6, 7, 8, 10, 12, 16])
@pytest.mark.parametrize("sigma", [2.0, 1.25])
def test_horner_pswf_approximates_kernel(ns, sigma):
    """Each interval's polynomial reproduces psi inside the support to well
    below the tolerance this width serves (the width law's own prefactor),
    and the library's fit equals the oracle's."""
    beta = math.pi * ns * (1 - 1 / (2 * sigma)) - 0.05
    C = _lib_horner_pswf(ns, beta)
    assert C.shape[0] == ns + 2 + (ns <= 7) - (ns == 2)
    t = np.linspace(-1, 1, 301)
    worst = 0.0
    for i in range(ns):
        z = -1 + (2 * i + 1) / ns + t / ns
        inside = np.abs(z) < 1
        p = np.polynomial.polynomial.polyval(t, C[:, i])
        worst = max(worst, np.abs(p - kernel.pswf(beta, z))[inside].max())
    served = 0.18 * math.exp(-(ns - 1) * math.pi * math.sqrt(1 - 1 / sigma))
    assert worst < max(0.15 * served, 2e-15), (worst, served)
    Co = kernel.horner_coeffs_pswf(ns, beta, C.shape[0])
    assert np.abs(C - Co).max() < 1e-9 * max(1.0, np.abs(Co).max())


@needs_ref
@pytest.mark.parametrize("ns", [4, 7, 8, 12])
def test_horner_pswf_agrees_with_reference_fit(ns):
    """The reference CPU library interpolates each interval at min(19, ns+3)
    Chebyshev points; its polynomial and the library's agree on the interval
    to the accuracy of either fit."""
    beta = math.pi * ns * 0.75 - 0.05
    C = _lib_horner_pswf(ns, beta)
    n = min(19, ns + 3)
    t = np.linspace(-1, 1, 201)
    served = 0.18 * math.exp(-(ns - 1) * math.pi * math.sqrt(0.5))
    for j in range(ns):
        rc = ref.pswf_panel_coeffs(beta, ns, j, n)
        diff = np.polyval(rc, t) - np.polynomial.polynomial.polyval(t, C[:, j])
        assert np.abs(diff).max() < 0.05 * served, (j, np.abs(diff).max())
