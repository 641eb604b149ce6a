"""Oracle of the plan-time kernel math.

* ``es(x, ns, beta)``: ES kernel exp(beta (sqrt(1 - (2x/ns)^2) - 1)), |x| < ns/2
  (include/cufinufft/spreadinterp.hpp:29-39).
* ``kernel_shape``: width and beta rules of src/cuda/makeplan.cu:43-119 (the
  sigma = 2 width ceil(-log10(eps/10)) uses eps/10 formed in the plan
  precision and log10 in double).
* ``horner_coeffs``: the collocation least-squares fit of
  devel/ker_ppval_coeff_mat.m with the degree rule of get_degree_and_beta.m,
  solved here with numpy's SVD least squares (the library uses Householder
  QR in extended precision: two independent solvers of the same problem).
* ``next235``: brute-force smallest 2,3,5-smooth even number (the reference
  table test/testutils.cpp:39-44 pins it).
* ``fine_grid_len``: src/cuda/fft.cu:166-179.
* ``fourier_series``: the kernel's continuous Fourier transform at integer
  frequencies by adaptive scipy quadrature (independent of the library's
  Gauss-Legendre rule, src/cuda/fft.cu:29-49 and :185-202).
* ``pswf``: the prolate spheroidal wave function psi_0^c normalised to 1 at
  0, the reference CPU library's kernel (include/finufft_common/pswf.h,
  src/common/kernel.cpp:41-57): lowest eigenvector of the prolate operator
  in the orthonormal even Legendre basis, by LAPACK's tridiagonal eigensolver
  (scipy.linalg.eigh_tridiagonal), summed with numpy's Legendre series.
* ``kernel_shape_pswf``: width law and shape parameter of the reference CPU
  plan (theoretical_kernel_ns and clamp_kernel_ns, src/common/kernel.cpp:60-96
  and include/finufft_common/kernel.h; set_kernel_shape_given_ns kerformula
  8, kernel.cpp:98-150).
* ``horner_coeffs_pswf``: per-interval least-squares polynomial fit of the
  PSWF at Chebyshev points.

Test infrastructure only (see oracle/__init__.py).
"""

from __future__ import annotations

import math

import numpy as np


def es(x, ns, beta):
    z = 2.0 * np.asarray(x, np.float64) / ns
    with np.errstate(invalid="ignore"):
        v = np.exp(beta * (np.sqrt(1.0 - z * z) - 1.0))
    return np.where(np.abs(z) < 1.0, v, 0.0)


def kernel_shape(tol, upsampfac=2.0, is_double=False):
    """(ns, beta) as chosen by makeplan."""
    if is_double:
        eps = max(float(tol), float(np.finfo(np.float64).eps))
        eps10 = eps / 10.0
        max_ns = 16
    else:
        eps = max(np.float32(tol), np.finfo(np.float32).eps)
        eps10 = float(np.float32(eps) / np.float32(10.0))
        eps = float(eps)
        max_ns = 12
    if upsampfac == 2.0:
        ns = math.ceil(-math.log10(eps10))
    else:
        ns = math.ceil(-math.log(eps) / (math.pi * math.sqrt(1.0 - 1.0 / upsampfac)))
    ns = min(max(2, ns), max_ns)
    bo = 2.30
    if ns == 2:
        bo = 2.20
    elif ns == 3:
        bo = 2.26
    elif ns == 4:
        bo = 2.38
    if upsampfac != 2.0:
        bo = 0.97 * math.pi * (1.0 - 1.0 / (2.0 * upsampfac))
    return ns, bo * ns


def horner_degree(ns, upsampfac):
    if upsampfac == 2.0:
        return ns + 1 + (ns <= 7) - (ns == 2)
    return math.ceil(0.7 * ns + 1.3)


def horner_coeffs(ns, upsampfac):
    """(degree+1, ns) coefficients, row k multiplies z^k."""
    bo = {2: 2.20, 3: 2.26, 4: 2.38}.get(ns, 2.30) if upsampfac == 2.0 else (
        0.97 * math.pi * (1.0 - 1.0 / (2.0 * upsampfac)))
    beta = bo * ns
    d = horner_degree(ns, upsampfac)
    fitd, m = 20, 7
    h = 1.0 / ns
    l = 2 * (np.arange(1, m + 1) - 0.5) / m - 1
    z = np.concatenate([l - 1j, 1 + 1j * l, l + 1j, -1 + 1j * l])
    V = z[:, None] ** np.arange(fitd + 1)[None, :]
    R = np.stack(
        [np.exp(beta * (np.sqrt(1 - (-1 + h * (2 * i - 1) + z * h) ** 2) - 1))
         for i in range(1, ns + 1)], axis=1)
    C, *_ = np.linalg.lstsq(V, R, rcond=None)
    return C[: d + 1].real


def pswf(c, z):
    """psi_0^c(z) / psi_0^c(0), 0 for |z| > 1."""
    from numpy.polynomial import legendre
    from scipy.linalg import eigh_tridiagonal

    n = int(c) + 40
    k = np.arange(0, 2 * n, 2).astype(np.float64)
    b = (k + 1) ** 2 / ((2 * k + 1) * (2 * k + 3))
    b[1:] += k[1:] ** 2 / ((2 * k[1:] + 1) * (2 * k[1:] - 1))
    d = k * (k + 1) + c * c * b
    kk = k[:-1]
    e = c * c * (kk + 1) * (kk + 2) / ((2 * kk + 3) * np.sqrt((2 * kk + 1) * (2 * kk + 5)))
    _, v = eigh_tridiagonal(d, e, select="i", select_range=(0, 0))
    coef = np.zeros(2 * n)
    coef[0::2] = v[:, 0] * np.sqrt(k + 0.5)
    z = np.asarray(z, np.float64)
    val = legendre.legval(np.clip(z, -1.0, 1.0), coef) / legendre.legval(0.0, coef)
    return np.where(np.abs(z) <= 1.0, val, 0.0)


def kernel_shape_pswf(tol, upsampfac, dim, ttype, is_double):
    """(ns, beta) of a Horner (kerevalmeth 1) plan: the PSWF kernel."""
    eps = max(float(tol), float(np.finfo(np.float64 if is_double else np.float32).eps))
    if not is_double:
        eps = float(np.float32(eps))
    tolfac = 0.18 * 1.4 ** (dim - 1) * (1.4 if ttype == 3 else 1.0)
    ns = math.ceil(math.log(tolfac / eps) / (math.pi * math.sqrt(1.0 - 1.0 / upsampfac)) + 1.0)
    ns = min(max(2, ns), 16 if is_double else 12)
    if not is_double and upsampfac < 1.4:
        ns = min(ns, 8)
    return ns, math.pi * ns * (1.0 - 1.0 / (2.0 * upsampfac)) - 0.05


def horner_coeffs_pswf(ns, beta, nc):
    """(nc, ns) coefficients, row k multiplies z^k on each interval."""
    t = np.cos(np.pi * (2 * np.arange(4 * nc) + 1) / (8 * nc))
    V = t[:, None] ** np.arange(nc)[None, :]
    R = np.stack([pswf(beta, -1 + (2 * i + 1) / ns + t / ns) for i in range(ns)], axis=1)
    C, *_ = np.linalg.lstsq(V, R, rcond=None)
    return C


def horner_eval(coef, z):
    """ker[i] at local offset z in [-1, 1] for each interval i."""
    out = np.zeros(coef.shape[1])
    for k in range(coef.shape[0] - 1, -1, -1):
        out = out * z + coef[k]
    return out


def next235(n, f=2):
    n = max(int(n), 1)
    f = max(int(f), 1)
    m = -(-n // f)
    while True:
        r = m
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        if r == 1:
            return m * f
        m += 1


def fine_grid_len(ms, upsampfac, ns):
    nf = max(math.ceil(upsampfac * ms), 2 * ns)
    return next235(nf, 2)


def fourier_series(ns, beta, nf, kmax, formula="es"):
    """phihat(k) = int phi(x) exp(-2 pi i k x / nf) dx over |x| < ns/2 (grid units),
    times (-1)^k, for k = 0..kmax; phi the ES or the PSWF kernel."""
    from scipy.integrate import quad

    if formula == "es":
        def phi(x):
            return float(es(x, ns, beta))
    else:
        def phi(x):
            return float(pswf(beta, 2.0 * x / ns))
    out = np.zeros(kmax + 1)
    for k in range(kmax + 1):
        w = 2 * math.pi * k / nf
        val, _ = quad(lambda x: phi(x) * math.cos(w * x), -ns / 2, ns / 2,
                      limit=200, epsabs=1e-14, epsrel=1e-13)
        out[k] = val * (-1) ** k
    return out
