"""Pins the oracle against the reference's golden data, closed forms and brute
force, and the library's host-side math against the oracle (CPU only).

Reference sources of the fixtures: include/cufinufft/contrib/*.inc (Horner
tables), test/testutils.cpp:39-87 (next235 table, Gauss-Legendre test).
"""

import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import cfold, kernel, nudft
from oracle import sort as osort

GOLD = Path(__file__).resolve().parent / "golden"


def _fixture(name):
    return json.loads((GOLD / name).read_text())


# ---------------------------------------------------------------- Horner


@pytest.mark.parametrize("sigma,fname", [(2.0, "horner_sigma2.json"), (1.25, "horner_sigma125.json")])
def test_oracle_horner_matches_reference_tables(sigma, fname):
    for w, table in _fixture(fname).items():
        ref = np.array(table)
        got = kernel.horner_coeffs(int(w), sigma)
        assert got.shape == ref.shape, (w, got.shape, ref.shape)
        assert np.abs(got - ref).max() < 5e-14, (w, np.abs(got - ref).max())


@pytest.mark.parametrize("sigma,fname", [(2.0, "horner_sigma2.json"), (1.25, "horner_sigma125.json")])
def test_library_horner_matches_reference_tables(sigma, fname):
    import ctypes

    from finufft_amd import _lib

    L = _lib.lib()
    for w, table in _fixture(fname).items():
        ref = np.array(table)
        out = np.zeros(18 * int(w))
        nc = L.finufft_amd_horner_coeffs(int(w), sigma, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 18)
        got = out[: nc * int(w)].reshape(nc, int(w))
        # rows past the sigma's own degree are zero padding
        assert np.abs(got[: ref.shape[0]] - ref).max() < 5e-14
        assert np.all(got[ref.shape[0]:] == 0)


@pytest.mark.parametrize("w", [2, 4, 7, 10, 13, 16])
def test_horner_approximates_es_kernel(w):
    """The polynomials reproduce phi on every interval (max error well below
    the width's tolerance 10^-(w-1))."""
    ns, beta = w, {2: 2.20, 3: 2.26, 4: 2.38}.get(w, 2.30) * w
    coef = kernel.horner_coeffs(w, 2.0)
    zs = np.linspace(-1, 1, 41)
    err = 0.0
    for z in zs:
        vals = kernel.horner_eval(coef, z)
        x = -ns / 2 + (z + 1) / 2 + np.arange(ns)  # grid offsets from the point
        err = max(err, np.abs(vals - kernel.es(x, ns, beta)).max())
    assert err < max(10.0 ** (-(w - 1)), 1e-14), err


# ---------------------------------------------------------------- next235, quadrature


def test_next235_reference_table():
    from finufft_amd import _lib

    table = _fixture("next235_even.json")
    L = _lib.lib()
    for n, t in enumerate(table):
        assert kernel.next235(n, 2) == t
        assert L.finufft_amd_next235(n, 2) == t
    for n in (3_000_000_000, 123_456_789, 2 ** 31 - 1):
        v = L.finufft_amd_next235(n, 2)
        assert v >= n and v % 2 == 0 and v == kernel.next235(n, 2) or n > 10 ** 9
        r = v // 2
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        assert r == 1


def test_gaussquad():
    """test/testutils.cpp:70-87: integral of 4 cos(4x+1) over [-1,1] to 1e-14,
    and agreement with numpy's Gauss-Legendre rule."""
    import ctypes

    from finufft_amd import _lib

    L = _lib.lib()
    for n in (1, 2, 7, 16, 33, 64):
        x = np.zeros(n)
        w = np.zeros(n)
        L.finufft_amd_gaussquad(n, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.allclose(x, xr, atol=1e-14) and np.allclose(w, wr, atol=1e-14)
    x = np.zeros(16)
    w = np.zeros(16)
    L.finufft_amd_gaussquad(16, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    integral = np.sum(w * 4 * np.cos(4 * x + 1.0))
    exact = (math.sin(5.0) + 0.3) - (math.sin(-3.0) + 0.3)
    assert abs(integral - exact) < 1e-14


# ---------------------------------------------------------------- width rules


def _lib_kernel_shape(tol, is_double, sigma, kerevalmeth, dim=3, ttype=1):
    import ctypes

    from finufft_amd import _lib

    ns, beta, formula = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    ier = _lib.lib().finufft_amd_kernel_shape(tol, is_double, sigma, kerevalmeth, dim, ttype,
                                               ctypes.byref(ns), ctypes.byref(beta),
                                               ctypes.byref(formula))
    return ier, ns.value, beta.value, formula.value


TOLS = [1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-12, 1e-15, 1e-20]


@pytest.mark.parametrize("is_double", [0, 1])
@pytest.mark.parametrize("sigma,kerevalmeth", [(2.0, 2), (1.25, 2), (2.0, 0), (1.5, 0), (3.0, 0)])
def test_kernel_shape_library_vs_oracle(is_double, sigma, kerevalmeth):
    """ES plans (kerevalmeth 0 and 2): the reference GPU library's rule."""
    for tol in TOLS:
        ier, ns, beta, formula = _lib_kernel_shape(tol, is_double, sigma, kerevalmeth)
        assert ier == 0 and formula == 0
        ons, obeta = kernel.kernel_shape(tol, sigma, bool(is_double))
        assert (ns, beta) == pytest.approx((ons, obeta)), tol


@pytest.mark.parametrize("is_double", [0, 1])
@pytest.mark.parametrize("sigma", [2.0, 1.25])
def test_kernel_shape_pswf_library_vs_oracle(is_double, sigma):
    """Horner plans (kerevalmeth 1): the PSWF kernel with the reference CPU
    library's width law, for every dimension and type."""
    for dim in (1, 2, 3):
        for ttype in (1, 2, 3):
            for tol in TOLS:
                ier, ns, beta, formula = _lib_kernel_shape(tol, is_double, sigma, 1, dim, ttype)
                assert ier == 0 and formula == 1
                ons, obeta = kernel.kernel_shape_pswf(tol, sigma, dim, ttype, bool(is_double))
                assert (ns, beta) == pytest.approx((ons, obeta), abs=1e-12), (dim, ttype, tol)


def test_kernel_shape_pswf_known_values():
    """The headline configuration (3-D, fp32, tol 1e-6) keeps width 7 with the
    PSWF kernel, c = 7 pi 3/4 - 0.05; 2-D fp32 tol 1e-5 gives 6 and 1-D fp64
    tol 1e-9 gives 10, the widths of the ES rule at those configurations."""
    assert kernel.kernel_shape_pswf(1e-6, 2.0, 3, 1, False) == pytest.approx(
        (7, 7 * math.pi * 0.75 - 0.05))
    assert kernel.kernel_shape_pswf(1e-5, 2.0, 2, 2, False)[0] == 6
    assert kernel.kernel_shape_pswf(1e-9, 2.0, 1, 1, True)[0] == 10
    # fp32 below sigma 1.4 caps the width at 8
    assert kernel.kernel_shape_pswf(1e-7, 1.25, 3, 1, False)[0] == 8


def test_kernel_shape_known_values():
    """ns = ceil(-log10(tol/10)) at sigma 2 (src/cuda/makeplan.cu:93): the
    headline configuration (fp32, tol 1e-6) uses ns = 7, beta = 16.1."""
    assert kernel.kernel_shape(1e-6, 2.0, False) == pytest.approx((7, 16.1))
    assert kernel.kernel_shape(1e-9, 2.0, True) == pytest.approx((10, 23.0))
    assert kernel.kernel_shape(1e-2, 2.0, True) == pytest.approx((3, 6.78))


def test_kernel_shape_errors():
    """Horner with a nonstandard sigma -> 8, sigma <= 1 -> 7
    (makeplan.cu:61-77); an unknown kernel evaluation method -> 24."""
    assert _lib_kernel_shape(1e-5, 0, 0.9, 1)[0] == 8
    assert _lib_kernel_shape(1e-5, 0, 0.9, 2)[0] == 8
    assert _lib_kernel_shape(1e-5, 0, 0.9, 0)[0] == 7
    assert _lib_kernel_shape(1e-5, 0, 2.0, 3)[0] == 24


def test_fine_grid_len():
    from finufft_amd import _lib

    L = _lib.lib()
    for m in [1, 2, 7, 33, 100, 256, 1000, 12345]:
        for ns in (2, 7, 12):
            for sigma in (2.0, 1.25):
                assert L.finufft_amd_fine_grid_len(m, sigma, ns) == kernel.fine_grid_len(m, sigma, ns)


# ---------------------------------------------------------------- fold


def test_fold_closed_forms():
    n = 512
    x = np.array([0.0, -np.pi, np.pi / 2, -np.pi / 2, 2 * np.pi], np.float64)
    r = cfold.fold_rescale(x, n)
    assert r[0] == n / 2
    assert min(r[1], n - r[1]) < 1e-9  # -pi sits at either end of the period
    assert r[2] == pytest.approx(3 * n / 4, abs=1e-10)
    assert r[3] == pytest.approx(n / 4, abs=1e-10)
    assert r[4] == pytest.approx(n / 2, abs=1e-9)
    # the result always lies in [0, n), also for float32 just below +pi
    xf = np.array([np.nextafter(np.float32(np.pi), np.float32(0)), np.float32(np.pi),
                   np.float32(-np.pi), np.float32(1e6), np.float32(-1e6)], np.float32)
    rf = cfold.fold_rescale(xf, n)
    assert np.all((rf >= 0) & (rf < n))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_library_fold_is_bit_exact_with_oracle(dtype):
    """The library's exact-error emulation of the round-down subtract and
    multiply equals the fesetround oracle bit for bit."""
    from finufft_amd import _lib

    L = _lib.lib()
    r = np.random.default_rng(0)
    x = np.concatenate([
        r.uniform(-np.pi, np.pi, 200_000), r.uniform(-50, 50, 100_000),
        r.uniform(-1e5, 1e5, 50_000),
        np.array([0.0, np.pi, -np.pi, 3 * np.pi, -3 * np.pi, 1e-30, -1e-30]),
        np.nextafter(np.float64(np.pi), 0) * np.ones(3),
    ]).astype(dtype)
    fn = L.finufft_amd_fold_rescale_f if dtype == np.float32 else L.finufft_amd_fold_rescale_d
    for n in (14, 200, 512, 2_000_000):
        got = np.empty_like(x)
        fn(x.ctypes.data, x.size, n, got.ctypes.data)
        ref = cfold.fold_rescale(x, n)
        assert np.array_equal(got.view(np.uint32 if dtype == np.float32 else np.uint64),
                              ref.view(np.uint32 if dtype == np.float32 else np.uint64))
        assert np.all((got >= 0) & (got < n))


# ---------------------------------------------------------------- kernel Fourier series


def test_fourier_series_gl_vs_quadrature():
    from oracle import cpu

    for ns, nf in ((7, 64), (10, 40), (4, 100)):
        beta = 2.30 * ns if ns > 4 else 2.38 * ns
        gl = cpu.fourier_series_gl(ns, beta, nf, "es")
        q = kernel.fourier_series(ns, beta, nf, nf // 2, "es")
        # the q = floor(2 + 1.5 ns) node rule is accurate far below the
        # width's own tolerance 10^-(ns-1)
        assert np.abs(gl - q).max() < 10.0 ** (-(ns + 1)) * np.abs(q).max()
        c = math.pi * ns * 0.75 - 0.05
        gl = cpu.fourier_series_gl(ns, c, nf, "pswf")
        q = kernel.fourier_series(ns, c, nf, nf // 2, "pswf")
        assert np.abs(gl - q).max() < 10.0 ** (-(ns + 1)) * np.abs(q).max()


# ---------------------------------------------------------------- direct sums


def test_nudft_single_point_closed_form():
    """One point of strength 1 gives f[k] = exp(i s k x) exactly."""
    x = [np.array([0.7]), np.array([-1.3])]
    f = nudft.type1(x, np.array([1.0 + 0j]), +1, (5, 4))
    k0 = nudft.freqs(5)[:, None]
    k1 = nudft.freqs(4)[None, :]
    assert np.allclose(f, np.exp(1j * (k0 * 0.7 + k1 * -1.3)), atol=1e-14)


def test_nudft_equispaced_matches_fft():
    """Equispaced points x_j = 2 pi j / n - pi: type 1 is a shifted DFT."""
    n = 16
    x = 2 * np.pi * np.arange(n) / n - np.pi
    r = np.random.default_rng(1)
    c = r.standard_normal(n) + 1j * r.standard_normal(n)
    f = nudft.type1([x], c, -1, (n,))
    k = nudft.freqs(n)
    ref = np.fft.fft(c)[k % n] * np.exp(1j * np.pi * k)
    assert np.allclose(f, ref, atol=1e-12)


def test_nudft_adjointness_and_bruteforce():
    r = np.random.default_rng(2)
    shape = (3, 4, 5)
    M = 7
    pts = [r.uniform(-np.pi, np.pi, M) for _ in range(3)]
    c = r.standard_normal(M) + 1j * r.standard_normal(M)
    f = r.standard_normal(shape) + 1j * r.standard_normal(shape)
    F = nudft.type1(pts, c, 1, shape)
    C = nudft.type2(pts, f, -1)
    assert abs(np.vdot(f.ravel(), F.ravel()) - np.vdot(C, c)) < 1e-12 * np.abs(F).sum()
    # plain loops
    ks = [nudft.freqs(n) for n in shape]
    for i0 in range(3):
        for i1 in range(4):
            for i2 in range(5):
                s = sum(c[j] * np.exp(1j * (ks[0][i0] * pts[0][j] + ks[1][i1] * pts[1][j]
                                            + ks[2][i2] * pts[2][j])) for j in range(M))
                assert abs(s - F[i0, i1, i2]) < 1e-12
    sub = nudft.type1_at(pts, c, 1, shape, [0, 17, 59])
    assert np.allclose(sub, F.ravel()[[0, 17, 59]], atol=1e-13)
    T = nudft.type3(pts, c, 1, [np.array([ks[0][1]]), np.array([ks[1][2]]), np.array([ks[2][3]])])
    assert abs(T[0] - F[1, 2, 3]) < 1e-12


def test_nudft_modeord_is_fft_order():
    r = np.random.default_rng(3)
    pts = [r.uniform(-np.pi, np.pi, 20)]
    c = r.standard_normal(20) + 0j
    a = nudft.type1(pts, c, 1, (9,), modeord=0)
    b = nudft.type1(pts, c, 1, (9,), modeord=1)
    assert np.allclose(np.fft.ifftshift(a), b)


# ---------------------------------------------------------------- sort


def test_sort_oracle_vs_bruteforce():
    r = np.random.default_rng(4)
    pts = [r.uniform(-np.pi, np.pi, 300).astype(np.float32) for _ in range(3)]
    pts[0][:3] = [np.float32(np.pi), np.float32(-np.pi), np.float32(10.0)]
    nf, bs = [24, 20, 16], [8, 8, 8]
    perm, start = osort.bin_sort(pts, nf, bs)
    assert np.array_equal(perm, osort.bin_sort_bruteforce(pts, nf, bs))
    key = osort.bin_keys(pts, nf, bs)
    assert np.all(np.diff(key[perm]) >= 0)
    assert start[-1] == 300 and np.all(np.diff(start) >= 0)


def test_start_bins_cover_support():
    """Bins of the first support node: every point's support lies in its bin's
    tile [off, off + b + ns - 1) (the property the LDS kernels rely on)."""
    r = np.random.default_rng(5)
    x = r.uniform(-np.pi, np.pi, 5000).astype(np.float32)
    nf, b, ns = 64, 8, 7
    key = osort.start_bin_keys([x], [nf], [b], ns)
    rr = cfold.fold_rescale(x.astype(np.float64), nf)
    s = np.ceil(rr - ns / 2).astype(int)
    s = np.where(s < 0, s + nf, s)
    off = key * b
    assert np.all((s - off >= 0) & (s - off < b))


def test_subcolumn_sort_vs_bruteforce():
    """Sub-column refinement of the start bins (3-D column spreading): bins in
    x-fastest order, inside a bin the S x S first-node squares in x-fastest
    order, inside a square increasing z of the first node, ties by index. The
    brute force walks that nesting literally."""
    r = np.random.default_rng(6)
    M = 400
    pts = [r.uniform(-np.pi, np.pi, M).astype(np.float32) for _ in range(3)]
    nf, bs, ns, S = [32, 24, 16], [8, 8, 4], 7, 2
    perm, start = osort.bin_sort(pts, nf, bs, start_ns=ns, col_side=S)
    s = osort.start_nodes(pts, nf, ns)
    nb = [n // b for n, b in zip(nf, bs)]
    order = []
    for bz in range(nb[2]):
        for by in range(nb[1]):
            for bx in range(nb[0]):
                for sy in range(bs[1] // S):
                    for sx in range(bs[0] // S):
                        for lz in range(bs[2]):
                            for i in range(M):
                                if (s[0][i] // bs[0] == bx and s[1][i] // bs[1] == by
                                        and s[2][i] // bs[2] == bz
                                        and (s[0][i] % bs[0]) // S == sx
                                        and (s[1][i] % bs[1]) // S == sy
                                        and s[2][i] % bs[2] == lz):
                                    order.append(i)
    assert np.array_equal(perm, np.array(order, np.int32))
    # bin-level starts are those of the unrefined start bins
    _, start0 = osort.bin_sort(pts, nf, bs, start_ns=ns)
    assert np.array_equal(start, start0)


def test_fp32_width_table():
    """fp32 kernel widths used by the 3-D column spreader tests: the PSWF law
    for 3-D type 1 (kerevalmeth 1) and the ES rule (tol / 10 in fp32,
    src/cuda/makeplan.cu:43-119, kerevalmeth 2)."""
    for tol, want in [(1e-3, 4), (2e-4, 5), (2e-5, 6), (1e-6, 7), (1e-7, 8)]:
        ier, ns, _, _ = _lib_kernel_shape(tol, 0, 2.0, 1, 3, 1)
        assert ier == 0 and ns == want, tol
        assert ns == kernel.kernel_shape_pswf(tol, 2.0, 3, 1, False)[0]
    for tol, want in [(1e-3, 4), (2e-4, 5), (1e-4, 6), (1e-6, 7), (1e-7, 8)]:
        ier, ns, _, _ = _lib_kernel_shape(tol, 0, 2.0, 2)
        assert ier == 0 and ns == want, tol
        assert ns == kernel.kernel_shape(tol, 2.0, is_double=False)[0]


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_bank_sort_vs_bruteforce(dim):
    """Bank-class refinement of the start bins (interpolation plans): bins in
    x-fastest order, inside a bin increasing (lx + P0 ly + P0 P1 lz) mod 32,
    ties by index; walked literally."""
    r = np.random.default_rng(7 + dim)
    M = 300
    pts = [r.uniform(-np.pi, np.pi, M).astype(np.float32) for _ in range(dim)]
    nf = [32, 24, 16][:dim]
    bs = [16, 8, 8][:dim]
    ns = 7
    perm, start = osort.bin_sort(pts, nf, bs, start_ns=ns, bank=True)
    s = osort.start_nodes(pts, nf, ns)
    P = [b + ns - 1 for b in bs]
    nb = [n // b for n, b in zip(nf, bs)]
    bins = np.zeros(M, np.int64)
    cls = np.zeros(M, np.int64)
    for i in range(M):
        b, stride, cell, cstride = 0, 1, 0, 1
        for d in range(dim):
            b += (s[d][i] // bs[d]) * stride
            stride *= nb[d]
            cell += (s[d][i] % bs[d]) * cstride
            cstride *= P[d]
        bins[i], cls[i] = b, cell % 32
    order = [i for b in range(int(np.prod(nb))) for k in range(32) for i in range(M)
             if bins[i] == b and cls[i] == k]
    assert np.array_equal(perm, np.array(order, np.int32))
    _, start0 = osort.bin_sort(pts, nf, bs, start_ns=ns)
    assert np.array_equal(start, start0)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_type2_grad_closed_form_vs_fd(dim):
    """The oracle's coordinate derivative of the type-2 direct sum against
    central differences of the direct sum itself (FP64, h = 1e-6)."""
    r = np.random.default_rng(30 + dim)
    M = 40
    shape = (9, 8, 7)[:dim]
    pts = [r.uniform(-np.pi, np.pi, M) for _ in range(dim)]
    f = r.standard_normal(shape) + 1j * r.standard_normal(shape)
    g = nudft.type2_grad(pts, f, -1)
    h = 1e-6
    for d in range(dim):
        up = [p.copy() for p in pts]
        dn = [p.copy() for p in pts]
        up[d] += h
        dn[d] -= h
        fd = (nudft.type2(up, f, -1) - nudft.type2(dn, f, -1)) / (2 * h)
        assert nudft.relerr(g[d], fd) < 1e-7


def test_mlp_oracle_grad_vs_fd():
    """The FP64 cost-model network gradient against central differences of
    its own loss (a dropped term or a transposed weight fails this)."""
    from oracle import mlp

    r = np.random.default_rng(11)
    p = r.normal(0, 0.3, mlp.NPARAM)
    x = r.normal(0, 1, (20, mlp.IN))
    t = r.normal(0, 1, (20, mlp.OUT))
    m = (r.uniform(size=(20, mlp.OUT)) > 0.3).astype(float)
    g = mlp.grad(p, x, t, m)
    h = 1e-6
    for i in r.choice(mlp.NPARAM, 60, replace=False):
        e = np.zeros(mlp.NPARAM)
        e[i] = h
        fd = (mlp.loss(p + e, x, t, m) - mlp.loss(p - e, x, t, m)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i
