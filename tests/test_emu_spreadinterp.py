"""The spreading and interpolation kernels (finufft_amd/csrc/spreadinterp.hip),
compiled from their own source for the host with AddressSanitizer and run
one OS thread per lane (tests/emu/hip_shim.hpp), against a direct FP64
spread / interpolation with the exponential of semicircle kernel in numpy.

This is the reference's spread-only / interp-only known-answer check
(test/cuda/cufinufft1dspreadinterponly_test.cu:68-152) restated for the
kernel source itself: point-driven spreading (method 1), LDS subproblem
spreading with an atomic flush, whole-bin parity-colored spreading with a
plain flush, the 3-D column-window spreader (colored bins, dense-bin
subproblems and their mix), LDS interpolation by bins and by subproblems,
the batch path's multi-vector interpolation, point-driven interpolation
and its coordinate derivatives (the autograd kernel), in 1-3 dimensions and
both precisions, with uniform and clustered points, bins that do not divide
the grid, tiles that wrap twice, and several vectors per launch. ASan checks
every global and LDS access. The GPU tests check the same kernels on the
device (tests/test_gpu_parity.py, tests/test_gpu_batch.py).
"""

import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "finufft_amd" / "csrc"
EMU = Path(__file__).resolve().parent / "emu"
CLANG = Path("/opt/rocm/llvm/bin/clang++")


def _strip(text):
    keep = [ln for ln in text.splitlines()
            if not ln.startswith("#include") and ln.strip() != "#pragma once"]
    out = "\n".join(keep)
    for a in ("__attribute__((address_space(1)))", "__attribute__((address_space(4)))"):
        out = out.replace(a, "")
    return out


def _kernel_source():
    """kernel_math.hpp, spreadinterp.hpp and the device part of
    spreadinterp.hip (up to its launchers, which use the HIP runtime); device
    address-space qualifiers dropped so that ASan sees every access."""
    km = _strip((CSRC / "kernel_math.hpp").read_text())
    sh = _strip((CSRC / "spreadinterp.hpp").read_text())
    hip = (CSRC / "spreadinterp.hip").read_text()
    start = hip.index("namespace fa {")
    end = hip.index("// ---------------------------------------------------------------- launchers")
    body = _strip(hip[start:end]) + "\n}  // namespace\n}  // namespace fa\n"
    n = body.count("extern __shared__ __align__(16) unsigned char lds_raw[];")
    assert n >= 4
    body = body.replace("extern __shared__ __align__(16) unsigned char lds_raw[];",
                        "unsigned char *lds_raw = emu_lds;")
    return km + "\n" + sh + "\n" + body


_BUILT = {}


def _exe(tmp_path_factory, real, dim, ns):
    if not CLANG.exists() or shutil.which("ld") is None:
        pytest.skip("host clang++ not available")
    key = (real, dim, ns)
    if key not in _BUILT:
        d = tmp_path_factory.mktemp(f"emu_si_{real}{dim}_{ns}")
        (d / "spreadinterp_kernels.inc").write_text(_kernel_source())
        exe = d / "spreadinterp_emu"
        cmd = [str(CLANG), "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer",
               "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
               "-DEMU_NO_FA_TYPES", f"-DFA_REAL={real}", f"-DFA_DIM={dim}", f"-DEMU_NS={ns}",
               f"-I{EMU}", f"-I{d}", str(EMU / "spreadinterp_emu.cpp"), "-o", str(exe),
               "-lpthread"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        assert res.returncode == 0, res.stderr[-4000:]
        _BUILT[key] = exe
    return _BUILT[key]


def _direct(x, nf, ns, beta):
    """Per point and dimension: first node s = ceil(r - ns/2) of the folded
    coordinate r in [0, nf) and the ns kernel values phi(s + a - r)."""
    r = np.mod(x / (2 * np.pi) + 0.5, 1.0) * nf
    s = np.ceil(r - ns / 2.0)
    a = s[:, None] + np.arange(ns)[None, :] - r[:, None]
    c = 4.0 / (ns * ns)
    z = c * a * a
    ker = np.where(z < 1, np.exp(beta * (np.sqrt(np.maximum(1 - z, 0)) - 1)), 0.0)
    cells = np.mod(s[:, None].astype(np.int64) + np.arange(ns)[None, :], nf)
    return cells, ker


def _weights(xs, nf, ns, beta):
    """Flat grid index (x fastest) and weight of every (point, stencil cell)."""
    dim = len(xs)
    per = [_direct(x, nf[d], ns, beta) for d, x in enumerate(xs)]
    M = len(xs[0])
    idx = np.zeros((M, 1), np.int64)
    w = np.ones((M, 1))
    stride = 1
    for d in range(dim):
        cells, ker = per[d]
        idx = (idx[:, :, None] + stride * cells[:, None, :]).reshape(M, -1)
        w = (w[:, :, None] * ker[:, None, :]).reshape(M, -1)
        stride *= nf[d]
    return idx, w


def _run(exe, tmp, var, nf, bs, M, maxsub, blk, clustered, seed, np_od=64):
    """np_od: points staged per batch by the output-driven spreader (od)."""
    dim = len(nf)
    nf3 = list(nf) + [1] * (3 - dim)
    bs3 = list(bs) + [1] * (3 - dim)
    out = tmp / f"{var}.bin"
    res = subprocess.run([str(exe), var, *map(str, nf3), *map(str, bs3), str(M), str(maxsub),
                          str(blk), str(clustered), str(seed), str(out)],
                         capture_output=True, text=True, timeout=900,
                         env=dict(os.environ, EMU_NP=str(np_od)))
    assert res.returncode == 0, (res.returncode, res.stderr[-4000:])
    v = np.fromfile(out, dtype=np.float64)
    nsub = int(v[0])
    p = 1
    xs = [v[p + d * M:p + (d + 1) * M] for d in range(dim)]
    p += dim * M
    c = v[p:p + 2 * M * blk]
    p += 2 * M * blk
    g = v[p:]
    c = (c[0::2] + 1j * c[1::2]).reshape(blk, M)
    fw = (g[0::2] + 1j * g[1::2]).reshape(blk, -1)
    return nsub, xs, c, fw


# (real, dim, ns, nf, bins, M, maxsub, blk, clustered)
GEOMS = [
    ("float", 1, 7, (200,), (50,), 3000, 1024, 1, 0),
    ("double", 1, 12, (100,), (20,), 1500, 300, 2, 1),
    ("float", 2, 7, (48, 40), (12, 10), 4000, 1024, 2, 1),
    ("double", 2, 5, (36, 30), (9, 5), 2500, 200, 1, 0),
    ("float", 3, 4, (24, 20, 16), (6, 5, 4), 3000, 256, 1, 1),
    ("double", 3, 7, (20, 18, 16), (10, 9, 8), 1500, 1024, 2, 0),
]


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: f"{g[0]}-{g[1]}d-ns{g[2]}")
@pytest.mark.parametrize("var", ["gm", "sub", "colored", "smv"])
def test_spread_kernels_vs_direct(tmp_path_factory, tmp_path, geom, var):
    real, dim, ns, nf, bs, M, maxsub, blk, clustered = geom
    if var == "colored" and not all(n % b == 0 and (n // b) % 2 == 0 and n // b >= 2
                                    and b >= ns - 1 for n, b in zip(nf, bs)):
        pytest.skip("bins not colorable (the library spreads such plans by subproblems)")
    exe = _exe(tmp_path_factory, real, dim, ns)
    nsub, xs, c, fw = _run(exe, tmp_path, var, nf, bs, M, maxsub, blk, clustered, 11)
    idx, w = _weights(xs, nf, ns, 2.30 * ns)
    ntot = int(np.prod(nf))
    for t in range(blk):
        ref = np.zeros(ntot, complex)
        np.add.at(ref, idx.ravel(), (w * c[t][:, None]).ravel())
        err = np.linalg.norm(fw[t] - ref) / np.linalg.norm(ref)
        print(var, geom, f"subproblems {nsub} relerr {err:.2e}")
        assert err < (2e-6 if real == "float" else 1e-13)


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: f"{g[0]}-{g[1]}d-ns{g[2]}")
@pytest.mark.parametrize("np_od", [16, 64, 300])
def test_output_driven_spread_vs_direct(tmp_path_factory, tmp_path, geom, np_od):
    """Output-driven spreading (gpu_method 3): subproblems of maxsub points
    staged np at a time (several batches per subproblem; np above the
    workgroup's 256 threads too), against the direct spread."""
    real, dim, ns, nf, bs, M, maxsub, blk, clustered = geom
    exe = _exe(tmp_path_factory, real, dim, ns)
    nsub, xs, c, fw = _run(exe, tmp_path, "od", nf, bs, M, maxsub, blk, clustered, 15,
                           np_od=np_od)
    idx, w = _weights(xs, nf, ns, 2.30 * ns)
    ntot = int(np.prod(nf))
    for t in range(blk):
        ref = np.zeros(ntot, complex)
        np.add.at(ref, idx.ravel(), (w * c[t][:, None]).ravel())
        err = np.linalg.norm(fw[t] - ref) / np.linalg.norm(ref)
        print("od", np_od, geom, f"subproblems {nsub} relerr {err:.2e}")
        assert err < (2e-6 if real == "float" else 1e-13)


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: f"{g[0]}-{g[1]}d-ns{g[2]}")
@pytest.mark.parametrize("var", ["igm", "bin", "binsub", "bbin", "bsub", "mv"])
def test_interp_kernels_vs_direct(tmp_path_factory, tmp_path, geom, var):
    real, dim, ns, nf, bs, M, maxsub, blk, clustered = geom
    exe = _exe(tmp_path_factory, real, dim, ns)
    nsub, xs, c, fw = _run(exe, tmp_path, var, nf, bs, M, maxsub, blk, clustered, 12)
    idx, w = _weights(xs, nf, ns, 2.30 * ns)
    for t in range(blk):
        ref = (fw[t][idx] * w).sum(axis=1)
        err = np.linalg.norm(c[t] - ref) / np.linalg.norm(ref)
        print(var, geom, f"relerr {err:.2e}")
        assert err < (2e-6 if real == "float" else 1e-13)


# fp32 3-D column-window spreader: (ns, nf, bins) with every sub-column side
# S = 4, 2, 1 (ns <= 5, <= 7, 8), bins of 8 x 8 first nodes in x, y.
COL_GEOMS = [
    (4, (32, 32, 24), (8, 8, 6)),
    (7, (32, 32, 24), (8, 8, 6)),
    (8, (32, 16, 32), (8, 8, 8)),
]


@pytest.mark.parametrize("geom", COL_GEOMS, ids=lambda g: f"ns{g[0]}")
@pytest.mark.parametrize("var,maxsub,clustered", [("col", 1024, 0), ("colsub", 100, 1),
                                                  ("colmix", 8, 1)])
def test_column_spreader_vs_direct(tmp_path_factory, tmp_path, geom, var, maxsub, clustered):
    """Whole bins in parity colors (col), every bin as subproblems (colsub),
    and the mix the library runs on clustered points: bins above maxsub
    points as subproblems, the rest colored (colmix)."""
    ns, nf, bs = geom
    exe = _exe(tmp_path_factory, "float", 3, ns)
    blk = 2 if var == "col" and ns == 4 else 1  # one multi-vector case (keeps the suite short)
    nsub, xs, c, fw = _run(exe, tmp_path, var, nf, bs, 2000, maxsub, blk, clustered, 14)
    if var == "colmix":
        assert nsub > 0  # the cluster's bins went to the subproblem kernel
    idx, w = _weights(xs, nf, ns, 2.30 * ns)
    for t in range(blk):
        ref = np.zeros(int(np.prod(nf)), complex)
        np.add.at(ref, idx.ravel(), (w * c[t][:, None]).ravel())
        err = np.linalg.norm(fw[t] - ref) / np.linalg.norm(ref)
        print(var, geom, f"dense subproblems {nsub} relerr {err:.2e}")
        assert err < 2e-6


@pytest.mark.parametrize("ns", [4, 7])
def test_column_spreader_many_points_per_wave(tmp_path_factory, tmp_path, ns):
    """Whole bins of about 470 points: every wave walks several passes of 64
    points, each as two staged halves of 32 (the other tests keep waves
    below one half)."""
    nf, bs = (32, 32, 24), (8, 8, 6)
    exe = _exe(tmp_path_factory, "float", 3, ns)
    nsub, xs, c, fw = _run(exe, tmp_path, "col", nf, bs, 30000, 1024, 1, 0, 15)
    idx, w = _weights(xs, nf, ns, 2.30 * ns)
    ref = np.zeros(int(np.prod(nf)), complex)
    np.add.at(ref, idx.ravel(), (w * c[0][:, None]).ravel())
    err = np.linalg.norm(fw[0] - ref) / np.linalg.norm(ref)
    print(ns, f"relerr {err:.2e}")
    assert err < 2e-6


@pytest.mark.parametrize("var,maxsub", [("bbin", 1024), ("bsub", 700), ("bsub", 64)])
def test_bank_ordered_interp_dense_bins(tmp_path_factory, tmp_path, var, maxsub):
    """Bank-ordered interpolation on bins of about 1900 points: full rounds of
    one point per bank class, then the remaining points strided; whole bins
    (bbin) and subproblems whose class bounds are clipped to the unit (bsub)."""
    nf, bs = (32, 32, 32), (16, 8, 8)
    exe = _exe(tmp_path_factory, "float", 3, 7)
    nsub, xs, c, fw = _run(exe, tmp_path, var, nf, bs, 60000, maxsub, 1, 0, 16)
    idx, w = _weights(xs, nf, 7, 2.30 * 7)
    ref = (fw[0][idx] * w).sum(axis=1)
    err = np.linalg.norm(c[0] - ref) / np.linalg.norm(ref)
    print(var, maxsub, f"subproblems {nsub} relerr {err:.2e}")
    assert np.isfinite(c[0]).all() and err < 2e-6


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: f"{g[0]}-{g[1]}d-ns{g[2]}")
def test_interp_grad_kernel_vs_direct(tmp_path_factory, tmp_path, geom):
    """Values and coordinate derivatives of the interpolation (the coordinate
    part of the autograd backward): dc/dx_d = sum fw * d/dx_d prod phi, with
    d phi(s + a - r)/dx = -phi'(s + a - r) nf / (2 pi)."""
    real, dim, ns, nf, bs, M, maxsub, blk, clustered = geom
    exe = _exe(tmp_path_factory, real, dim, ns)
    out = tmp_path / "grad.bin"
    nf3 = list(nf) + [1] * (3 - dim)
    bs3 = list(bs) + [1] * (3 - dim)
    res = subprocess.run([str(exe), "grad", *map(str, nf3), *map(str, bs3), str(M), str(maxsub),
                          str(blk), str(clustered), "15", str(out)],
                         capture_output=True, text=True, timeout=900)
    assert res.returncode == 0, (res.returncode, res.stderr[-4000:])
    v = np.fromfile(out, dtype=np.float64)
    p = 1
    xs = [v[p + d * M:p + (d + 1) * M] for d in range(dim)]
    p += dim * M
    c = v[p:p + 2 * M * blk]
    p += 2 * M * blk
    ng = int(np.prod(nf))
    g = v[p:p + 2 * ng * blk]
    p += 2 * ng * blk
    dcv = v[p:]
    c = (c[0::2] + 1j * c[1::2]).reshape(blk, M)
    fw = (g[0::2] + 1j * g[1::2]).reshape(blk, ng)
    dc = (dcv[0::2] + 1j * dcv[1::2]).reshape(dim, blk, M)
    beta, cc = 2.30 * ns, 4.0 / (ns * ns)
    per = []
    for d in range(dim):
        r = np.mod(xs[d] / (2 * np.pi) + 0.5, 1.0) * nf[d]
        s = np.ceil(r - ns / 2.0)
        a = s[:, None] + np.arange(ns)[None, :] - r[:, None]
        z = np.maximum(1 - cc * a * a, 0)
        phi = np.where(cc * a * a < 1, np.exp(beta * (np.sqrt(z) - 1)), 0.0)
        dphi = np.where(cc * a * a < 1, -phi * beta * cc * a / np.sqrt(np.where(z > 0, z, 1)), 0.0)
        dx = -dphi * nf[d] / (2 * np.pi)
        cells = np.mod(s[:, None].astype(np.int64) + np.arange(ns)[None, :], nf[d])
        per.append((cells, phi, dx))

    def outer(which):
        idx = np.zeros((M, 1), np.int64)
        w = np.ones((M, 1))
        stride = 1
        for d in range(dim):
            cells, phi, dx = per[d]
            k = dx if d == which else phi
            idx = (idx[:, :, None] + stride * cells[:, None, :]).reshape(M, -1)
            w = (w[:, :, None] * k[:, None, :]).reshape(M, -1)
            stride *= nf[d]
        return idx, w

    tol = 2e-6 if real == "float" else 1e-12
    for t in range(blk):
        idx, w = outer(-1)
        ref = (fw[t][idx] * w).sum(axis=1)
        assert np.linalg.norm(c[t] - ref) / np.linalg.norm(ref) < tol
        for d in range(dim):
            idx, w = outer(d)
            ref = (fw[t][idx] * w).sum(axis=1)
            err = np.linalg.norm(dc[d, t] - ref) / np.linalg.norm(ref)
            assert err < tol, (d, err)


def test_column_spreader_spread_only_grid(tmp_path_factory, tmp_path):
    """The geometry of a spread-only 3-D fp32 plan on a 48 x 40 x 44 mode grid
    (nf = the mode counts, x fastest: 44 x 40 x 48): bins of 8 x 8 x 8 whose
    x count does not divide the grid, so every bin runs as subproblems of the
    column spreader, on clustered points."""
    exe = _exe(tmp_path_factory, "float", 3, 7)
    nf, bs = (44, 40, 48), (8, 8, 8)
    nsub, xs, c, fw = _run(exe, tmp_path, "colsub", nf, bs, 2000, 1024, 1, 1, 16)
    idx, w = _weights(xs, nf, 7, 2.30 * 7)
    ref = np.zeros(int(np.prod(nf)), complex)
    np.add.at(ref, idx.ravel(), (w * c[0][:, None]).ravel())
    assert np.linalg.norm(fw[0] - ref) / np.linalg.norm(ref) < 2e-6


def test_bins_wider_than_half_the_grid(tmp_path_factory, tmp_path):
    """Bins of 16 on a grid of 18: the padded tile reaches past 2 nf, and the
    interpolation stages every tile cell (zero padding included), whose grid
    index wraps twice."""
    exe = _exe(tmp_path_factory, "float", 2, 7)
    nf, bs = (18, 18), (16, 16)
    idx = w = None
    for var in ("sub", "bin"):
        nsub, xs, c, fw = _run(exe, tmp_path, var, nf, bs, 800, 1024, 1, 0, 13)
        if idx is None:
            idx, w = _weights(xs, nf, 7, 2.30 * 7)
        if var == "sub":
            ref = np.zeros(int(np.prod(nf)), complex)
            np.add.at(ref, idx.ravel(), (w * c[0][:, None]).ravel())
            err = np.linalg.norm(fw[0] - ref) / np.linalg.norm(ref)
        else:
            ref = (fw[0][idx] * w).sum(axis=1)
            err = np.linalg.norm(c[0] - ref) / np.linalg.norm(ref)
        assert err < 2e-6, (var, err)
