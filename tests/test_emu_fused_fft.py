"""The batch path's fused column / row transform kernels, compiled from their
own source (finufft_amd/csrc/fused_fft.hip) for the host with
AddressSanitizer and run one OS thread per lane (tests/emu/hip_shim.hpp),
against numpy's FFT of the amplified, zero-padded modes.

This pins the kernels' index arithmetic and the mixed-radix (4, 2, 3, 5)
Stockham passes on the CPU, for grid lengths with every radix, odd and even
mode counts, both mode orders and both exponent signs; ASan checks every
global and LDS access (the LDS block is exactly the launch's dynamic size).
The GPU tests check the same kernels on the device (tests/test_gpu_batch.py).
"""

import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "finufft_amd" / "csrc"
EMU = Path(__file__).resolve().parent / "emu"
CLANG = Path("/opt/rocm/llvm/bin/clang++")


def _kernel_source():
    """fused_fft.hpp's declarations and fused_fft.hip's device code (up to the
    end of its anonymous namespace; the launchers after it use the HIP
    runtime), with the dynamic LDS array bound to the emulator's block."""
    hpp = (CSRC / "fused_fft.hpp").read_text()
    hip = (CSRC / "fused_fft.hip").read_text()
    strip = lambda s: "\n".join(ln for ln in s.splitlines()
                                if not ln.startswith("#include") and ln != "#pragma once")
    start = hip.index("namespace fa {")
    anon = hip.index("namespace {", start)
    end = hip.index("\n}  // namespace\n", anon) + len("\n}  // namespace\n")
    body = hip[start:end] + "\n}  // namespace fa\n"
    body = body.replace("extern __shared__ __align__(16) unsigned char lds_raw[];",
                        "unsigned char *lds_raw = emu_lds;")
    assert "lds_raw = emu_lds" in body
    return strip(hpp) + "\n" + strip(body)


@pytest.fixture(scope="module")
def emu_exe(tmp_path_factory):
    if not CLANG.exists() or shutil.which("ld") is None:
        pytest.skip("host clang++ not available")
    d = tmp_path_factory.mktemp("emu_fft")
    (d / "fused_fft_kernels.inc").write_text(_kernel_source())
    exe = d / "fused_fft_emu"
    cmd = [str(CLANG), "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           f"-I{EMU}", f"-I{d}", str(EMU / "fused_fft_emu.cpp"), "-o", str(exe), "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    return exe


def _expected(fk, rk0, rk1, nf0, nf1, m0, m1, modeord, sign):
    """sum over modes of fk * rker1[|fy|] rker0[|fx|] exp(sign 2 pi i (fx gx /
    nf0 + fy gy / nf1)), on the nf1 x nf0 grid (x fastest)."""
    fx = np.arange(-(m0 // 2), m0 - m0 // 2)
    fy = np.arange(-(m1 // 2), m1 - m1 // 2)
    kx = fx + m0 // 2 if modeord == 0 else np.where(fx >= 0, fx, m0 + fx)
    ky = fy + m1 // 2 if modeord == 0 else np.where(fy >= 0, fy, m1 + fy)
    P = np.zeros((nf1, nf0), complex)
    amp = fk[np.ix_(ky, kx)] * rk1[np.abs(fy)][:, None] * rk0[np.abs(fx)][None, :]
    P[np.ix_(fy % nf1, fx % nf0)] = amp
    return np.fft.fft2(P) if sign < 0 else np.fft.ifft2(P) * (nf0 * nf1)


CASES = [
    # (precision, nf0, nf1, m0, m1, modeord, sign)
    ("f", 64, 48, 30, 21, 0, 1),     # radices 4 4 4 / 4 4 3
    ("f", 60, 90, 27, 44, 1, -1),    # 4 3 5 / 2 3 3 5
    ("f", 100, 32, 50, 15, 0, -1),   # 4 5 5 / 4 4 2
    ("f", 16, 250, 7, 120, 1, 1),    # long columns: one column per workgroup
    ("d", 40, 36, 19, 18, 0, 1),     # fp64: smaller LDS budget
    ("d", 96, 50, 45, 25, 1, -1),
    ("f", 12, 12, 5, 6, 0, 1),       # tiny: workgroups far wider than the grid
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c)))
def test_fused_transforms_vs_numpy(emu_exe, tmp_path, case):
    prec, nf0, nf1, m0, m1, modeord, sign = case
    out = tmp_path / "out.bin"
    res = subprocess.run([str(emu_exe), prec, str(nf0), str(nf1), str(m0), str(m1),
                          str(modeord), str(sign), "7", str(out)],
                         capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, (res.returncode, res.stderr[-4000:])
    v = np.fromfile(out, dtype=np.float64)
    nm, n0, n1 = m0 * m1, nf0 // 2 + 1, nf1 // 2 + 1
    fk = (v[0:2 * nm:2] + 1j * v[1:2 * nm:2]).reshape(m1, m0)
    rk0 = v[2 * nm:2 * nm + n0]
    rk1 = v[2 * nm + n0:2 * nm + n0 + n1]
    g = v[2 * nm + n0 + n1:]
    fw = (g[0::2] + 1j * g[1::2]).reshape(nf1, nf0)
    assert np.isfinite(fw).all(), "grid cells left unwritten"
    ref = _expected(fk, rk0, rk1, nf0, nf1, m0, m1, modeord, sign)
    err = np.linalg.norm(fw - ref) / np.linalg.norm(ref)
    print(case, f"relerr {err:.2e}")
    assert err < (2e-6 if prec == "f" else 1e-13)


def test_extraction_covers_both_kernels():
    src = _kernel_source()
    assert re.search(r"fused_cols_kernel\(", src) and re.search(r"fused_rows_kernel\(", src)
    assert "<<<" not in src  # no host launch code
