"""The pruned 3-D line-transform kernels (finufft_amd/csrc/fft3.hip), with
the pass descriptors built by the library's own f3_passes, compiled from
their source for the host with AddressSanitizer and run one OS thread per
lane (tests/emu/hip_shim.hpp), against numpy's FFT.

Type-1 direction (grid -> modes): the kept modes of the full 3-D transform
of the grid, divided by the kernel coefficients, in either mode order.
Type-2 direction (modes -> grid): the full transform of the zero-padded,
amplified modes. Every line length 16..1024 and every workgroup shape
(64, 32, 16, 8 lines in fp32; 64..8 in fp64) is covered, odd and even mode
counts, both signs, several vectors. ASan bounds every global and LDS access
(the LDS block is exactly the launch's dynamic size); unwritten outputs
would show as NaN. The GPU runs the same kernels in tests/test_gpu_fft3.py.
"""

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
    """fft3.hpp's pass descriptors and fft3.hip's device code and host-pure
    helpers (up to its HIP launches), with the dynamic LDS array bound to the
    emulator's block."""
    hpp = (CSRC / "fft3.hpp").read_text()
    hip = (CSRC / "fft3.hip").read_text()
    h0 = hpp.index("constexpr int F3_THREADS")
    h1 = hpp.index("// True when every side")
    head = "namespace fa {\n" + hpp[h0:h1] + "}  // namespace fa\n"
    start = hip.index("namespace fa {")
    end = hip.index("// ---- HIP launches")
    body = hip[start:end] + "\n}  // namespace fa\n"
    body = body.replace("extern __shared__ __align__(16) unsigned char lds_raw[];",
                        "unsigned char *lds_raw = emu_lds;")
    assert body.count("lds_raw = emu_lds") == 2
    return head + body


@pytest.fixture(scope="module")
def emu_exe(tmp_path_factory):
    if not CLANG.exists() or shutil.which("ld") is None:
        pytest.skip("host clang++ not available")
    d = tmp_path_factory.mktemp("emu_fft3")
    (d / "fft3_kernels.inc").write_text(_kernel_source())
    exe = d / "fft3_emu"
    cmd = [str(CLANG), "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           f"-I{EMU}", f"-I{d}", str(EMU / "fft3_emu.cpp"), "-o", str(exe), "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    return exe


def _axis_modes(m, n, modeord):
    """Frequencies of the m modes in storage order, their cells in n."""
    f = np.arange(-(m // 2), m - m // 2)
    if modeord == 1:
        f = np.concatenate([f[f >= 0], f[f < 0]])
    return f, f % n


def _reference(x, ker, nf, ms, modeord, sign, to_modes):
    """numpy: the grid (nvec, nf2, nf1, nf0) -> modes (nvec, m2, m1, m0) or
    the reverse."""
    axes = [_axis_modes(ms[d], nf[d], modeord) for d in range(3)]
    kp = (ker[2][np.abs(axes[2][0])][:, None, None] * ker[1][np.abs(axes[1][0])][None, :, None]
          * ker[0][np.abs(axes[0][0])][None, None, :])
    cells = np.ix_(axes[2][1], axes[1][1], axes[0][1])
    tot = nf[0] * nf[1] * nf[2]
    if to_modes:
        F = np.fft.fftn(x, axes=(1, 2, 3)) if sign < 0 else np.fft.ifftn(x, axes=(1, 2, 3)) * tot
        return np.stack([F[t][cells] / kp for t in range(x.shape[0])])
    out = []
    for t in range(x.shape[0]):
        P = np.zeros((nf[2], nf[1], nf[0]), complex)
        P[cells] = x[t] / kp
        out.append(np.fft.fftn(P) if sign < 0 else np.fft.ifftn(P) * tot)
    return np.stack(out)


CASES = [
    # (precision, nf0, nf1, nf2, m0, m1, m2, modeord, sign, to_modes, nvec)
    ("f", 32, 16, 64, 15, 8, 30, 0, 1, 1, 1),      # G 64 everywhere
    ("f", 16, 64, 32, 9, 31, 16, 1, -1, 0, 2),     # two vectors
    ("f", 16, 16, 512, 5, 6, 256, 0, -1, 1, 1),    # z lines of 512: G 16, padded layout
    ("f", 16, 256, 16, 7, 100, 8, 1, 1, 0, 1),     # y lines of 256: G 32
    ("f", 1024, 16, 16, 500, 7, 8, 1, -1, 0, 1),   # x rows of 1024: G 8
    ("f", 1024, 16, 16, 333, 9, 5, 0, 1, 1, 1),
    ("d", 32, 32, 16, 17, 12, 7, 0, -1, 1, 1),     # fp64: G 64
    ("d", 16, 16, 128, 8, 9, 64, 1, 1, 0, 1),      # fp64 z lines of 128: G 32
    ("d", 16, 512, 16, 8, 201, 9, 0, 1, 1, 1),     # fp64 y lines of 512: G 8
    ("d", 256, 16, 16, 127, 5, 4, 1, -1, 1, 1),    # fp64 x rows of 256: G 16
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(map(str, c)))
def test_fft3_passes_vs_numpy(emu_exe, tmp_path, case):
    prec, nf0, nf1, nf2, m0, m1, m2, modeord, sign, to_modes, nvec = case
    nf, ms = (nf0, nf1, nf2), (m0, m1, m2)
    r = np.random.default_rng(11)
    shape = (nvec,) + ((nf2, nf1, nf0) if to_modes else (m2, m1, m0))
    x = r.standard_normal(shape) + 1j * r.standard_normal(shape)
    ker = [r.uniform(0.5, 1.5, n // 2 + 1) for n in nf]
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    np.concatenate([np.stack([x.real.ravel(), x.imag.ravel()], 1).ravel()] + ker).tofile(inp)
    res = subprocess.run([str(emu_exe), prec, *map(str, (nf0, nf1, nf2, m0, m1, m2, modeord,
                                                         sign, to_modes, nvec)),
                          str(inp), str(out)], capture_output=True, text=True, timeout=900)
    assert res.returncode == 0, (res.returncode, res.stderr[-4000:])
    v = np.fromfile(out, dtype=np.float64)
    got = (v[0::2] + 1j * v[1::2]).reshape((nvec,) + ((m2, m1, m0) if to_modes else (nf2, nf1, nf0)))
    assert np.isfinite(got).all(), "outputs left unwritten"
    ref = _reference(x, ker, nf, ms, modeord, sign, to_modes)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(case, f"relerr {err:.2e}")
    assert err < (2e-6 if prec == "f" else 1e-13)


def test_extraction_covers_both_kernels():
    src = _kernel_source()
    assert "fft3_cols_kernel(" in src and "fft3_rows_kernel(" in src and "f3_passes(" in src
    assert "<<<" not in src  # no host launch code
