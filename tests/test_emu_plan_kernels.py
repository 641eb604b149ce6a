"""The deconvolution and amplification kernels (finufft_amd/csrc/plan_kernels.hip:
fine grid -> modes for type 1, modes -> zero-padded fine grid for type 2, and
the batch path's multi-problem amplification), compiled from their own source
for the host with AddressSanitizer (tests/emu/hip_shim.hpp), against numpy:
fk[k] = fw[w(k)] / prod_d ker_d[|k_d|] and its inverse with zeros outside
the modes, in both mode orders (src/cuda/execute.cu:24-89), 1-3 dimensions,
both precisions, odd and even mode counts, several vectors."""

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
    """The headers the kernels use and plan_kernels.hip up to the end of its
    anonymous namespace (its launchers use the HIP runtime)."""
    parts = [_strip((CSRC / h).read_text()) for h in
             ("host_math.hpp", "kernel_math.hpp", "spreadinterp.hpp", "plan_kernels.hpp")]
    hip = (CSRC / "plan_kernels.hip").read_text()
    start = hip.index("namespace fa {")
    anon = hip.index("namespace {", start)
    end = hip.index("\n}  // namespace\n", anon) + len("\n}  // namespace\n")
    parts.append(_strip(hip[start:end]) + "\n}  // namespace fa\n")
    return "\n".join(parts)


_BUILT = {}


def _exe(tmp_path_factory, real, dim):
    if not CLANG.exists() or shutil.which("ld") is None:
        pytest.skip("host clang++ not available")
    if (real, dim) not in _BUILT:
        d = tmp_path_factory.mktemp(f"emu_pk_{real}{dim}")
        (d / "plan_kernels_kernels.inc").write_text(_kernel_source())
        exe = d / "plan_kernels_emu"
        cmd = [str(CLANG), "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer",
               "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
               "-DEMU_NO_FA_TYPES", f"-DFA_REAL={real}", f"-DFA_DIM={dim}",
               f"-I{EMU}", f"-I{d}", str(EMU / "plan_kernels_emu.cpp"), "-o", str(exe),
               "-lpthread"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        assert res.returncode == 0, res.stderr[-4000:]
        _BUILT[(real, dim)] = exe
    return _BUILT[(real, dim)]


def _maps(ms, nf, modeord):
    """Per axis: the mode index of every frequency, its grid cell and |freq|."""
    out = []
    for m, n in zip(ms, nf):
        fr = np.arange(-(m // 2), m - m // 2)
        k = fr + m // 2 if modeord == 0 else np.where(fr >= 0, fr, m + fr)
        out.append((k, np.mod(fr, n), np.abs(fr)))
    return out


CASES = [
    ("float", (21,), (48,), 0, 2),
    ("double", (20,), (40,), 1, 1),
    ("float", (13, 10), (30, 24), 1, 2),
    ("double", (16, 9), (32, 20), 0, 1),
    ("float", (9, 8, 7), (20, 16, 14), 0, 1),
    ("double", (6, 11, 8), (12, 24, 16), 1, 2),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{len(c[1])}d-modeord{c[3]}")
def test_deconvolve_and_amplify_vs_numpy(tmp_path_factory, tmp_path, case):
    real, ms, nf, modeord, blk = case
    dim = len(ms)
    exe = _exe(tmp_path_factory, real, dim)
    ms3 = list(ms) + [1] * (3 - dim)
    nf3 = list(nf) + [1] * (3 - dim)
    out = tmp_path / "o.bin"
    res = subprocess.run([str(exe), *map(str, ms3), *map(str, nf3), str(modeord), str(blk), "5",
                          str(out)], capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, (res.returncode, res.stderr[-3000:])
    v = np.fromfile(out, dtype=np.float64)
    p = 0
    ker = []
    for n in nf:
        ker.append(v[p:p + n // 2 + 1])
        p += n // 2 + 1
    nm, ng = int(np.prod(ms)), int(np.prod(nf))

    def take(n):
        nonlocal p
        a = v[p:p + 2 * n * blk]
        p += 2 * n * blk
        return a[0::2] + 1j * a[1::2]

    # arrays are x fastest: numpy axes (z, y, x) = reversed dimension order
    fw_in = take(ng).reshape((blk,) + tuple(nf[::-1]))
    fk_out = take(nm).reshape((blk,) + tuple(ms[::-1]))
    fk_in = take(nm).reshape((blk,) + tuple(ms[::-1]))
    fw_out = take(ng).reshape((blk,) + tuple(nf[::-1]))
    maps = _maps(ms, nf, modeord)[::-1]  # numpy axis order
    kers = ker[::-1]
    kidx = np.ix_(*[m[0] for m in maps])
    cidx = np.ix_(*[m[1] for m in maps])
    kv = np.ones([len(m[0]) for m in maps])
    for ax, (m, k) in enumerate(zip(maps, kers)):
        shape = [1] * dim
        shape[ax] = -1
        kv = kv * k[m[2]].reshape(shape)
    tol = 1e-6 if real == "float" else 1e-14
    for t in range(blk):
        ref_fk = np.empty(ms[::-1], complex)
        ref_fk[kidx] = fw_in[t][cidx] / kv
        assert np.allclose(fk_out[t], ref_fk, rtol=tol, atol=0), "deconvolution"
        ref_fw = np.zeros(nf[::-1], complex)
        ref_fw[cidx] = fk_in[t][kidx] / kv
        assert np.allclose(fw_out[t], ref_fw, rtol=tol, atol=0), "amplification"
