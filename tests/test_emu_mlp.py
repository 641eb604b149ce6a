"""The cost model's network kernels (finufft_amd/csrc/mlp.hip: forward,
gradient of the masked squared error with atomic accumulation, Adam step),
compiled from their own source for the host with AddressSanitizer
(tests/emu/hip_shim.hpp), against the FP64 numpy restatement oracle/mlp.py
(itself pinned by finite differences)."""

import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import mlp as omlp

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "finufft_amd" / "csrc"
EMU = Path(__file__).resolve().parent / "emu"
CLANG = Path("/opt/rocm/llvm/bin/clang++")


def _kernel_source():
    """The network-size macros of include/finufft_amd.h and mlp.hip up to the
    end of its anonymous namespace (its launchers use the HIP runtime)."""
    hdr = (ROOT / "include" / "finufft_amd.h").read_text()
    i = hdr.index("#define FINUFFT_AMD_MLP_IN")
    j = hdr.index("FINUFFT_EXPORT int finufft_amd_mlp_forward")
    macros = hdr[i:j]
    assert re.search(r"#define FINUFFT_AMD_MLP_NPARAM", macros)
    hip = (CSRC / "mlp.hip").read_text()
    start = hip.index("namespace fa {")
    anon = hip.index("namespace {", start)
    end = hip.index("\n}  // namespace\n", anon) + len("\n}  // namespace\n")
    body = "\n".join(ln for ln in hip[start:end].splitlines() if not ln.startswith("#include"))
    return macros + "\n" + body + "\n}  // namespace fa\n"


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if not CLANG.exists() or shutil.which("ld") is None:
        pytest.skip("host clang++ not available")
    d = tmp_path_factory.mktemp("emu_mlp")
    (d / "mlp_kernels.inc").write_text(_kernel_source())
    out = d / "mlp_emu"
    cmd = [str(CLANG), "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           f"-I{EMU}", f"-I{d}", str(EMU / "mlp_emu.cpp"), "-o", str(out), "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    return out


@pytest.mark.parametrize("n", [1, 130, 300])
def test_network_kernels_vs_oracle(exe, tmp_path, n):
    out = tmp_path / "o.bin"
    res = subprocess.run([str(exe), str(n), "3", str(out)], capture_output=True, text=True,
                         timeout=600)
    assert res.returncode == 0, res.stderr[-3000:]
    v = np.fromfile(out, dtype=np.float64)
    P, I, O = omlp.NPARAM, omlp.IN, omlp.OUT
    sizes = [P, n * I, n * O, n * O, n * O, P, 1, P, P, P]
    parts, o = [], 0
    for s in sizes:
        parts.append(v[o:o + s])
        o += s
    p, x, t, m, y, g, loss, m1, m2, p2 = parts
    x, t, m, y = x.reshape(n, I), t.reshape(n, O), m.reshape(n, O), y.reshape(n, O)
    y_ref, _ = omlp.forward(p, x)
    assert np.max(np.abs(y - y_ref)) < 1e-4 * max(1.0, np.max(np.abs(y_ref)))
    g_ref = omlp.grad(p, x, t, m)
    assert np.linalg.norm(g - g_ref) < 1e-4 * np.linalg.norm(g_ref)
    l_ref = omlp.loss(p, x, t, m)
    assert abs(loss[0] - l_ref) < 1e-4 * max(1.0, l_ref)
    # Adam at step 3: m1' = b1 m1 + (1 - b1) g, m2' = b2 m2 + (1 - b2) g^2,
    # p' = p - lr (m1' / c1) / (sqrt(m2' / c2) + eps)
    b1, b2, lr = 0.9, 0.999, 1e-2
    a = b1 * m1 + (1 - b1) * g
    s = b2 * m2 + (1 - b2) * g * g
    p_ref = p - lr * (a / (1 - b1 ** 3)) / (np.sqrt(s / (1 - b2 ** 3)) + 1e-8)
    assert np.max(np.abs(p2 - p_ref)) < 1e-5 * max(1.0, np.max(np.abs(p)))
