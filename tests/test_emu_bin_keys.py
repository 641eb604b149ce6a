"""The sort's bin-key kernel (finufft_amd/csrc/sort.hip: bins of the folded
coordinate as include/cufinufft/spreadinterp.hpp:214-229, bins of the first
support node, and their sub-column and LDS-bank refinements), compiled from
its own source for the host with AddressSanitizer (tests/emu/hip_shim.hpp),
bit-exact against oracle/sort.py; and, in 2-D and 3-D, the coordinates the
key kernel packs per point, gathered into the stable key order by the
sort's gather kernel, equal to the permuted inputs. The keys fix the sort
permutation; the radix passes that order them are checked bit-exact on the
GPU (tests/test_gpu_parity.py::test_sort_permutation_bit_exact)."""

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import sort as osort

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "finufft_amd" / "csrc"
EMU = Path(__file__).resolve().parent / "emu"
CLANG = Path("/opt/rocm/llvm/bin/clang++")


def _strip(text):
    return "\n".join(ln for ln in text.splitlines()
                     if not ln.startswith("#include") and ln.strip() != "#pragma once")


def _kernel_source():
    """kernel_math.hpp, BinGeom (sort.hpp) and the key kernel with its fold
    helper (sort.hip)."""
    km = _strip((CSRC / "kernel_math.hpp").read_text())
    hpp = (CSRC / "sort.hpp").read_text()
    geom = hpp[hpp.index("struct BinGeom {"):hpp.index("// Scratch reused across setpts")]
    hip = (CSRC / "sort.hip").read_text()
    i = hip.index("// Folded coordinate, with non-finite input pinned to 0")
    j = hip.index("// binstart[g] = first sorted position")
    return km + "\nnamespace fa {\n" + geom + "\n" + hip[i:j] + "\n}  // namespace fa\n"


_BUILT = {}


def _exe(tmp_path_factory, real, dim):
    if not CLANG.exists() or shutil.which("ld") is None:
        pytest.skip("host clang++ not available")
    if (real, dim) not in _BUILT:
        d = tmp_path_factory.mktemp(f"emu_key_{real}{dim}")
        (d / "bin_key_kernels.inc").write_text(_kernel_source())
        exe = d / "bin_key_emu"
        cmd = [str(CLANG), "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined",
               "-fno-sanitize-recover=undefined", "-DEMU_NO_FA_TYPES", f"-DFA_REAL={real}",
               f"-DFA_DIM={dim}", f"-I{EMU}", f"-I{d}", str(EMU / "bin_key_emu.cpp"),
               "-o", str(exe), "-lpthread"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        assert res.returncode == 0, res.stderr[-4000:]
        _BUILT[(real, dim)] = exe
    return _BUILT[(real, dim)]


CASES = [
    # (real, nf, binsize, mode, ns, col_side)
    ("float", (200,), (64,), "ref", 0, 0),
    ("double", (96, 80), (40, 40), "ref", 0, 0),
    ("float", (32, 48, 40), (8, 8, 8), "ref", 0, 0),
    ("float", (120,), (30,), "start", 7, 0),
    ("double", (60, 50), (12, 9), "start", 12, 0),
    ("float", (32, 32, 24), (8, 8, 6), "col", 7, 2),
    ("float", (40, 32, 48), (8, 8, 12), "col", 4, 4),
    ("float", (48, 40), (32, 8), "bank", 7, 0),
    ("double", (20, 18, 16), (8, 6, 4), "bank", 9, 0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-{len(c[1])}d-{c[3]}")
def test_bin_keys_vs_oracle(tmp_path_factory, tmp_path, case):
    real, nf, bs, mode, ns, cols = case
    dim = len(nf)
    exe = _exe(tmp_path_factory, real, dim)
    r = np.random.default_rng(17)
    M = 5000
    dt = np.float32 if real == "float" else np.float64
    pts = [r.uniform(-np.pi, np.pi, M) for _ in range(dim)]
    # points on the period's edges and far outside it (folded the same way)
    for p in pts:
        p[:8] = [-np.pi, np.pi, 0.0, 3 * np.pi, -7.5, 1e3, np.nextafter(np.pi, 0), -1e-30]
    pts = [p.astype(dt) for p in pts]
    start = mode != "ref"
    fold64 = real == "float"
    hdr = np.array([M, int(start), ns, int(fold64), cols, int(mode == "bank")]
                   + list(nf) + [1] * (3 - dim) + list(bs) + [1] * (3 - dim), np.int64)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(hdr.tobytes())
        for p in pts:
            f.write(p.astype(np.float64).tobytes())
    out = tmp_path / "keys.bin"
    res = subprocess.run([str(exe), str(inp), str(out)], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 0, res.stderr[-3000:]
    raw = np.fromfile(out, dtype=np.uint8)
    keys = raw[:4 * M].view(np.int32).astype(np.int64)
    if dim > 1:
        # the coordinates packed by the key kernel, gathered in the stable
        # key order by the sort's gather kernel: the permuted inputs, exactly
        got = raw[4 * M:].view(np.float64).reshape(dim, M)
        order = np.argsort(keys, kind="stable")
        for d in range(dim):
            assert np.array_equal(got[d], pts[d].astype(np.float64)[order])
    if mode == "ref":
        ref = osort.bin_keys(pts, list(nf), list(bs))
    else:
        ref = osort.start_bin_keys(pts, list(nf), list(bs), ns, fold64=fold64,
                                   col_side=cols, bank=(mode == "bank"))
    assert np.array_equal(keys, np.asarray(ref, np.int64))
