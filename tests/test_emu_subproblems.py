"""The bin bounds, the multi-level exclusive scan and the subproblem table
(finufft_amd/csrc/sort.hip: bin_bounds_kernel, two_point_bounds_kernel,
scan_reduce_kernel / scan_apply_kernel, numsub_kernel, sub2bin_kernel; the
reference's calc_subprob / map_b_into_subprob with a Thrust scan,
include/cufinufft/spreadinterp.hpp:333-350), compiled from their own source
for the host with AddressSanitizer (tests/emu/hip_shim.hpp; wave shuffles
become collectives over the wave's threads), against numpy: every subproblem
covers at most maxsub points of one bin, in order, and the counts sum to
the number of points."""

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
SCAN_TILE = 256 * 8


def _kernel_source():
    """sort.hip's anonymous namespace without the host recursion of the scan
    (its launches are replaced by the harness's)."""
    hip = (CSRC / "sort.hip").read_text()
    start = hip.index("namespace fa {")
    anon = hip.index("namespace {", start)
    end = hip.index("\n}  // namespace\n", anon) + len("\n}  // namespace\n")
    body = hip[start:end]
    i = body.index("void scan_rec(")
    j = body.index("\n}\n", i) + 3
    body = body[:i] + body[j:]
    assert "<<<" not in body
    km = "\n".join(ln for ln in (CSRC / "kernel_math.hpp").read_text().splitlines()
                   if not ln.startswith("#include") and ln.strip() != "#pragma once")
    hpp = (CSRC / "sort.hpp").read_text()
    geom = hpp[hpp.index("struct BinGeom {"):hpp.index("// Scratch reused across setpts")]
    return km + "\nnamespace fa {\n" + geom + "\n}\n" + body + "\n}  // namespace fa\n"


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if not CLANG.exists() or shutil.which("ld") is None:
        pytest.skip("host clang++ not available")
    d = tmp_path_factory.mktemp("emu_sub")
    (d / "subproblem_kernels.inc").write_text(_kernel_source())
    out = d / "subproblems_emu"
    cmd = [str(CLANG), "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-DEMU_NO_FA_TYPES", f"-I{EMU}", f"-I{d}",
           str(EMU / "subproblems_emu.cpp"), "-o", str(out), "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    return out


@pytest.mark.parametrize("M,nb,div,maxsub,dense_min", [
    (5000, 37, 1, 64, 0),       # uneven bins, several subproblems per dense bin
    (3000, 2500, 1, 1024, 0),   # more bins than one scan tile's worth of bins
    (4000, 20, 32, 100, 150),   # bank-refined keys (key / 32 = bin), dense bins only
    (0, 9, 1, 1024, 0),         # no points
    (1, 1, 1, 1024, 0),
])
def test_bounds_scan_subproblems(exe, tmp_path, M, nb, div, maxsub, dense_min):
    r = np.random.default_rng(M + nb)
    # skewed bin occupancy: a few crowded bins, many empty ones
    bins = np.minimum((r.pareto(1.2, M) * nb / 8).astype(np.int64), nb - 1)
    keys = np.sort(bins * div + r.integers(0, div, M)).astype(np.int32)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.array([M, nb, div, maxsub, dense_min], np.int64).tobytes())
        f.write(keys.tobytes())
    out = tmp_path / "out.bin"
    res = subprocess.run([str(exe), str(inp), str(out)], capture_output=True, text=True,
                         timeout=600)
    assert res.returncode == 0, res.stderr[-3000:]
    v = np.fromfile(out, dtype=np.int32)
    cap = nb + (M + maxsub - 1) // maxsub
    binstart, substart = v[:nb + 1], v[nb + 1:2 * nb + 2]
    sub2bin = v[2 * nb + 2:2 * nb + 2 + cap]
    long_scan = v[2 * nb + 2 + cap:]
    # bin bounds: first sorted position of each bin
    ref_start = np.searchsorted(keys // div, np.arange(nb + 1), side="left")
    assert np.array_equal(binstart, ref_start)
    cnt = np.diff(ref_start)
    nsub = np.where(cnt > dense_min, (cnt + maxsub - 1) // maxsub, 0)
    ref_sub = np.concatenate([[0], np.cumsum(nsub)])
    assert np.array_equal(substart, ref_sub)
    total = int(ref_sub[-1])
    assert np.array_equal(sub2bin[:total], np.repeat(np.arange(nb), nsub))
    assert np.all(sub2bin[total:] == -1)  # nothing written past the table
    if dense_min == 0:
        assert np.sum(np.minimum(cnt[:, None] - maxsub * np.arange(nsub.max(initial=0) + 1)
                                 [None, :], maxsub).clip(0)) == M
    nl = 40 * SCAN_TILE + 17
    a = (np.arange(nl, dtype=np.int64) * 7919) % 13
    ref_scan = np.concatenate([[0], np.cumsum(a)[:-1]])
    assert np.array_equal(long_scan.astype(np.int64), ref_scan)


def test_scan_extraction():
    assert re.search(r"scan_apply_kernel\(", _kernel_source())
