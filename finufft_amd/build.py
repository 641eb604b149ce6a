"""Build recipe of the native library ``finufft_amd/libcufinufft.so`` for gfx950.

Every translation unit is compiled with hipcc (in parallel) and linked against
rocFFT and the HIP runtime. ``python -m finufft_amd.build`` rebuilds
everything; ``--incremental`` recompiles only sources newer than their object
(headers are tracked coarsely: any header change forces a full rebuild).
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import time
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libcufinufft.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = "gfx950"

CXXFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-fvisibility=hidden",
    "-munsafe-fp-atomics",
    # packed fp32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{REPO / 'include'}",
    f"-I{CSRC}",
]


def _units():
    """(source, object, extra defines) for every translation unit."""
    units = []
    for src in ["host_math.cpp", "fft.cpp", "c_api.cpp", "batch.cpp", "plan.hip",
                "plan_kernels.hip", "sort.hip", "mlp.hip", "fused_fft.hip", "probe.hip",
                "density.hip", "fft3.hip"]:
        units.append((CSRC / src, BUILD / (Path(src).stem + ".o"), []))
    for real in ("float", "double"):
        for dim in (1, 2, 3):
            units.append(
                (
                    CSRC / "spreadinterp.hip",
                    BUILD / f"spreadinterp_{real[0]}{dim}.o",
                    [f"-DFA_REAL={real}", f"-DFA_DIM={dim}"],
                )
            )
    return units


def _compile(src: Path, obj: Path, defines: list[str]) -> tuple[Path, float, str]:
    t0 = time.time()
    lang = ["-x", "hip"] if src.suffix == ".hip" else []
    cmd = [HIPCC, *CXXFLAGS, *defines, *lang, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stderr}")
    # hipcc's host pass does not know the packed-fp32 target feature; that
    # warning is expected and dropped.
    warn = "\n".join(
        ln for ln in res.stderr.splitlines() if ln.strip() and "packed-fp32-ops" not in ln
    )
    return obj, time.time() - t0, warn


def build(incremental: bool = False, jobs: int | None = None, verbose: bool = False) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    units = _units()
    newest_header = max(
        [p.stat().st_mtime for p in CSRC.glob("*.hpp")]
        + [p.stat().st_mtime for p in (REPO / "include").glob("*.h")]
    )
    todo = []
    for src, obj, defs in units:
        if (
            incremental
            and obj.exists()
            and obj.stat().st_mtime > src.stat().st_mtime
            and obj.stat().st_mtime > newest_header
        ):
            continue
        todo.append((src, obj, defs))
    jobs = jobs or min(len(todo) or 1, os.cpu_count() or 4, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, dt, warn in ex.map(lambda u: _compile(*u), todo):
            if verbose:
                print(f"  {obj.name}: {dt:.1f}s", flush=True)
            if warn and verbose:
                print(warn)
    objs = [str(obj) for _, obj, _ in units]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [
        HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs,
        f"-L{ROCM / 'lib'}", "-lrocfft", "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}",
    ]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=None)
    args = ap.parse_args(argv)
    t0 = time.time()
    lib = build(incremental=args.incremental, jobs=args.jobs, verbose=True)
    print(f"built {lib} in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    sys.exit(main())
