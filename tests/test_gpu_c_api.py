"""A C program compiled against include/cufinufft.h with the host C
compiler (no HIP compiler, no C++) and run: the guru and one-shot
interfaces checked against direct sums and the reference's error cases
(tests/c/public_api.c; the reference's own C caller of this interface is
test/cuda/public_api_test.c)."""

import os
import shutil
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))


def test_c_program(tmp_path, torch_gpu):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lib = ROOT / "finufft_amd" / "libcufinufft.so"
    assert lib.exists(), "build the library first"
    exe = tmp_path / "public_api"
    cmd = [cc, "-std=c11", "-O1", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
           f"-I{ROOT / 'include'}", f"-I{ROCM / 'include'}", str(ROOT / "tests" / "c" / "public_api.c"),
           "-o", str(exe), f"-L{lib.parent}", "-lcufinufft", f"-L{ROCM / 'lib'}", "-lamdhip64",
           f"-Wl,-rpath,{lib.parent}", f"-Wl,-rpath,{ROCM / 'lib'}", "-lm"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(run.stdout, run.stderr)
    assert run.returncode == 0 and "PASS" in run.stdout
