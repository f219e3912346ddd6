"""The CPU tests of the C-ABI host layer and of the oracle, rerun with both
built under AddressSanitizer and UndefinedBehaviorSanitizer (host code only:
clgpu_amd/_build/san/libiclBLAS.so from `python -m clgpu_amd.build --sanitize`,
oracle/_san/liboracle.so from `make -C oracle san`), loaded into a Python with
libasan and libubsan preloaded. Any report fails the run: UBSan is built
without recovery and ASan aborts on its first error. First, an out-of-bounds
read planted in a call of the oracle must be reported, so that a silently
uninstrumented build cannot pass."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SUITE = ["tests/test_capi_cpu.py", "tests/test_oracle.py", "tests/test_oracle_level3.py", "tests/test_oracle_gemv.py"]


def _runtime(name):
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture
def san_env():
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("no host C/C++ compiler")
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("no ASan / UBSan runtime")
    sys.path.insert(0, str(ROOT))
    from clgpu_amd import build as b
    lib = b.build_sanitized(verbose=False)
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "san"], check=True)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": f"{asan}:{ubsan}",
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "ICLBLAS_LIBRARY": str(lib),
        "ICLBLAS_ORACLE_LIBRARY": str(ROOT / "oracle" / "_san" / "liboracle.so"),
    })
    return env


def test_cpu_suite_under_asan_and_ubsan(san_env):
    _planted_overflow_is_reported(san_env)
    r = subprocess.run([sys.executable, "-m", "pytest", *SUITE, "-x", "-q", "-m", "not gpu", "-p", "no:cacheprovider"],
                       cwd=ROOT, env=san_env, capture_output=True, text=True, timeout=900)
    out = r.stdout[-4000:] + r.stderr[-4000:]
    assert r.returncode == 0, out
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, out
    assert " passed" in r.stdout, out


def _planted_overflow_is_reported(san_env):
    # SGEMM 4 x 4 x 4 with a B of only two columns: the oracle reads past it.
    code = ("import numpy as np, oracle\n"
            "A = np.zeros((4, 4), np.float32, order='F'); B = np.zeros((4, 2), np.float32, order='F')\n"
            "C = np.zeros((4, 4), np.float32, order='F')\n"
            "oracle.lib().orc_sgemm(0, 0, 4, 4, 4, 1.0, A.ctypes.data, 4, B.ctypes.data, 4, 0.0, C.ctypes.data, 4)\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=san_env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "AddressSanitizer: heap-buffer-overflow" in r.stderr, r.stderr[-3000:]
