"""The triangle tile selection of the level-3 launches, on the host: the C++
header that the kernels, the launcher and the scoring model share is compiled
with g++ into tests/tri_tiles_check.cpp's checker (no GPU), which enumerates
the tiles of the off-band and band launches and checks that together they
cover the written triangle exactly once."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None or not Path("/opt/rocm/include").is_dir(),
                    reason="needs g++ and the ROCm headers")
def test_band_split_covers_triangle_once(tmp_path):
    exe = tmp_path / "tri_tiles_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{ROOT / 'clgpu_amd' / 'csrc'}", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "tri_tiles_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout
