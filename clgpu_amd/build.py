"""Builds libiclBLAS.so for gfx950 in-tree.

Kernels (*.hip) are compiled by hipcc for --offload-arch=gfx950; the host layer
(*.cpp) by hipcc as plain C++; everything is linked into
clgpu_amd/libiclBLAS.so, which travels with the source tree. Objects go to
clgpu_amd/_build/.

    python -m clgpu_amd.build            # full rebuild
    python -m clgpu_amd.build --incremental
    python -m clgpu_amd.build --sanitize # + clgpu_amd/_build/san/libiclBLAS.so

The sanitizer build compiles the host layer (*.cpp) with g++ and
AddressSanitizer + UndefinedBehaviorSanitizer (host code only; the kernels
are the regular objects) into clgpu_amd/_build/san/libiclBLAS.so, which the
CPU tests load with libasan preloaded (tests/test_sanitizers.py).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
LIB = PKG / "libiclBLAS.so"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-DEXPORT_ICLBLAS_SYMBOLS",
          f"-I{INCLUDE}", f"-I{CSRC}", "-Wall", "-Wno-unused-function"]
# Packed FP32 instructions are disabled for all device code.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HOST = ["-x", "c++", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]


def _headers_digest() -> str:
    h = hashlib.sha256()
    for p in sorted(list(CSRC.glob("*.h")) + list(CSRC.glob("*.hpp")) + list(INCLUDE.glob("*.h"))):
        h.update(p.read_bytes())
    return h.hexdigest()


def _compile(src: Path, incremental: bool, hdr: str) -> Path:
    obj = BUILD / (src.name + ".o")
    stamp = BUILD / (src.name + ".stamp")
    key = hashlib.sha256(src.read_bytes() + hdr.encode() + " ".join(COMMON + DEVICE).encode()).hexdigest()
    if incremental and obj.exists() and stamp.exists() and stamp.read_text() == key:
        return obj
    if src.suffix == ".hip":
        cmd = [HIPCC] + COMMON + DEVICE + ["-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC] + COMMON + HOST + ["-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(line for line in (res.stdout + res.stderr).splitlines()
                     if "packed-fp32-ops" not in line)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{msgs}")
    if msgs.strip():
        print(msgs, file=sys.stderr)
    stamp.write_text(key)
    return obj


def build(incremental: bool = False, jobs: int | None = None, verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    hdr = _headers_digest()
    jobs = jobs or min(8, os.cpu_count() or 4)
    # longest compiles first
    sources.sort(key=lambda p: (p.suffix != ".hip", -p.stat().st_size))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental, hdr), sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs] + \
          ["-Wl,--no-undefined", "-Wl,-soname,libiclBLAS.so"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{res.stdout}{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} ({LIB.stat().st_size // 1024} KiB, {len(objs)} objects)")
    return LIB


SAN = BUILD / "san"
SAN_LIB = SAN / "libiclBLAS.so"
SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined"]


def build_sanitized(jobs: int | None = None, verbose: bool = True) -> Path:
    """Host layer under ASan + UBSan (g++), linked with the regular kernel objects."""
    kernel_objs = [BUILD / (p.name + ".o") for p in sorted(CSRC.glob("*.hip"))]
    if not all(o.exists() for o in kernel_objs):
        build(incremental=True, jobs=jobs, verbose=False)
    SAN.mkdir(parents=True, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    flags = SAN_FLAGS + ["-std=c++17", "-fPIC", "-fvisibility=hidden", "-DEXPORT_ICLBLAS_SYMBOLS",
                         "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{INCLUDE}", f"-I{CSRC}"]

    def one(src: Path) -> Path:
        obj = SAN / (src.name + ".o")
        tmp = SAN / (src.name + f".{os.getpid()}.o")
        res = subprocess.run([cxx] + flags + ["-c", str(src), "-o", str(tmp)], capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"sanitizer compile failed: {src.name}\n{res.stdout}{res.stderr}")
        os.replace(tmp, obj)
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs or min(8, os.cpu_count() or 4)) as ex:
        host_objs = list(ex.map(one, sorted(CSRC.glob("*.cpp"))))
    tmp = SAN_LIB.with_suffix(f".so.tmp{os.getpid()}")
    cmd = [cxx, "-shared", "-fsanitize=address,undefined", "-o", str(tmp)] + [str(o) for o in host_objs] + \
          [str(o) for o in kernel_objs] + [f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib",
                                          "-Wl,--no-undefined", "-Wl,-soname,libiclBLAS.so"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"sanitizer link failed\n{' '.join(cmd)}\n{res.stdout}{res.stderr}")
    os.replace(tmp, SAN_LIB)
    if verbose:
        print(f"built {SAN_LIB}")
    return SAN_LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=None)
    ap.add_argument("--sanitize", action="store_true", help="also build the ASan/UBSan host layer")
    a = ap.parse_args()
    build(incremental=a.incremental, jobs=a.jobs)
    if a.sanitize:
        build_sanitized(jobs=a.jobs)


if __name__ == "__main__":
    main()
