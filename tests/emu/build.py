"""Builds the CPU kernel emulator: the library's HIP sources compiled as host
C++ against tests/emu/include/hip/hip_runtime.h, with AddressSanitizer, into
tests/emu/lib/libtetra_hip_emu.so (same C ABI as libtetra_hip.so).

Test infrastructure only: nothing in the package loads this library.  Run
`python tests/emu/build.py`; tests/test_emu_kernels.py builds it on demand.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
HIP_SRC = ROOT / "tetragpu_amd" / "csrc" / "hip"
LIB = HERE / "lib"
OBJ = LIB / "obj"
CLANG = "/opt/rocm/llvm/bin/clang++"
# pool.hip is replaced by the exact-size allocator in emu_runtime.cpp
SKIP = {"pool.hip"}
FLAGS = ["-x", "c++", "-std=c++17", "-O1", "-g", "-fPIC", "-fsanitize=address",
         "-fno-omit-frame-pointer", "-Wno-unused-function", "-Wno-unused-variable",
         "-Wno-unknown-attributes", f"-I{HERE / 'include'}", f"-I{ROOT / 'include'}",
         f"-I{HIP_SRC}"]


def asan_runtime() -> str:
    r = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"],
                       capture_output=True, text=True, check=True)
    return r.stdout.strip()


def _cc(src: Path) -> Path:
    out = OBJ / (src.stem + ".o")
    if out.exists() and out.stat().st_mtime > max(src.stat().st_mtime, _newest_header()):
        return out
    r = subprocess.run([CLANG, *FLAGS, "-c", str(src), "-o", str(out)],
                       capture_output=True, text=True)
    if r.returncode:
        sys.stderr.write(r.stderr)
        raise RuntimeError(f"emulator build failed: {src.name}")
    return out


def _newest_header() -> float:
    hs = list(HIP_SRC.glob("*.h")) + list((HERE / "include").rglob("*.h")) + \
        list((ROOT / "include").rglob("*.h"))
    return max(p.stat().st_mtime for p in hs)


def build(jobs: int = 8) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = [p for p in sorted(HIP_SRC.glob("*.hip")) if p.name not in SKIP]
    srcs.append(HERE / "emu_runtime.cpp")
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_cc, srcs))
    so = LIB / "libtetra_hip_emu.so"
    tmp = LIB / "libtetra_hip_emu.so.tmp"
    r = subprocess.run([CLANG, "-shared", "-fsanitize=address", "-shared-libasan",
                        *map(str, objs), "-o", str(tmp)], capture_output=True, text=True)
    if r.returncode:
        sys.stderr.write(r.stderr)
        raise RuntimeError("emulator link failed")
    os.replace(tmp, so)
    return so


if __name__ == "__main__":
    print(build())
