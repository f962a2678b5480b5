"""Builds the native libraries of tetragpu_amd in-tree.

  tetragpu_amd/lib/libtetra_hip.so   HIP kernels for gfx950 + the C ABI
                                     (include/tetra_c.h)
  tetragpu_amd/lib/libtetra_host.so  host C++: mesh generator
                                     (include/tetra_meshgen.h) and the C++
                                     drop-in layer of the reference API
  tetragpu_amd/lib/tetra_main        command-line driver (reference options)

Run `python -m tetragpu_amd.build` (or __graft_entry__.build()).  hipcc
cross-compiles gfx950 code without a GPU present.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "lib"
OBJ = LIB / "obj"
HIP_SRC = PKG / "csrc" / "hip"
HOST_SRC = PKG / "csrc" / "host"
INCLUDE = ROOT / "include"

ARCH = os.environ.get("TETRA_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in every device compile
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}", f"-I{HIP_SRC}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-fno-fast-math",
              "-Wall", "-Wextra", "-Wno-unused-parameter", f"-I{INCLUDE}",
              f"-I{INCLUDE / 'tetra'}", f"-I{HOST_SRC}"]


def _run(cmd: list[str], quiet: bool = True) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    if not quiet and r.stderr.strip():
        sys.stderr.write(r.stderr)


def _compile_hip(src: Path) -> Path:
    out = OBJ / (src.stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(out)])
    return out


def build_hip(jobs: int = 8) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = sorted(HIP_SRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_hip, srcs))
    so = LIB / "libtetra_hip.so"
    tmp = LIB / "libtetra_hip.so.tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs),
          "-o", str(tmp)])
    os.replace(tmp, so)
    return so


def build_host(jobs: int = 8) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = [p for p in sorted(HOST_SRC.glob("*.cpp")) if p.name != "tetra_main.cpp"]

    def comp(src: Path) -> Path:
        out = OBJ / ("host_" + src.stem + ".o")
        _run([CXX, *HOST_FLAGS, "-c", str(src), "-o", str(out)])
        return out

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(comp, srcs))
    so = LIB / "libtetra_host.so"
    tmp = LIB / "libtetra_host.so.tmp"
    _run([CXX, "-shared", "-fopenmp", *map(str, objs), "-o", str(tmp),
          f"-L{LIB}", "-ltetra_hip", "-lz", f"-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, so)
    main_src = HOST_SRC / "tetra_main.cpp"
    if main_src.exists():
        exe = LIB / "tetra_main"
        _run([CXX, *HOST_FLAGS, str(main_src), "-o", str(exe) + ".tmp",
              f"-L{LIB}", "-ltetra_host", "-ltetra_hip", "-Wl,-rpath,$ORIGIN"])
        os.replace(str(exe) + ".tmp", exe)
    return so


def build_all(jobs: int = 8) -> None:
    build_hip(jobs)
    build_host(jobs)


if __name__ == "__main__":
    build_all()
    print("built", *(str(p.relative_to(ROOT)) for p in sorted(LIB.glob("*")) if p.is_file()))
