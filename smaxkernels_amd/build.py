"""Builds the native libraries of smaxkernels_amd in-tree.

  _build/libsmax_hip.so     HIP kernels + C ABI (include/smax_hip.h), gfx950
  _build/libsmax_matgen.so  seeded workload generators (host C)
  _build/libsmaxkernels_hip.so  C++ SMAX::Interface with the HIP platform

Run `python -m smaxkernels_amd.build` (or call build_all()).  Every call
recompiles; objects are compiled in parallel.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
GEN = PKG / "gen"
OUT = PKG / "_build"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]

HOST_CFLAGS = ["-std=gnu11", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC",
               "-Wall", "-Wextra"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def hip_sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def build_hip(jobs: int | None = None) -> Path:
    OUT.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in hip_sources():
        obj = OUT / (src.stem + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    jobs = jobs or min(len(cmds), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        list(pool.map(_run, cmds))
    lib = OUT / "libsmax_hip.so"
    tmp = OUT / "libsmax_hip.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC",
          *[str(o) for o in objs], "-o", str(tmp)])
    os.replace(tmp, lib)
    for obj in objs:
        obj.unlink(missing_ok=True)
    return lib


def build_matgen() -> Path:
    OUT.mkdir(exist_ok=True)
    lib = OUT / "libsmax_matgen.so"
    tmp = OUT / "libsmax_matgen.so.tmp"
    cc = shutil.which("gcc") or "cc"
    _run([cc, *HOST_CFLAGS, "-shared", str(GEN / "matgen.c"), str(GEN / "mtx.c"),
          "-o", str(tmp), "-lm"])
    os.replace(tmp, lib)
    return lib


def build_interface() -> Path:
    """The C++ SMAX::Interface with the CPU and HIP platforms (src/), linked
    against libsmax_hip.so."""
    OUT.mkdir(exist_ok=True)
    lib = OUT / "libsmaxkernels_hip.so"
    tmp = OUT / "libsmaxkernels_hip.so.tmp"
    cxx = shutil.which("g++") or "c++"
    src = PKG / "src"
    # the CPU platform's kernels: OpenMP, and hardware fma for their chains
    _run([cxx, "-std=c++17", "-O2", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-Wall",
          "-Wextra", "-shared", f"-I{INCLUDE}", f"-I{src}", str(src / "interface.cpp"),
          str(src / "error_handler.cpp"), "-o", str(tmp), f"-L{OUT}",
          "-lsmax_hip", f"-Wl,-rpath,{OUT}"])
    os.replace(tmp, lib)
    return lib


def build_vendor_cmp() -> Path | None:
    """rocSPARSE CSR SpMV, CSR SpMM and BSR SpMM wrappers used by bench.py as a library baseline
    (comparison only); skipped when rocSPARSE is not installed."""
    rocm = Path(HIPCC).resolve().parent.parent
    if not (rocm / "lib" / "librocsparse.so").exists():
        return None
    OUT.mkdir(exist_ok=True)
    lib = OUT / "librocsparse_cmp.so"
    tmp = OUT / "librocsparse_cmp.so.tmp"
    cxx = shutil.which("g++") or "c++"
    _run([cxx, "-std=c++17", "-O2", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__",
          "-Wno-deprecated-declarations", f"-I{rocm / 'include'}",
          str(PKG / "vendor" / "rocsparse_cmp.cpp"), "-o", str(tmp),
          f"-L{rocm / 'lib'}", "-lrocsparse", "-lamdhip64",
          f"-Wl,-rpath,{rocm / 'lib'}"])
    os.replace(tmp, lib)
    return lib


def build_all() -> None:
    build_matgen()
    build_hip()
    build_interface()
    build_vendor_cmp()


if __name__ == "__main__":
    build_all()
    print(f"built {OUT}")
