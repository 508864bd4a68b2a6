"""Builds libhsg.so in-tree: host C++ (ordering, symbolic analysis) with g++, the HIP kernels and
the C ABI with hipcc for gfx950; and libhsg_solve.so, the C entry of the whole solve
(include/hsg_solve.h), with g++ against the Python headers (it does not link libpython: the
interpreter's symbols come from the process that loads it). Always recompiles (a fresh build is
the build check).

    python -m hsolve_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "libhsg.so"
SOLVE_LIB = HERE / "libhsg_solve.so"
ARCH = os.environ.get("HSG_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # no packed FP32 instructions (v_pk_*_f32) in any kernel: some of their forms have
             # given silently wrong results on this hardware; scalar FP32 code is used instead.
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-pthread"]


def _hipcc() -> str:
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and Path(c).exists():
            return c
    raise RuntimeError("hipcc not found: the HIP toolchain (ROCm) is required to build libhsg")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    hipcc = _hipcc()
    cxx = shutil.which("g++") or "g++"
    jobs = []
    for src in sorted((CSRC / "host").glob("*.cpp")):
        jobs.append([cxx, *CXX_FLAGS, "-c", str(src), "-o", str(BUILD / (src.stem + ".o"))])
    for src in sorted((CSRC / "kernels").glob("*.hip")) + sorted((CSRC / "kernels").glob("*.cpp")):
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(BUILD / (src.stem + ".o"))])
    jobs.append([hipcc, *HIP_FLAGS, "-c", str(CSRC / "capi.cpp"), "-o", str(BUILD / "capi.o")])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, "-shared", "-fPIC", "-pthread", f"--offload-arch={ARCH}", *objs, "-o", str(tmp)])
    os.replace(tmp, LIB)
    build_solve_lib()
    if verbose:
        print(f"built {LIB} and {SOLVE_LIB}")
    return LIB


def build_solve_lib() -> Path:
    """libhsg_solve.so from csrc/upper/solve_capi.cpp."""
    import sysconfig
    cxx = shutil.which("g++") or "g++"
    tmp = SOLVE_LIB.with_suffix(".so.tmp")
    _run([cxx, *CXX_FLAGS, "-shared", f"-I{sysconfig.get_paths()['include']}", str(CSRC / "upper" / "solve_capi.cpp"),
          "-o", str(tmp), "-ldl"])
    os.replace(tmp, SOLVE_LIB)
    return SOLVE_LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
