"""Build recipe for libsk_gpu.so (HIP for gfx950) from sih26_amd/csrc.

    python sih26_amd/build.py            # compile sih26_amd/lib/libsk_gpu.so and the
                                         # CPU-only sih26_amd/lib/libsk_gpu_stub.so

Device sources (*.hip) go through hipcc for gfx950 only; host sources (*.cpp) through g++.
Everything links into one shared library with a plain C ABI (include/sk_gpu.h).
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
LIBDIR = PKG / "lib"
OBJDIR = PKG / "lib" / "obj"
LIBNAME = "libsk_gpu.so"
STUB_LIBNAME = "libsk_gpu_stub.so"
ARCH = os.environ.get("SK_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", shutil.which("g++") or "g++")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # keep packed-FP32 instructions out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unknown-pragmas",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    # hipcc's host pass does not know the packed-fp32 target feature; that warning is noise
    noise = ("packed-fp32-ops", "not a recognized feature")
    for line in (proc.stdout + proc.stderr).splitlines():
        if line.strip() and not any(n in line for n in noise):
            sys.stderr.write(line + "\n")


def build_stub(verbose: bool = False) -> Path:
    """The CPU-only library, sih26_amd/lib/libsk_gpu_stub.so: every host source with g++, and
    csrc/stub/ in place of the HIP translation units (no device code; device entry points return
    SK_GPU_UNAVAILABLE). The counterpart of the reference's no-CUDA build, which links
    src/gpu/gpu_runtime_stub.cpp instead of the .cu files (CMakeLists.txt:212-222)."""
    LIBDIR.mkdir(parents=True, exist_ok=True)
    objdir = OBJDIR / "stub"
    objdir.mkdir(parents=True, exist_ok=True)
    jobs = []
    for src in sorted(CSRC.glob("*.cpp")) + sorted((CSRC / "stub").glob("*.cpp")):
        obj = objdir / (src.stem + ".cpp.o")
        jobs.append([CXX, *CXX_FLAGS, "-I", str(ROOT / "include"), "-c", "-o", str(obj), str(src)])
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        for fut in [pool.submit(_run, j) for j in jobs]:
            fut.result()
    objs = [j[j.index("-o") + 1] for j in jobs]
    out = LIBDIR / STUB_LIBNAME
    tmp = LIBDIR / (STUB_LIBNAME + ".tmp")
    _run([CXX, "-shared", "-fPIC", "-pthread", "-o", str(tmp), *objs])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def build(verbose: bool = False) -> Path:
    """Compile every source and link sih26_amd/lib/libsk_gpu.so. Always recompiles."""
    LIBDIR.mkdir(parents=True, exist_ok=True)
    OBJDIR.mkdir(parents=True, exist_ok=True)
    jobs = []
    for src in sorted(CSRC.glob("*.hip")):
        obj = OBJDIR / (src.stem + ".hip.o")
        jobs.append([HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), "-c", "-o", str(obj), str(src)])
    for src in sorted(CSRC.glob("*.cpp")):
        obj = OBJDIR / (src.stem + ".cpp.o")
        jobs.append([CXX, *CXX_FLAGS, "-I", str(ROOT / "include"), "-c", "-o", str(obj), str(src)])
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        for fut in [pool.submit(_run, j) for j in jobs]:
            fut.result()
    objs = [j[j.index("-o") + 1] for j in jobs]
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", str(tmp), *objs,
          "-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_stub(verbose=True)
