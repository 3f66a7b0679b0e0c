"""Builds the native library in-tree: lightbaker_amd/liblightbaker_amd.so.

Device code (*.hip) goes through hipcc for gfx950; host-only C++ (*.cpp) through g++. Every unit
is built with -ffp-contract=off: the ray passes keep the reference's FP32 operation order
(crates/lightbaker/src/gpu/trace.wgsl) and the BVH build keeps glam's (bvh.rs), so no fused
multiply-add may change a rounding.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "liblightbaker_amd.so"
OBJ = HERE / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
CXX_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
    "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _hipcc() -> str:
    p = shutil.which("hipcc") or str(ROCM / "bin" / "hipcc")
    if not Path(p).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build")
    return p


def build(verbose: bool = False) -> Path:
    """Compiles every unit under csrc/ and links the shared library; returns its path."""
    OBJ.mkdir(exist_ok=True)
    hipcc = _hipcc()
    jobs = []
    for src in sorted(CSRC.glob("*.hip")):
        obj = OBJ / (src.stem + ".hip.o")
        jobs.append([hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    for src in sorted(CSRC.glob("*.cpp")):
        obj = OBJ / (src.stem + ".cpp.o")
        jobs.append(["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)])
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        for f in cf.as_completed([ex.submit(_run, j) for j in jobs]):
            f.result()
    objs = [j[-1] for j in jobs]
    tmp = OUT.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs])
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
