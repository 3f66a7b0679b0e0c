"""Builds the oracle (test infrastructure): oracle/liboracle.so from the C restatement.

Built with gcc -ffp-contract=off (WGSL's operation order, no fused multiply-add) and OpenMP over
items for the CPU baseline. The reference itself (Rust + wgpu) has no toolchain in this image, so
there is no oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "liboracle.so"
SOURCES = ["trace_ref.c", "filter_ref.c", "fp64_ref.c", "adjoint_ref.c", "finish_ref.c", "deps_ref.c"]


def build(verbose: bool = False) -> Path:
    srcs = [str(HERE / s) for s in SOURCES if (HERE / s).exists()]
    tmp = OUT.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fopenmp",
           "-Wall", "-Wno-unused-function", "-o", str(tmp), *srcs, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed: {' '.join(cmd)}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
