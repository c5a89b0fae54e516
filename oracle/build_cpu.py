"""Build ``oracle/libpfcpu.so``, the CPU FDTD stepper of ``cpu_fdtd.cpp``
(test and baseline infrastructure): g++ with OpenMP, no FP contraction,
x86-64-v3 (AVX2) so the library runs on this container and on the GPU
box's host alike. ``python -m oracle.build_cpu``."""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "libpfcpu.so"


def build(verbose: bool = False) -> Path:
    cxx = os.environ.get("CXX") or shutil.which("g++") or "g++"
    # a private temporary per process, then an atomic rename: parallel test
    # workers building at once never see a half-written library
    tmp = OUT.with_name(f".{OUT.name}.{os.getpid()}.tmp")
    cmd = [cxx, "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-march=x86-64-v3", f"-I{HERE.parent / 'include'}", str(HERE / "cpu_fdtd.cpp"),
           "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
