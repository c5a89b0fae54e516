"""Build libpfhip.so for gfx950 in-tree with hipcc (no JIT cache).

``python -m photonfdtd_amd.build`` compiles every HIP translation unit under
``photonfdtd_amd/csrc`` and links ``photonfdtd_amd/libpfhip.so``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libpfhip.so"
BUILD = PKG.parent / "_build"

# -packed-fp32-ops keeps the compiler from emitting packed FP32 adds/muls/fmas;
# -ffp-contract=off keeps the reference operation order (no implicit FMA) --
# kernels that want FMAs call fma() explicitly.
COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", f"-I{PKG.parent / 'include'}"]

SOURCES = ["kernels_generic.hip", "yee3d_fused.hip", "yee3d_march64.hip", "yee3d_march2.hip", "adjoint.hip", "batch2d.hip", "nn_mfma.hip", "monitor_q.hip", "pfhip.cpp", "stream_stepper.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the HIP library cannot be built")


def build(verbose: bool = False, jobs: int = 0) -> Path:
    """Compile every translation unit (in parallel, ``jobs`` at a time; 0:
    one per source up to the CPU count) and link the library."""
    cc = hipcc()
    BUILD.mkdir(exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        cmds.append([cc, *COMMON, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    workers = jobs or max(1, min(len(cmds), os.cpu_count() or 1, 8))
    running = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        running.append(subprocess.Popen(cmd))
        while len([p for p in running if p.poll() is None]) >= workers:
            running[[p.poll() is None for p in running].index(True)].wait()
    failed = [c for p, c in zip(running, cmds) if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
