"""Benchmark: 3-D dispersive slab stepping on one MI355X.

Workload (the headline configuration): 512^3 cells at dx = 20 nm, 12-cell
CPML on all six faces, float32; silica cladding (eps 2.085), a silicon
Sellmeier box 200 nm x 500 nm x 8 um, a 300 nm gold Drude cube, an Ey point
dipole (Gaussian pulse at 1.55 um, 10 fs FWHM) and two x-plane DFT monitors
(Ey, Ez, Hy, Hz at 32 frequencies, complex128 accumulation).

A step is one full Yee time step of the whole grid (H and E updates, CPML,
ADE poles, sources and DFT accumulation). ``value`` is Gcell-updates/s over
the timed steps, fields resident in HBM. Printed as one JSON line, together
with the roofline of the stencil kernels (HIP events on the stepper's stream),
the CPU baseline (the NumPy float64 oracle on a bounded 96^3 sample) and the
float64/float32 error of the GPU stepper against that oracle run.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Gcell-updates/s & % HBM peak; verified design iters/s surrogate vs plain"
HBM_PEAK_GBS = 8000.0
DX = 20e-9
L_REF = 512 * DX          # positions are fractions of the 512-cell domain


COURANT = 0.9


def dispersive_slab(pf, n=512, steps=2000, precision="float32", backend="hip"):
    """The 3-D dispersive slab scene on an n^3 grid (positions scale with n).

    Courant number 0.9: with the reference's explicit ADE update, both the
    silicon Sellmeier poles and the gold Drude pole grow without bound at its
    default 0.99 on a 3-D 20 nm grid (the oracle overflows within ~400
    steps), and stay bounded at 0.9.
    """
    L = n * DX
    s = L / L_REF
    gold_drude = pf.DispersiveMedium(eps_inf=1.0, poles=(pf.gold().poles[0],),
                                     name="Au Drude term (Rakic row 0)")
    structures = [
        pf.Box(center=(0, 0, 0), size=(3 * L,) * 3, medium=pf.Medium(2.085)),
        pf.Box(center=(0, 0, 0), size=(200e-9, 500e-9, 8e-6), medium=pf.silicon()),
        pf.Box(center=(1.5e-6 * s, 1.5e-6 * s, 0), size=(300e-9,) * 3, medium=gold_drude),
    ]
    src = pf.PointDipole(position=(-1.5e-6 * s, 0, 0), component="Ey",
                         waveform=pf.GaussianPulse(freq0=pf.C_0 / 1.55e-6, fwhm=10e-15))
    freqs = np.linspace(pf.C_0 / 1.65e-6, pf.C_0 / 1.45e-6, 32)
    mons = [pf.DFTMonitor(name=nm, components=("Ey", "Ez", "Hy", "Hz"), freqs=freqs,
                          plane_axis="x", plane_position=x)
            for nm, x in (("in", -3e-6 * s), ("out", 3e-6 * s))]
    grid = pf.Grid(size=(L, L, L), cell_size=DX, pml_layers=(12, 12, 12))
    dt = COURANT / (pf.C_0 * math.sqrt(3.0 / DX ** 2))
    return pf.Simulation(grid, structures=structures, sources=[src], monitors=mons,
                         run_time=(steps - 0.5) * dt, courant=COURANT, precision=precision,
                         backend=backend)


def _rel(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) /
                 max(np.linalg.norm(np.asarray(b).ravel()), 1e-300))


def parity_and_cpu(n=80, steps=100):
    """Oracle run on an n^3 sample (timed: the CPU baseline) and the GPU
    float64 / float32 errors against it (DFT planes and final fields)."""
    import photonfdtd_amd as pf
    from photonfdtd_amd.hip_backend import HipEngine
    from oracle.fdtd_numpy import run_oracle

    sim64 = dispersive_slab(pf, n, steps, "float64")
    t0 = time.perf_counter()
    ref, state = run_oracle(sim64, return_state=True)
    cpu_s = time.perf_counter() - t0
    cells = sim64.grid.n_cells

    errs = {}
    for prec, exact in (("float64", True), ("float32", False)):
        sim = dispersive_slab(pf, n, steps, prec)
        eng = HipEngine(sim, np.float64 if prec == "float64" else np.float32, exact)
        eng.advance(0, sim.n_steps)
        worst = 0.0
        for nm in ("in", "out"):
            got = eng.dft(nm)
            for c, v in got.items():
                worst = max(worst, _rel(v, ref["dft"][nm][c]))
        for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
            worst = max(worst, _rel(eng.get(c), state[c]))
        eng.close()
        errs[prec] = worst
    return dict(cpu_seconds=cpu_s, cpu_value=cells * steps / cpu_s / 1e9,
                sample=f"NumPy float64 oracle, {n}^3 cells x {steps} steps of the same scene "
                       f"({cells * steps:.3g} cell-updates), 1 thread",
                fp64_rel_l2=errs["float64"], fp32_rel_l2=errs["float32"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--precision", default="float32", choices=("float32", "float64"))
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this stepper runs on one GPU; multi-GPU runs are not "
                                   "part of this build (see the sharding design notes)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import photonfdtd_amd as pf
    from photonfdtd_amd import _native as N
    from photonfdtd_amd.hip_backend import HipEngine, device_info, exact_mode

    check = None
    if not args.no_cpu:
        check = parity_and_cpu()
        print(f"[bench] oracle sample done: {check}", file=sys.stderr, flush=True)

    total = args.warmup + args.steps
    t_setup = time.perf_counter()
    sim = dispersive_slab(pf, args.n, total, args.precision)
    npdt = np.float32 if args.precision == "float32" else np.float64
    exact = exact_mode(npdt)
    eng = HipEngine(sim, npdt, exact)
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] setup {setup_s:.1f} s, grid {sim.grid.shape}, {total} steps", file=sys.stderr,
          flush=True)
    eng.advance(0, args.warmup)
    lib = eng.lib
    N.check(lib, lib.pf_timing_enable(eng.h, 1), "pf_timing_enable")
    t0 = time.perf_counter()
    eng.advance(args.warmup, args.steps)
    wall = time.perf_counter() - t0
    ms = (np.zeros(4, dtype=np.float64))
    cnt = np.zeros(4, dtype=np.int64)
    N.check(lib, lib.pf_timing_read(eng.h, N.as_ptr(ms, N.C.c_double), N.as_ptr(cnt, N.C.c_int64)),
            "pf_timing_read")
    # sanity: the DFT accumulators must be finite
    finite = all(np.isfinite(v).all() for v in eng.dft("out").values())
    eng.close()

    cells = sim.grid.n_cells
    value = cells * args.steps / wall / 1e9
    itemsize = np.dtype(npdt).itemsize
    bytes_per_update = 13 * itemsize            # read 6 fields + ce, write 6 fields
    k = dict(zip(N.KERNEL_CLASSES, zip(ms.tolist(), cnt.tolist())))
    if cnt[2] > 0:
        step_ms = ms[2] / cnt[2]
        kernel = "yee_fused"
    else:
        step_ms = ms[0] / max(cnt[0], 1) + ms[1] / max(cnt[1], 1)
        kernel = "yee_h + yee_e"
    achieved = cells * bytes_per_update / (step_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if npdt == np.float32 else "f64",
        "data": "synthetic",
        "config": {"workload": f"3-D dispersive slab {args.n}^3, dx 20 nm, 12-cell CPML x6, "
                               "Si Sellmeier box + Au Drude cube in silica, Ey dipole, "
                               "2 x-plane DFT monitors x 4 comps x 32 freqs",
                   "cells": cells, "exact_order": exact, "device": device_info()[0]},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_cell_update": bytes_per_update,
                     "kernel_ms_per_step": round(step_ms, 4)},
        "kernel_ms": {name: {"total_ms": round(t, 3), "launch_groups": int(c)}
                      for name, (t, c) in k.items()},
        "hbm_fraction_of_value": round(value * 1e9 * bytes_per_update / 1e9 / HBM_PEAK_GBS, 4),
        "dft_finite": bool(finite),
        "setup_s": round(setup_s, 1),
    }
    if check is not None:
        out["cpu_baseline"] = {"value": round(check["cpu_value"], 5), "unit": "Gcell-updates/s",
                               "cores": 1, "kind": "port", "sample": check["sample"]}
        out["fp64_rel_l2_vs_oracle"] = check["fp64_rel_l2"]
        out["fp32_rel_l2_vs_oracle"] = check["fp32_rel_l2"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
