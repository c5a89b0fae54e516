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
the CPU baseline (the C++/OpenMP restatement of the reference's compiled CPU stepper on
a bounded sample of the full workload, host threads stated), the NumPy float64 oracle on a
64^3 x 1000-step sample, the float64/float32 error of the GPU stepper against that oracle
run, and the full-size float32 error against the float64 GPU stepper.
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


def dielectric_slab(pf, n=512, steps=128, precision="float32", backend="hip"):
    """The slab scene without dispersion (the silicon box as eps 12.1, no gold
    cube): the workload of the beyond-HBM streaming stepper, which does not
    step ADE media."""
    sim = dispersive_slab(pf, n, steps, precision, backend)
    sim.structures = [sim.structures[0],
                      pf.Box(center=sim.structures[1].center, size=sim.structures[1].size,
                             medium=pf.Medium(12.1))]
    return type(sim)(sim.grid, structures=sim.structures, sources=sim.sources,
                     monitors=sim.monitors, run_time=sim.run_time, courant=COURANT,
                     precision=precision, backend=backend)


def batched_2d_scenes(pf, n_sims=64, lo=256, hi=1024, steps=4000, precision="float32", seed=0,
                      backend="hip", component="Ex"):
    """Independent 2-D scenes of uneven size for batched stepping.

    Sides uniform in [lo, hi] cells (square), dx = 20 nm, 12-cell CPML on x
    and y, the 2-D default Courant number. Each scene holds 1-4 PolySlab
    scatterers (3-8 vertices at sorted random angles around a random centre,
    radius 0.3-1.5 um, clipped to the region between the DFT planes), index
    uniform in [1.44, 3.48], 20% of them a single-pole Lorentz medium; a
    point dipole on ``component`` (Gaussian pulse at 1.55 um, 10 fs FWHM);
    two x-plane DFT monitors (Ey, Hz for TE, Ez, Hy for TM) at
    x = +-(L/2 - 1 um), 16 frequencies in [c/1.7 um, c/1.4 um].
    """
    rng = np.random.default_rng(seed)
    te = component in ("Ex", "Ey", "Hz")
    comps = ("Ey", "Hz") if te else ("Ez", "Hy")
    freqs = np.linspace(pf.C_0 / 1.7e-6, pf.C_0 / 1.4e-6, 16)
    sims = []
    for _ in range(n_sims):
        n = int(rng.integers(lo, hi + 1))
        L = n * DX
        sc = min(1.0, L / (256 * DX))     # margins shrink only below 256 cells
        xp = L / 2 - 1e-6 * sc
        ylim = L / 2 - 14 * DX
        structures = []
        for _p in range(int(rng.integers(1, 5))):
            nv = int(rng.integers(3, 9))
            c = rng.uniform(-(xp - 0.5e-6 * sc), xp - 0.5e-6 * sc, size=2)
            r = rng.uniform(0.3e-6, 1.5e-6) * sc
            ang = np.sort(rng.uniform(0, 2 * np.pi, size=nv))
            vx = np.clip(c[0] + r * np.cos(ang), -(xp - 2 * DX), xp - 2 * DX)
            vy = np.clip(c[1] + r * np.sin(ang), -ylim, ylim)
            idx = rng.uniform(1.44, 3.48)
            if rng.uniform() < 0.2:
                med = pf.DispersiveMedium.lorentz(
                    idx ** 2, [(rng.uniform(0.1, 1.0), rng.uniform(100e12, 400e12),
                                rng.uniform(0.0, 20e12))])
            else:
                med = pf.Medium.from_index(idx)
            structures.append(pf.PolySlab(vertices=tuple(zip(vx, vy)), z_bounds=(-1.0, 1.0),
                                          medium=med))
        src = pf.PointDipole(position=(-(L / 2 - 0.6e-6 * sc), 0.0), component=component,
                             waveform=pf.GaussianPulse(freq0=pf.C_0 / 1.55e-6, fwhm=10e-15))
        mons = [pf.DFTMonitor(name=nm, components=comps, freqs=freqs, plane_axis="x",
                              plane_position=x) for nm, x in (("in", -xp), ("out", xp))]
        grid = pf.Grid(size=(L, L), cell_size=DX, pml_layers=(12, 12, 0))
        probe = pf.Simulation(grid, structures=structures, sources=[src], monitors=mons,
                              run_time=1e-15, precision=precision, backend=backend)
        sims.append(pf.Simulation(grid, structures=structures, sources=[src], monitors=mons,
                                  run_time=(steps - 0.5) * probe.dt, precision=precision,
                                  backend=backend))
    return sims


def _rel(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) /
                 max(np.linalg.norm(np.asarray(b).ravel()), 1e-300))


def parity_and_cpu(n=64, steps=1000):
    """Oracle run on an n^3 sample (timed: the CPU baseline) and the GPU
    float64 / float32 errors against it (DFT planes and final fields). The
    reference operation order in float32 (the oracle run at float32) is
    reported beside the GPU float32 error as the intrinsic float32 level."""
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
    ref32, state32 = run_oracle(dispersive_slab(pf, n, steps, "float32"), return_state=True)
    intrinsic = 0.0
    for nm in ("in", "out"):
        for c, v in ref32["dft"][nm].items():
            intrinsic = max(intrinsic, _rel(v, ref["dft"][nm][c]))
    for c in state:
        intrinsic = max(intrinsic, _rel(state32[c], state[c]))
    errs["ref_float32"] = intrinsic
    return dict(cpu_seconds=cpu_s, cpu_value=cells * steps / cpu_s / 1e9,
                sample=f"NumPy float64 oracle, {n}^3 cells x {steps} steps of the same scene "
                       f"({cells * steps:.3g} cell-updates), 1 thread",
                fp64_rel_l2=errs["float64"], fp32_rel_l2=errs["float32"],
                ref_fp32_rel_l2=errs["ref_float32"])


def fullsize_fp32_error(n, steps, exact32=False):
    """Relative L2 of the float32 stepper (fast order, or the reference order
    with ``exact32``) against the float64 stepper (reference order, bitwise
    equal to the oracle on every scene the tests compare) on the full
    workload: final six fields and both DFT planes, n^3 cells x ``steps``
    steps, both on the GPU. Returns the worst array, the worst field, the
    worst DFT plane and every array's value."""
    import photonfdtd_amd as pf
    from photonfdtd_amd.hip_backend import HipEngine

    t0 = time.perf_counter()
    sim32 = dispersive_slab(pf, n, steps, "float32")
    eng = HipEngine(sim32, np.float32, exact32)
    eng.advance(0, sim32.n_steps)
    f32 = {c: eng.get(c).copy() for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
    d32 = {nm: eng.dft(nm) for nm in ("in", "out")}
    eng.close()
    sim64 = dispersive_slab(pf, n, steps, "float64")
    eng = HipEngine(sim64, np.float64, True)
    eng.advance(0, sim64.n_steps)
    per = {}
    for c, a in f32.items():
        b = eng.get(c)
        num = den = 0.0
        for i in range(0, b.shape[0], 32):          # bounded host temporaries
            d = a[i:i + 32].astype(np.float64) - b[i:i + 32]
            num += float(np.vdot(d, d).real)
            den += float(np.vdot(b[i:i + 32], b[i:i + 32]).real)
        per[c] = math.sqrt(num / max(den, 1e-300))
    for nm in ("in", "out"):
        got = eng.dft(nm)
        for c, v in got.items():
            per[f"dft_{nm}_{c}"] = _rel(d32[nm][c], v)
    eng.close()
    worst = max(per.values())
    fields = max(v for k, v in per.items() if not k.startswith("dft"))
    dfts = max(v for k, v in per.items() if k.startswith("dft"))
    return dict(rel_l2=worst, fields_rel_l2=fields, dft_rel_l2=dfts, steps=steps, cells=n ** 3,
                seconds=round(time.perf_counter() - t0, 1), worst_array=max(per, key=per.get),
                per_array={k: float(f"{v:.3e}") for k, v in per.items()})


def host_cpu():
    """CPU model name and logical CPU count of this host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_stepper_sample(n, seconds=15.0, max_steps=200):
    """The compiled CPU baseline: the OpenMP restatement of the reference's
    fused CPU stepper (rust/src/kernel.rs order) with the ADE and per-step
    DFT of this workload, on the full n^3 float32 scene, timed over as many
    steps as fit in ``seconds`` (after one untimed step) on the OpenMP
    threads the host grants (OMP_NUM_THREADS)."""
    import photonfdtd_amd as pf
    from oracle.fdtd_cpu import CpuStepper
    sim = dispersive_slab(pf, n, max_steps + 1, "float32")
    eng = CpuStepper(sim, np.float32, threads=0)
    eng.advance(0, 1)
    done, t0 = 0, time.perf_counter()
    while done < max_steps and time.perf_counter() - t0 < seconds:
        eng.advance(1 + done, 1)
        done += 1
    wall = time.perf_counter() - t0
    threads = eng.threads
    eng.close()
    model, ncpu = host_cpu()
    return dict(value=n ** 3 * done / wall / 1e9, steps=done, seconds=wall, threads=threads,
                sample=(f"C++/OpenMP restatement of the reference's compiled CPU stepper "
                        f"(rust/src/kernel.rs operation order) plus this workload's ADE and "
                        f"per-step DFT: the full {n}^3 float32 scene, {done} steps after one "
                        f"untimed step, {threads} OpenMP threads on {model} "
                        f"({ncpu} logical CPUs visible)"))


PROFILES = ROOT / "profiles"


def measured_traffic(name):
    """HBM bytes per launch of the step kernel from the committed counter
    summary of this workload (FETCH_SIZE and WRITE_SIZE, each scaled by what
    the counter reads for the full-field gather copy), or None."""
    f = PROFILES / f"{name}_traffic.json"
    if not f.exists():
        return None
    d = json.loads(f.read_text())
    return d.get("hbm_bytes_per_step", d.get("hbm_bytes_per_launch"))


def summarize(prof_dir, kind, out=None):
    """Summarize rocprofv3 CSV output (kernel stats or one PMC counter)."""
    import csv
    import glob
    prof_dir = Path(prof_dir)
    lines = []
    if kind == "stats":
        path = sorted(glob.glob(str(prof_dir / "**" / "*kernel_stats.csv"), recursive=True))[0]
        rows = list(csv.DictReader(open(path)))
        lines.append("kernel | calls | total_ms | avg_us | min_us | max_us")
        for r in rows:
            lines.append(f"{r['Name'][:90]} | {r['Calls']} | {float(r['TotalDurationNs'])/1e6:.3f} | "
                         f"{float(r['AverageNs'])/1e3:.2f} | {float(r['MinNs'])/1e3:.2f} | "
                         f"{float(r['MaxNs'])/1e3:.2f}")
    else:
        path = sorted(glob.glob(str(prof_dir / "**" / "*counter_collection.csv"), recursive=True))[0]
        agg = {}
        for r in csv.DictReader(open(path)):
            key = (r["Kernel_Name"], r["Counter_Name"])
            agg.setdefault(key, []).append(float(r["Counter_Value"]))
        lines.append("kernel | counter | dispatches | mean value (KiB per dispatch)")
        for (k, c), v in sorted(agg.items()):
            lines.append(f"{k[:90]} | {c} | {len(v)} | {sum(v) / len(v):.1f}")
    text = "\n".join(lines) + "\n"
    if out:
        Path(out).write_text(text)
    print(text)
    return text


def traffic_from_counters(fetch_dir, write_dir, kernel, copy_cells, alg_bytes, itemsize, name,
                          workload, copy_fetch_cells=None):
    """HBM bytes per step of `kernel` (all its variants) from one FETCH_SIZE
    and one WRITE_SIZE rocprofv3 run. Each counter is scaled by its reading
    for the full-field k_gather_box of the same run (a copy of copy_cells
    values each way; copy_fetch_cells: the values its reads cover when the
    field rows are padded, whole rows). Writes profiles/<name>_traffic.json."""
    import csv
    import glob

    def dispatches(d):
        path = sorted(glob.glob(str(Path(d) / "**" / "*counter_collection.csv"), recursive=True))[0]
        return [(r["Kernel_Name"], float(r["Counter_Value"]), int(r["Grid_Size"]))
                for r in csv.DictReader(open(path))]
    true_kib = copy_cells * itemsize / 1024
    true_fetch_kib = (copy_fetch_cells or copy_cells) * itemsize / 1024
    fe, wr = dispatches(fetch_dir), dispatches(write_dir)
    copy_f = [v for k, v, g in fe if "k_gather_box" in k and g >= copy_cells][0]
    copy_w = [v for k, v, g in wr if "k_gather_box" in k and g >= copy_cells][0]
    def per_step(rows):
        # every variant of the step kernel launches once per step (the 3-D
        # march: psi-free tiles and CPML tiles): mean per variant, summed
        by = {}
        for k, v, g in rows:
            if kernel in k:
                by.setdefault(k, []).append(v)
        return sum(sum(v) / len(v) for v in by.values()), sorted(by)
    step_f, variants = per_step(fe)
    step_w, _ = per_step(wr)
    fetch = step_f * (true_fetch_kib / copy_f) * 1024
    write = step_w * (true_kib / copy_w) * 1024
    out = {"kernel": kernel, "variants": variants, "workload": workload,
           "fetch_kib_raw_per_step": step_f,
           "write_kib_raw_per_step": step_w,
           "copy_calibration": {"kernel": f"k_gather_box full field ({copy_cells} values each way)",
                                "fetch_kib_raw": copy_f, "write_kib_raw": copy_w,
                                "true_fetch_kib": true_fetch_kib, "true_write_kib": true_kib,
                                "fetch_factor": true_fetch_kib / copy_f,
                                "write_factor": true_kib / copy_w},
           "hbm_bytes_per_step": round(fetch + write), "fetch_bytes": round(fetch),
           "write_bytes": round(write), "algorithmic_bytes_per_step": alg_bytes,
           "traffic_over_algorithmic": (fetch + write) / alg_bytes}
    PROFILES.mkdir(exist_ok=True)
    (PROFILES / f"{name}_traffic.json").write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out, indent=1))


def batched_check(n_sims, seed, budget_updates=2.0e8):
    """Oracle run (timed: the CPU baseline) of the smallest scene of the
    batched workload, truncated to a bounded step count, and the float32 and
    float64 errors of the batched GPU stepper on that scene (DFT planes)."""
    import photonfdtd_amd as pf
    from oracle.fdtd_numpy import run_oracle
    from photonfdtd_amd.batch import run_batch
    probe = batched_2d_scenes(pf, n_sims=n_sims, steps=10, seed=seed)
    idx = int(np.argmin([s.grid.n_cells for s in probe]))
    cells = probe[idx].grid.n_cells
    steps = int(min(4000, budget_updates // cells))
    sim64 = batched_2d_scenes(pf, n_sims=n_sims, steps=steps, seed=seed, precision="float64")[idx]
    t0 = time.perf_counter()
    ref = run_oracle(sim64)
    cpu_s = time.perf_counter() - t0
    errs = {}
    for prec in ("float64", "float32"):
        sim = batched_2d_scenes(pf, n_sims=n_sims, steps=steps, seed=seed, precision=prec)[idx]
        got = run_batch([sim])[0]
        worst = 0.0
        for name, comps in ref["dft"].items():
            for c, v in comps.items():
                worst = max(worst, _rel(got.dft[name][c], v))
        errs[prec] = worst
    nx, ny, _ = sim64.grid.shape
    return dict(cpu_value=cells * sim64.n_steps / cpu_s / 1e9, cpu_s=cpu_s,
                sample=f"scene {idx} of the batch ({nx}x{ny}, the smallest), {sim64.n_steps} "
                       "steps, NumPy float64 oracle, 1 core",
                fp64_rel_l2=errs["float64"], fp32_rel_l2=errs["float32"])


def bench_batched(args):
    """Batched 2-D workload: 64 TE scenes of uneven size in one launch."""
    import photonfdtd_amd as pf
    from photonfdtd_amd.batch import SimBatch
    from photonfdtd_amd.hip_backend import device_info

    n_sims, seed = 64, 0
    check = None
    if not args.no_cpu:
        check = batched_check(n_sims, seed)
        print(f"[bench] oracle sample done: {check}", file=sys.stderr, flush=True)
    total = args.warmup + args.steps
    t_setup = time.perf_counter()
    sims = batched_2d_scenes(pf, n_sims=n_sims, steps=total, seed=seed, precision=args.precision)
    b = SimBatch(sims)
    setup_s = time.perf_counter() - t_setup
    cells = sum(s.grid.n_cells for s in sims)
    print(f"[bench] setup {setup_s:.1f} s, {n_sims} scenes, {cells} cells, {total} steps",
          file=sys.stderr, flush=True)
    print(f"[bench] persistent grid {b.persistent_grid(4)} (4 waves per workgroup)",
          file=sys.stderr, flush=True)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)
    if args.batch_config:
        m, w, r = args.batch_config.split(",")
        b.configure(m, int(w), int(r))
        tuned, trials = dict(mode=m, waves=int(w), rows=int(r)), []
    else:
        tuned = b.tune(trial_steps=min(100, total), modes=args.batch_modes.split(","), log=log)
        trials = tuned.pop("trials")
    print(f"[bench] tuned {tuned}", file=sys.stderr, flush=True)
    b.advance(args.warmup)
    t0 = time.perf_counter()
    b.advance(args.steps)
    wall = time.perf_counter() - t0
    dev_ms = b.last_ms()
    finite = all(np.isfinite(v).all() for i in range(n_sims) for v in b.dft(i, "out").values())
    # one full-field read of the largest scene: a plain copy kernel (k_gather_box) of known size
    # that calibrates the FETCH_SIZE / WRITE_SIZE counters
    big = int(np.argmax([s.grid.n_cells for s in sims]))
    b.field(big, "Ey")
    b.close()
    itemsize = 4 if args.precision == "float32" else 8
    bytes_per_update = 7 * itemsize          # TE: read Ex, Ey, Hz, ce; write Ex, Ey, Hz
    achieved = cells * args.steps * bytes_per_update / (dev_ms * 1e-3) / 1e9
    value = cells * args.steps / wall / 1e9
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
        "dtype": "f32" if itemsize == 4 else "f64",
        "data": "synthetic",
        "config": {"workload": f"batched 2-D TE: {n_sims} scenes, sides 256-1024 cells, dx 20 nm, "
                               "12-cell CPML, 1-4 PolySlabs (20% Lorentz), Ex dipole, "
                               "2 x-plane DFT monitors x (Ey, Hz) x 16 freqs",
                   "cells": cells, "seed": seed, "launch": tuned, "tune_trials": trials,
                   "device": device_info()[0]},
        "roofline": {"bound": "hbm", "kernel": f"k_batch_march ({tuned['mode']})",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (measured_traffic(f"batched2d_te_{args.precision}")
                                 if tuned["mode"] != "persistent" else None),
                     "bytes_per_cell_update": bytes_per_update,
                     "kernel_ms_per_step": round(dev_ms / args.steps, 4)},
        "march": {"psi_free_cell_fraction": round(info["cells_bulk"] / cells, 4),
                  "tiles_psi_free": info["tiles_bulk"], "tiles_cpml": info["tiles_cpml"],
                  "cpml_lists": info["cpml_lists"], "zrow": info["zrow"],
                  "dft_ring": info["dft_ring"]},
        "dft_finite": bool(finite),
        "setup_s": round(setup_s, 1),
    }
    if cpu is not None:
        out["cpu_baseline"] = {"value": round(cpu["value"], 5), "unit": "Gcell-updates/s",
                               "cores": cpu["threads"], "kind": "port", "sample": cpu["sample"]}
    if check is not None:
        out["numpy_oracle_baseline"] = {"value": round(check["cpu_value"], 5),
                                        "unit": "Gcell-updates/s", "cores": 1,
                                        "sample": check["sample"]}
        out["fp64_rel_l2_vs_oracle"] = check["fp64_rel_l2"]
        out["fp32_rel_l2_vs_oracle"] = check["fp32_rel_l2"]
    print(json.dumps(out), flush=True)


def batched_3d_scenes(pf, n_sims=16, lo=64, hi=128, steps=220, precision="float32", seed=16):
    """Independent 3-D scenes of uneven size for batched stepping.

    Node counts per axis uniform in [lo, hi], dx = 40 nm, CPML of 5 / 4 / 6
    cells on x / y / z, Courant number 0.9. Each scene holds a dielectric box
    (eps uniform in [2, 6]), a Drude box and a two-pole Lorentz box placed by
    the seed, an Ez and an Ex dipole (Gaussian pulse at 1.2 um, 4 fs FWHM),
    and DFT planes on x (Ex, Ez, Hy at 2 frequencies) and y (Hz, Ey at 1).
    """
    rng = np.random.default_rng(seed)
    dx = 40e-9
    out = []
    for _ in range(n_sims):
        shape = rng.integers(lo, hi + 1, 3)
        L = shape * dx
        grid = pf.Grid(size=tuple(L.tolist()), cell_size=dx, pml_layers=(5, 4, 6))

        def box(frac, medium):
            c = (rng.uniform(-0.15, 0.15, 3) * L).tolist()
            return pf.Box(center=tuple(c), size=tuple((np.asarray(frac) * L).tolist()),
                          medium=medium)

        st = [box((0.5, 0.4, 0.5), pf.Medium(float(rng.uniform(2.0, 6.0)))),
              box((0.2, 0.2, 0.15),
                  pf.DispersiveMedium.drude(1.0, f_plasma_Hz=80e12, f_collision_Hz=5e12)),
              box((0.15, 0.2, 0.2),
                  pf.DispersiveMedium.lorentz(2.0, [(1.5, 150e12, 5e12), (0.5, 300e12, 10e12)]))]
        wave = pf.GaussianPulse(freq0=pf.C_0 / 1.2e-6, fwhm=4e-15)
        p = (rng.uniform(-0.2, 0.2, 3) * L).tolist()
        src = [pf.PointDipole(position=tuple(p), component="Ez", waveform=wave),
               pf.PointDipole(position=(-p[0], p[1], -p[2]), component="Ex", waveform=wave,
                              amplitude=0.5)]
        mon = [pf.DFTMonitor(name="d", components=("Ex", "Ez", "Hy"),
                             freqs=[pf.C_0 / 1.2e-6, pf.C_0 / 1.0e-6], plane_axis="x",
                             plane_position=0.1 * L[0]),
               pf.DFTMonitor(name="y", components=("Hz", "Ey"), freqs=[pf.C_0 / 1.1e-6],
                             plane_axis="y", plane_position=-0.1 * L[1])]
        dt = pf.Simulation(grid, run_time=1e-15, precision=precision, backend="hip",
                           courant=0.9).dt
        out.append(pf.Simulation(grid, structures=st, sources=src, monitors=mon,
                                 run_time=(steps - 0.5) * dt, precision=precision, backend="hip",
                                 courant=0.9))   # exactly `steps` steps
    return out


def bench_stream3d(args):
    """Beyond-HBM streaming: the dielectric slab at n^3 kept in pinned host
    memory, x-slab tiles of --tile-rows planes with a 32-plane halo advanced
    32 steps per visit; against the device-resident stepper on the same scene
    (same bits required)."""
    import photonfdtd_amd as pf
    from photonfdtd_amd.hip_backend import HipEngine, device_info, exact_mode
    os.environ.setdefault("PHOTONFDTD_HIP_TB", "32")
    os.environ["PHOTONFDTD_HIP_TILE"] = str(args.tile_rows)
    total = args.warmup + args.steps
    npdt = np.float32 if args.precision == "float32" else np.float64
    out = {}
    for mode in ("streaming", "resident"):
        sim = dielectric_slab(pf, args.n, total, args.precision)
        eng = HipEngine(sim, npdt, exact_mode(npdt), streaming=mode == "streaming")
        eng.advance(0, args.warmup)
        t0 = time.perf_counter()
        eng.advance(args.warmup, args.steps)
        eng.dft_flush()
        wall = time.perf_counter() - t0
        info = eng.info()
        dft = eng.dft("out")
        eng.close()
        out[mode] = dict(wall=wall, info=info, dft=dft)
        print(f"[bench] {mode}: {wall:.2f} s for {args.steps} steps", file=sys.stderr, flush=True)
    cells = args.n ** 3
    item = np.dtype(npdt).itemsize
    T = int(os.environ["PHOTONFDTD_HIP_TB"])
    tiles = out["streaming"]["info"]["tiles_bulk"]
    # per chunk of T steps: every tile uploads its planes plus halo (six
    # fields and ce) and downloads its owned planes (six fields); the CPML
    # psi rows, a few percent more, are left out
    planes_up = args.n + 2 * T * (tiles - 1)
    link = (planes_up * 7 + args.n * 6) * args.n * args.n * item / T
    same = all(np.array_equal(out["streaming"]["dft"][c], out["resident"]["dft"][c])
               for c in out["resident"]["dft"])
    value = cells * args.steps / out["streaming"]["wall"] / 1e9
    line = {"metric": "Gcell-updates/s (beyond-HBM streaming)", "value": round(value, 3),
            "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(out["streaming"]["wall"] / args.steps * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32" if npdt == np.float32 else "f64", "data": "synthetic",
            "config": {"workload": f"3-D dielectric slab {args.n}^3 streamed from pinned host "
                                   f"memory, {args.tile_rows}-plane tiles, {T}-step visits",
                       "tiles": tiles, "t_block": T, "device": device_info()[0]},
            "host_link": {"bytes_per_step": round(link),
                          "GB_per_s": round(link * args.steps / out["streaming"]["wall"] / 1e9, 2)},
            "resident_value": round(cells * args.steps / out["resident"]["wall"] / 1e9, 3),
            "dft_bitwise_equal_to_resident": bool(same)}
    print(json.dumps(line), flush=True)


def bench_batched3d(args):
    """Batched 3-D workload: 16 uneven scenes of 64-128 nodes per axis in one
    batch, against the same scenes stepped one after another."""
    import photonfdtd_amd as pf
    from photonfdtd_amd.batch import SimBatch
    from photonfdtd_amd.hip_backend import HipEngine, device_info, exact_mode

    total = args.warmup + args.steps
    sims = batched_3d_scenes(pf, steps=total, precision=args.precision)
    cells = sum(int(np.prod(s.grid.shape)) for s in sims)
    npdt = np.float32 if args.precision == "float32" else np.float64
    # the same scenes, one stepper each, run one after another
    single = 0.0
    for s in sims:
        eng = HipEngine(s, npdt, exact_mode(npdt))
        eng.advance(0, args.warmup)
        t0 = time.perf_counter()
        eng.advance(args.warmup, args.steps)
        single += time.perf_counter() - t0
        eng.close()
    b = SimBatch(sims)
    tuned = b.tune(trial_steps=min(50, total))
    trials = tuned.pop("trials")
    b.advance(args.warmup)
    t0 = time.perf_counter()
    b.advance(args.steps)
    wall = time.perf_counter() - t0
    dev_ms = b.last_ms()
    finite = all(np.isfinite(v).all() for i in range(len(sims)) for v in b.dft(i, "d").values())
    b.close()
    out = {
        "metric": METRIC, "value": round(cells * args.steps / wall / 1e9, 3),
        "unit": "Gcell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if npdt == np.float32 else "f64", "data": "synthetic",
        "config": {"workload": f"batched 3-D: {len(sims)} scenes, 64-128 nodes per axis, dx 40 nm, "
                               "CPML 5/4/6, dielectric + Drude + Lorentz boxes, Ez + Ex dipoles, "
                               "DFT planes on x and y",
                   "cells": cells, "launch": tuned, "tune_trials": trials,
                   "device": device_info()[0]},
        "batch_device_ms_per_step": round(dev_ms / args.steps, 4),
        "single_runs_ms_per_step": round(single / args.steps * 1e3, 4),
        "speedup_vs_single_runs": round(single / wall, 3),
        "dft_finite": bool(finite),
    }
    print(json.dumps(out), flush=True)


_ADJ_MODE = {}


def adjoint3d_guide_mode(pf):
    """Fundamental mode of the 500 x 220 nm silicon core in silica at 1.55 um,
    solved on the 2.0 x 1.2 um source window (100 x 60 nodes of 20 nm),
    cached: (ModeSolver, ModeResult)."""
    if "m" not in _ADJ_MODE:
        core = pf.Box(center=(0.0, 0.0), size=(500e-9, 220e-9), medium=pf.Medium(12.1))
        ms = pf.ModeSolver(size=(2.0e-6, 1.2e-6), cell_size=DX, structures=[core],
                           wavelength=1.55e-6, background_eps=2.085, num_modes=1,
                           polarization="te")
        _ADJ_MODE["m"] = (ms, ms.solve())
    return _ADJ_MODE["m"]


def adjoint3d_scene(pf, precision="float32", steps=3000, rho=None):
    """The 3-D adjoint inverse-design scene: 256 x 256 x 128 cells, dx 20 nm,
    12-cell CPML, silica background, a 500 x 220 nm Si core along x, a
    64 x 64 x 8-cell design region at the centre (eps = 2.085 + rho 10.015),
    the guide's solved fundamental TE mode (Ey profile) as a ModeSource at
    x = -1.6 um and an x-plane DFT monitor (Ey, Ez, Hy, Hz; 8 frequencies in
    c/1.6..c/1.5 um) at x = +1.6 um."""
    shape = (256, 256, 128)
    grid = pf.Grid(size=tuple(n * DX for n in shape), cell_size=DX, pml_layers=(12, 12, 12))
    L = [n * DX for n in shape]
    structures = [pf.Box(center=(0, 0, 0), size=tuple(3 * x for x in L), medium=pf.Medium(2.085)),
                  pf.Box(center=(0, 0, 0), size=(3 * L[0], 500e-9, 220e-9), medium=pf.Medium(12.1))]
    ms, mr = adjoint3d_guide_mode(pf)
    src = pf.ModeSource(center=(-1.6e-6, 0, 0), size=(0, 2.0e-6, 1.2e-6), component="Ey",
                        waveform=pf.GaussianPulse(freq0=pf.C_0 / 1.55e-6, fwhm=10e-15),
                        profile=np.real(mr.Ey[0]), profile_coords=(ms.y, ms.z))
    mon = pf.DFTMonitor(name="out", components=("Ey", "Ez", "Hy", "Hz"),
                        freqs=np.linspace(pf.C_0 / 1.6e-6, pf.C_0 / 1.5e-6, 8), plane_axis="x",
                        plane_position=1.6e-6)
    probe = pf.Simulation(grid, structures=structures, sources=[src], monitors=[mon],
                          run_time=1e-15, precision=precision, backend="hip")
    sim = pf.Simulation(grid, structures=structures, sources=[src], monitors=[mon],
                        run_time=(steps - 0.5) * probe.dt, precision=precision, backend="hip")
    if rho is not None:
        eps = np.array(sim.eps_r, dtype=np.float64)
        c = [n // 2 for n in shape]
        eps[c[0] - 32:c[0] + 32, c[1] - 32:c[1] + 32, c[2] - 4:c[2] + 4] = 2.085 + rho * (12.1 - 2.085)
        sim.eps_r = eps.astype(sim.dtypes["eps_r"])
    return sim


def adjoint3d_mode(sim):
    """The overlap target on the monitor plane: the solved mode's tangential
    fields (Ey, Ez, Hy, Hz) placed in the 2.0 x 1.2 um window at the plane's
    centre, zero outside."""
    import photonfdtd_amd as pf
    ms, mr = adjoint3d_guide_mode(pf)
    ny, nz = sim.grid.shape[1], sim.grid.shape[2]
    oy, oz = (ny - ms.ny) // 2, (nz - ms.nz) // 2
    out = {}
    for c in ("Ey", "Ez", "Hy", "Hz"):
        plane = np.zeros((ny, nz), dtype=np.complex128)
        plane[oy:oy + ms.ny, oz:oz + ms.nz] = getattr(mr, c)[0]
        out[c] = plane
    return out


def bench_adjoint3d(args):
    """Forward + adjoint gradient of a mode-overlap objective at 256x256x128,
    float32, checkpoints every 100 steps; gradient checked in float64 against
    central differences on random design voxels."""
    import torch
    import photonfdtd_amd as pf
    from photonfdtd_amd.adjoint import value_and_grad_eps
    from photonfdtd_amd.hip_backend import device_info
    from photonfdtd_amd.smatrix import mode_amplitudes

    def log(msg):
        print(msg, file=sys.stderr, flush=True)
    steps = 3000
    rho = np.random.default_rng(0).uniform(0, 1, (64, 64, 8))
    sim0 = adjoint3d_scene(pf, "float32", steps, rho)
    mode = adjoint3d_mode(sim0)
    dA = DX * DX

    def loss_t(out):
        f = out["dft"]["out"]
        m = {k: torch.tensor(v) for k, v in mode.items()}
        ap, _ = mode_amplitudes(f, m, dA)
        return (ap.abs() ** 2).sum()

    def loss_np(res):
        ap, _ = mode_amplitudes(res.dft["out"], mode, dA)
        return float((np.abs(ap) ** 2).sum())

    c = [n // 2 for n in sim0.grid.shape]
    win = (slice(c[0] - 32, c[0] + 32), slice(c[1] - 32, c[1] + 32), slice(c[2] - 4, c[2] + 4))
    cells = sim0.grid.n_cells
    t0 = time.perf_counter()
    v32, g32 = value_and_grad_eps(sim0, loss_t, checkpoint_interval=100)
    wall32 = time.perf_counter() - t0
    work = 3 * cells * steps          # forward, checkpoint recompute, reverse step
    log(f"[adjoint3d] float32 value {v32:.6e}, forward+adjoint {wall32:.2f} s "
        f"({work / wall32 / 1e9:.1f} Gcell-updates/s)")
    out = {
        "metric": METRIC,
        "value": round(work / wall32 / 1e9, 3),
        "unit": "Gcell-updates/s (forward + checkpoint recompute + adjoint)",
        "n_gpus": 1, "steps": steps, "warmup": 0, "ms_per_step": round(wall32 / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "3-D adjoint inverse design 256x256x128, dx 20 nm, 12-cell CPML, "
                               "64x64x8 design (rho ~ U(0,1), seed 0), Ey Gaussian ModeSource, "
                               "mode-overlap objective at 8 frequencies, checkpoint every 100",
                   "cells": cells, "device": device_info()[0]},
        "gradient_wall_s": round(wall32, 3),
    }
    if args.fd_voxels > 0:
        t1 = time.perf_counter()
        sim64 = adjoint3d_scene(pf, "float64", steps, rho)
        v64, g64 = value_and_grad_eps(sim64, loss_t, checkpoint_interval=100)
        gd64 = g64[win] * (12.1 - 2.085)
        gd32 = g32[win] * (12.1 - 2.085)
        rel32 = float(np.linalg.norm(gd32 - gd64) / np.linalg.norm(gd64))
        log(f"[adjoint3d] float64 value {v64:.6e}; float32 vs float64 design gradient rel L2 "
            f"{rel32:.2e}")
        vox = np.random.default_rng(1).integers(0, [64, 64, 8], size=(args.fd_voxels, 3))
        h = 1e-3
        errs = []
        for v in vox:
            vals = []
            for sgn in (1, -1):
                r = rho.copy()
                r[tuple(v)] += sgn * h
                vals.append(loss_np(adjoint3d_scene(pf, "float64", steps, r).run()))
            fd = (vals[0] - vals[1]) / (2 * h)
            ad = float(gd64[tuple(v)])
            errs.append(abs(fd - ad) / max(abs(fd), 1e-300))
            log(f"[fd] voxel {tuple(int(x) for x in v)}: adjoint {ad:.6e} fd {fd:.6e} "
                f"rel {errs[-1]:.1e}")
        out["fd_check"] = {"voxels": int(args.fd_voxels), "h_rho": h, "precision": "float64",
                           "max_rel_err": max(errs), "median_rel_err": float(np.median(errs)),
                           "float32_vs_float64_grad_rel_l2": rel32,
                           "wall_s": round(time.perf_counter() - t1, 1)}
    print(json.dumps(out), flush=True)


def bench_surrogate(args):
    """Surrogate-guided inverse design against plain adjoint design (2-D TM,
    512^2 domain, 128^2 design region, float32)."""
    import torch
    import photonfdtd_amd as pf
    from photonfdtd_amd import design2d as D
    from photonfdtd_amd.hip_backend import device_info
    from photonfdtd_amd.surrogate import SpectralSurrogate

    def log(msg):
        print(msg, file=sys.stderr, flush=True)
    steps = 4000
    n_train, n_test = args.samples, 256
    t0 = time.perf_counter()
    ref = D.reference_flux(pf, steps)
    lab = D.Labeler(pf, ref, batch=256, steps=steps)
    log(f"[surrogate] setup {time.perf_counter() - t0:.1f} s")
    X = D.training_designs(n_train + n_test)
    t1 = time.perf_counter()
    parts = []
    for c0 in range(0, len(X), 2048):        # progress every 2048 designs
        parts.append(lab(X[c0:c0 + 2048]))
        log(f"[surrogate] labelled {min(c0 + 2048, len(X))}/{len(X)} designs "
            f"({time.perf_counter() - t1:.1f} s)")
    Y = np.concatenate(parts)
    t_label = time.perf_counter() - t1
    cells = lab.sim.grid.n_cells
    log(f"[surrogate] {len(X)} labels in {t_label:.1f} s "
        f"({len(X) / t_label:.1f} designs/s, {len(X) * cells * steps / t_label / 1e9:.1f} "
        f"Gcell-updates/s); mean T {Y.mean():.3f}")
    net = SpectralSurrogate(seed=0)
    t2 = time.perf_counter()
    losses = D.train(net, X[:n_train], Y[:n_train], epochs=args.epochs, batch=64, lr=1e-3,
                     log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t2
    P = D.predict(net, X[n_train:])
    Yt = Y[n_train:]
    rel_spec = float(np.linalg.norm(P - Yt) / np.linalg.norm(Yt))
    base_spec = float(np.linalg.norm(Y[:n_train].mean(0)[None] - Yt) / np.linalg.norm(Yt))
    jp, jt = D.objective(P), D.objective(Yt)
    rel_obj = float(np.mean(np.abs(jp - jt) / np.abs(jt)))
    base_obj = float(np.mean(np.abs(D.objective(Y[:n_train]).mean() - jt) / np.abs(jt)))
    corr = float(np.corrcoef(jp, jt)[0, 1])
    log(f"[surrogate] held-out spectrum rel L2 {rel_spec:.3f} (mean predictor {base_spec:.3f}); "
        f"objective rel err {rel_obj:.3f} (mean predictor {base_obj:.3f}), corr {corr:.3f}")
    rho0 = D.random_design(10 ** 6)
    t3 = time.perf_counter()
    plain = D.adjoint_design(pf, ref, rho0, iters=args.adjoint_iters, steps=steps, log=log)
    t_plain = time.perf_counter() - t3
    hist = plain["history"]
    target = max(h["objective"] for h in hist)
    t_plain_target = next(h["wall_s"] for h in hist if h["objective"] >= target)
    # the adjoint's objective at iteration k is verified when iteration k runs; the wall time to
    # reach it includes the forward and adjoint of that iteration
    it_target = next(h["iter"] for h in hist if h["objective"] >= target)
    t_plain_target = t_plain * (it_target + 1) / len(hist)
    # candidates per round are verified in one batch of their own size
    n_cand = args.candidates
    lab_v = D.Labeler(pf, ref, batch=n_cand, steps=steps)
    t4 = time.perf_counter()
    sur = D.surrogate_design(pf, net, lab_v, ref, rho0, rounds=args.rounds, target=target,
                             candidates=n_cand, pool=(X[:n_train], Y[:n_train]), steps=steps,
                             log=log)
    t_sur = time.perf_counter() - t4
    reached = sur["objective"] >= target
    # ablation without the network: the plain adjoint started from the best
    # labelled training design (the surrogate loop's best known start)
    kb = int(np.argmax(D.objective(Y[:n_train])))
    t5 = time.perf_counter()
    abl = D.adjoint_design(pf, ref, X[kb], iters=args.adjoint_iters, steps=steps, target=target)
    t_abl = time.perf_counter() - t5
    abl_reached = abl["history"][-1]["objective"] >= target
    log(f"[ablation] adjoint from the best training design (J {float(D.objective(Y[kb])[0]):.4f}): "
        f"J {abl['history'][-1]['objective']:.4f} after {len(abl['history'])} iterations, "
        f"{t_abl:.1f} s")
    lab.close()
    lab_v.close()
    plain_rate = len(hist) / t_plain
    sur_rate = (len(sur["history"]) - 1) * n_cand / max(t_sur, 1e-9)
    out = {
        "metric": METRIC,
        "value": round(t_plain_target / t_sur, 3) if reached else 0.0,
        "unit": "x less wall time to the same verified FDTD objective (surrogate vs plain adjoint)",
        "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 network / f32 FDTD", "data": "synthetic",
        "config": {"workload": "2-D TM 512^2 (dx 20 nm, 12-cell CPML), 128^2 design region "
                               "between an input and an 0.8 um-offset output Si guide, "
                               "16-frequency transmission 1.4-1.7 um",
                   "train_samples": n_train, "test_samples": n_test, "epochs": args.epochs,
                   "network_params": net.n_params(), "device": device_info()[0]},
        "verified_design_iters_per_s": {"surrogate": round(sur_rate, 2),
                                        "plain_adjoint": round(plain_rate, 3)},
        "target_objective": round(target, 4),
        "plain_adjoint": {"iters": len(hist), "wall_s_to_target": round(t_plain_target, 2),
                          "objective_first": round(hist[0]["objective"], 4)},
        "ablation_adjoint_from_best_training_design": {
            "start_objective": round(float(D.objective(Y[kb])[0]), 4),
            "iters": len(abl["history"]), "wall_s": round(t_abl, 2),
            "objective": round(abl["history"][-1]["objective"], 4),
            "reached_target": bool(abl_reached),
            "speedup_vs_plain": round(t_plain_target / t_abl, 3) if abl_reached else 0.0},
        "surrogate": {"rounds": len(sur["history"]) - 1, "wall_s": round(t_sur, 2),
                      "objective": round(sur["objective"], 4), "reached_target": bool(reached),
                      "fallbacks": sum(h.get("fallback", False) for h in sur["history"])},
        "dataset": {"designs": len(X), "label_s": round(t_label, 2),
                    "gcell_updates_per_s": round(len(X) * cells * steps / t_label / 1e9, 2)},
        "training": {"wall_s": round(t_train, 2), "final_mse": losses[-1],
                     "heldout_spectrum_rel_l2": rel_spec, "mean_predictor_rel_l2": base_spec,
                     "heldout_objective_rel_err": rel_obj, "mean_predictor_objective_rel_err": base_obj,
                     "heldout_objective_correlation": corr},
    }
    print(json.dumps(out), flush=True)


def bench_nn_gemm(args):
    """Training GEMMs of the surrogate network: every layer's forward, data-
    gradient and weight-gradient pass at the training batch size, timed with
    HIP events on the stream the kernels run on. FLOPs are the convolution's
    own (2 x output pixels x cout x (taps x cin) per pass), not the zero taps
    the data-gradient GEMM of a stride-2 layer multiplies."""
    import ctypes as C
    import torch
    from photonfdtd_amd import _native as N
    from photonfdtd_amd.surrogate import layer_table
    lib = N.load_library()
    dev = torch.device("cuda")
    B = args.nn_batch
    reps = 20
    st = torch.cuda.current_stream()
    sp = C.c_void_p(st.cuda_stream)
    ptr = lambda t: C.c_void_p(t.data_ptr())
    layers, h = [], 128
    table = layer_table(16)
    rows, tot_flops, tot_ms = [], 0.0, {"forward": 0.0, "data_grad": 0.0, "weight_grad": 0.0}
    input_grad_us = 0.0
    g = torch.Generator(device="cpu").manual_seed(0)
    for li, (ci, co, k, s, relu) in enumerate(table):
        if li == len(table) - 2:
            h = 1                               # head: 1x1 layers on a 1x1 map
        ho = (h - 1) // s + 1
        d = N.PfNnConv()
        d.n, d.h, d.w, d.cin, d.cout, d.ksize, d.stride, d.relu = B, h, h, ci, co, k, s, relu
        d.out_f32 = int(li == len(table) - 1)
        x = torch.randn(B, h, h, ci, generator=g).to(torch.bfloat16).to(dev)
        w = (torch.randn(co, k, k, ci, generator=g) / np.sqrt(k * k * ci)).to(torch.bfloat16).to(dev)
        b = torch.zeros(co, device=dev)
        y = torch.empty(B, ho, ho, co, dtype=torch.float32 if d.out_f32 else torch.bfloat16, device=dev)
        dy = torch.randn(B, ho, ho, co, generator=g).to(dev)
        dx = torch.empty(B, h, h, ci, device=dev)
        dw = torch.zeros(co, k, k, ci, device=dev)
        flops = 2.0 * B * ho * ho * co * k * k * ci
        times = {}
        for name, fn in (
                ("forward", lambda: lib.pf_nn_conv_forward(C.byref(d), ptr(x), ptr(w), ptr(b), ptr(y), sp)),
                ("data_grad", lambda: lib.pf_nn_conv_backward(C.byref(d), ptr(x), ptr(w), ptr(y), ptr(dy),
                                                             ptr(dx), None, None, sp)),
                ("weight_grad", lambda: lib.pf_nn_conv_backward(C.byref(d), ptr(x), ptr(w), ptr(y), ptr(dy),
                                                               None, ptr(dw), None, sp))):
            if li == 0 and name == "data_grad":
                # the input map needs no gradient in training (the design
                # loop's ascent on the network takes it): timed, not counted
                for _ in range(3):
                    assert fn() == 0
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(st)
                for _ in range(reps):
                    fn()
                e1.record(st)
                e1.synchronize()
                input_grad_us = e0.elapsed_time(e1) / reps * 1e3
                continue
            for _ in range(3):
                assert fn() == 0
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(reps):
                fn()
            e1.record(st)
            e1.synchronize()
            times[name] = e0.elapsed_time(e1) / reps
            tot_ms[name] += times[name]
            tot_flops += flops
        rows.append({"layer": li, "shape": [B, h, h, ci, co, k, s],
                     "gflop_per_pass": round(flops / 1e9, 3),
                     **{f"{nm}_us": round(t * 1e3, 1) for nm, t in times.items()},
                     **{f"{nm}_tflops": round(flops / (t * 1e-3) / 1e12, 1) for nm, t in times.items()}})
        h = ho
    total_ms = sum(tot_ms.values())
    out = {"metric": "surrogate training GEMM TFLOP/s (bf16 MFMA, FP32 accumulation)",
           "value": round(tot_flops / (total_ms * 1e-3) / 1e12, 2), "unit": "TFLOP/s",
           "batch": B, "ms_per_training_pass": round(total_ms, 3),
           "per_pass_ms": {k: round(v, 3) for k, v in tot_ms.items()},
           "input_grad_us": round(input_grad_us, 1),
           "roofline": {"bound": "mfma", "peak": 2500.0, "unit": "TFLOP/s",
                        "frac": round(tot_flops / (total_ms * 1e-3) / 1e12 / 2500.0, 4)},
           "layers": rows}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--samples", type=int, default=4096, help="surrogate training designs")
    ap.add_argument("--epochs", type=int, default=12, help="surrogate training epochs")
    ap.add_argument("--adjoint-iters", type=int, default=12, help="plain adjoint iterations")
    ap.add_argument("--rounds", type=int, default=6, help="surrogate design rounds")
    ap.add_argument("--candidates", type=int, default=16,
                    help="surrogate: designs verified per round (one batched FDTD run)")
    ap.add_argument("--fd-voxels", type=int, default=20, help="adjoint3d: central-difference voxels")
    ap.add_argument("--workload", default="slab3d",
                    choices=("slab3d", "batched2d", "batched3d", "surrogate", "adjoint3d",
                             "stream3d", "nn_gemm"),
                    help="slab3d: the 3-D dispersive slab (headline); batched2d: 64 uneven "
                         "2-D TE scenes in one launch; batched3d: 16 uneven 3-D scenes in one "
                         "batch")
    ap.add_argument("--nn-batch", type=int, default=256,
                    help="nn_gemm: design maps per training batch")
    ap.add_argument("--tile-rows", type=int, default=64,
                    help="stream3d: owned x planes per streamed tile")
    ap.add_argument("--batch-config", default=None, metavar="MODE,WAVES,ROWS",
                    help="fixed batched launch setting instead of the autotuner")
    ap.add_argument("--batch-modes", default="graph,stream,persistent",
                    help="launch modes the batched autotuner may pick from")
    ap.add_argument("--precision", default="float32", choices=("float32", "float64"))
    ap.add_argument("--timing-every", type=int, default=10,
                    help="slab3d: HIP events around the march and aux groups of every k-th timed "
                         "step (1: every step; each event pair costs ~17 us per step)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle leg")
    ap.add_argument("--check-steps", type=int, default=2000,
                    help="steps of the full-size float32-vs-float64 check (0: skip)")
    ap.add_argument("--check-exact32", action="store_true",
                    help="run the float32 side of that check in the reference operation order")
    ap.add_argument("--summarize", nargs=2, metavar=("DIR", "KIND"),
                    help="print a summary of rocprofv3 CSV output (KIND: stats | pmc)")
    ap.add_argument("--traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="derive HBM bytes per step launch from two rocprofv3 --pmc runs")
    args = ap.parse_args()
    if args.summarize:
        summarize(*args.summarize)
        return
    if args.traffic:
        item = 4 if args.precision == "float32" else 8
        if args.workload == "batched2d":
            import photonfdtd_amd as pf
            sims = batched_2d_scenes(pf, n_sims=64, steps=10, seed=0)
            cells = [s.grid.n_cells for s in sims]
            big = sims[cells.index(max(cells))].grid.shape
            # batched fields use padded y rows (64 slots per 62 columns): the
            # calibrating gather of the largest scene reads whole padded rows
            yrow = 64 * ((big[1] + 61) // 62)
            # one launch of the TE march variant covers every scene of this workload
            traffic_from_counters(*args.traffic, kernel="k_batch_march", copy_cells=max(cells),
                                  alg_bytes=sum(cells) * 7 * item, itemsize=item,
                                  name=f"batched2d_te_{args.precision}",
                                  workload=f"batched 2-D TE, 64 scenes, {sum(cells)} cells, "
                                           f"{args.precision}",
                                  copy_fetch_cells=big[0] * yrow)
        else:
            n = args.n
            # the march keeps z rows of a multiple of 64 slots: the
            # calibrating gather reads whole rows
            zrow = 64 * ((n + 63) // 64)
            traffic_from_counters(*args.traffic, kernel="k_yee_m64", copy_cells=n ** 3,
                                  alg_bytes=n ** 3 * 13 * item, itemsize=item,
                                  name=f"dispersive_slab_{n}_{args.precision}",
                                  workload=f"3-D dispersive slab {n}^3 {args.precision}",
                                  copy_fetch_cells=n * n * zrow)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this stepper runs on one GPU; runs on several GPUs are "
                                   "not supported (how the grid would shard is described in "
                                   "the design notes)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "batched2d":
        bench_batched(args)
        return
    if args.workload == "batched3d":
        bench_batched3d(args)
        return
    if args.workload == "surrogate":
        bench_surrogate(args)
        return
    if args.workload == "adjoint3d":
        bench_adjoint3d(args)
        return
    if args.workload == "stream3d":
        bench_stream3d(args)
        return
    if args.workload == "nn_gemm":
        bench_nn_gemm(args)
        return

    import photonfdtd_amd as pf
    from photonfdtd_amd import _native as N
    from photonfdtd_amd.hip_backend import HipEngine, device_info, exact_mode

    check = cpu = None
    if not args.no_cpu:
        check = parity_and_cpu()
        print(f"[bench] oracle sample done: {check}", file=sys.stderr, flush=True)
        cpu = cpu_stepper_sample(args.n)
        print(f"[bench] compiled CPU sample done: {cpu}", file=sys.stderr, flush=True)

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
    N.check(lib, lib.pf_timing_enable(eng.h, max(1, args.timing_every)), "pf_timing_enable")
    t0 = time.perf_counter()
    eng.advance(args.warmup, args.steps)
    # fold the DFT samples still buffered, so the timed region holds the
    # float64 accumulation of every timed step whatever --steps is
    eng.dft_flush()
    wall = time.perf_counter() - t0
    ms = np.zeros(len(N.KERNEL_CLASSES), dtype=np.float64)
    cnt = np.zeros(len(N.KERNEL_CLASSES), dtype=np.int64)
    N.check(lib, lib.pf_timing_read(eng.h, N.as_ptr(ms, N.C.c_double), N.as_ptr(cnt, N.C.c_int64)),
            "pf_timing_read")
    # sanity: the DFT accumulators must be finite
    finite = all(np.isfinite(v).all() for v in eng.dft("out").values())
    # one full-field gather: a plain copy kernel of known size (n cells read,
    # n cells written) that calibrates the FETCH_SIZE / WRITE_SIZE counters
    field_max = float(np.abs(eng.get("Ey")).max())
    info = eng.info()
    eng.close()
    full = None
    if args.check_steps > 0 and args.precision == "float32":
        full = fullsize_fp32_error(args.n, args.check_steps, args.check_exact32)
        print(f"[bench] full-size float32 vs float64: {full}", file=sys.stderr, flush=True)

    cells = sim.grid.n_cells
    value = cells * args.steps / wall / 1e9
    itemsize = np.dtype(npdt).itemsize
    bytes_per_update = 13 * itemsize            # read 6 fields + ce, write 6 fields
    k = dict(zip(N.KERNEL_CLASSES, zip(ms.tolist(), cnt.tolist())))
    if cnt[4] > 0:
        # two-step launches (plus a single step when the count is odd)
        step_ms = (ms[4] + ms[2]) / (2 * cnt[4] + cnt[2])
        kernel = "yee_pair"
    elif cnt[2] > 0:
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
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(f"dispersive_slab_{args.n}_{args.precision}"),
                     "bytes_per_cell_update": bytes_per_update,
                     "kernel_ms_per_step": round(step_ms, 4),
                     "timed_steps": f"HIP events on every {max(1, args.timing_every)}-th timed step"},
        "kernel_ms": {name: {"total_ms": round(t, 3), "launch_groups": int(c)}
                      for name, (t, c) in k.items()},
        "hbm_fraction_of_value": round(value * 1e9 * bytes_per_update / 1e9 / HBM_PEAK_GBS, 4),
        "march": {"psi_free_cell_fraction": round(info["cells_bulk"] / cells, 4),
                  "tiles_psi_free": info["tiles_bulk"], "tiles_cpml": info["tiles_cpml"],
                  "cpml_lists": info["cpml_lists"], "zrow": info["zrow"],
                  "dft_ring": info["dft_ring"]},
        "dft_finite": bool(finite),
        "ey_max_final": field_max,
        "setup_s": round(setup_s, 1),
    }
    if cpu is not None:
        out["cpu_baseline"] = {"value": round(cpu["value"], 5), "unit": "Gcell-updates/s",
                               "cores": cpu["threads"], "kind": "port", "sample": cpu["sample"]}
    if check is not None:
        out["numpy_oracle_baseline"] = {"value": round(check["cpu_value"], 5),
                                        "unit": "Gcell-updates/s", "cores": 1,
                                        "sample": check["sample"]}
        out["fp64_rel_l2_vs_oracle"] = check["fp64_rel_l2"]
        out["fp32_rel_l2_vs_oracle"] = check["fp32_rel_l2"]
        out["reference_order_fp32_rel_l2_vs_oracle"] = check["ref_fp32_rel_l2"]
    if full is not None:
        out["fullsize_fp32_vs_fp64"] = {
            "rel_l2": full["rel_l2"], "dft_planes_rel_l2": full["dft_rel_l2"],
            "final_fields_rel_l2": full["fields_rel_l2"], "worst_array": full["worst_array"],
            "check": (f"{full['cells']} cells x {full['steps']} steps, float32 "
                      f"{'reference' if args.check_exact32 else 'fast'} order vs float64 "
                      "reference order, both on the GPU; final six fields and both DFT planes")}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
