"""Write the golden fixtures from the reference's own NumPy float64 path.

Run in a container that holds the reference checkout (read-only, never
copied): ``python tests/golden/make_golden.py [/path/to/reference]
[--only=scene,scene]``.
For every scene in ``scenes.py`` this imports ``photonfdtd`` from the
reference's ``src/`` (no bytecode written), runs it with ``backend="numpy"``
and stores inputs-derived tables and outputs as ``<scene>.npz`` (float64) and
``<scene>_f32.npz`` (float32 scenes). Only data is written: monitor arrays,
DFT accumulators, flux totals, dt / n_steps, CPML tables, the stamped
permittivity and the expanded source list.
"""
import os
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
from scenes import F32_SCENES, FEATURE_SCENES, MODE_CASES, SCENES  # noqa: E402


def dump(sim, res):
    out = {"dt": np.float64(sim.dt), "n_steps": np.int64(sim.n_steps),
           "eps_r": np.asarray(sim.eps_r), "times_len": np.int64(res.times.size)}
    if getattr(sim, "subpixel", False):
        for c, e in zip(("xx", "yy", "zz"), sim._smoothed_eps_components()):
            out[f"eps_{c}"] = np.asarray(e)
    for a in range(3):
        out[f"cpml_b_e{a}"] = np.asarray(sim._b_e[a])
        out[f"cpml_c_e{a}"] = np.asarray(sim._c_e[a])
        out[f"cpml_b_h{a}"] = np.asarray(sim._b_h[a])
        out[f"cpml_c_h{a}"] = np.asarray(sim._c_h[a])
    out["src_pos"] = np.array([list(s.position) + [0.0] * (3 - len(s.position))
                               for s in sim.sources], dtype=np.float64).reshape(-1, 3)
    out["src_amp"] = np.array([s.amplitude for s in sim.sources], dtype=np.float64)
    out["src_comp"] = np.array([s.component for s in sim.sources])
    for name, comps in res.fields.items():
        out[f"mt/{name}"] = np.asarray(res.monitor_times[name])
        for c, v in comps.items():
            out[f"field/{name}/{c}"] = np.asarray(v)
            if type(v).__name__ == "CompressedFieldSeries":
                out[f"zbytes/{name}/{c}"] = np.int64(v.nbytes)
    for name, comps in res.dft.items():
        out[f"dftf/{name}"] = np.asarray(res.dft_freqs[name])
        for c, v in comps.items():
            out[f"dft/{name}/{c}"] = np.asarray(v)
    for name, v in res.flux.items():
        out[f"flux/{name}"] = np.float64(v)
    return out


def main(ref_root, only=None):
    sys.dont_write_bytecode = True
    sys.path.insert(0, str(Path(ref_root) / "src"))
    import photonfdtd as ref
    for key, builder in {**SCENES, **FEATURE_SCENES}.items():
        if only and key not in only:
            continue
        precs = ["float64"] + (["float32"] if key in F32_SCENES else [])
        for prec in precs:
            sim = builder(ref, "numpy", precision=prec)
            res = sim.run()
            name = key if prec == "float64" else key + "_f32"
            np.savez_compressed(HERE / f"{name}.npz", **dump(sim, res))
            print("wrote", name, sim.grid.shape, sim.n_steps, flush=True)
    # mode-solver results: n_eff and every field of the solved modes
    for key, builder in MODE_CASES.items():
        if only and key not in only:
            continue
        mr = builder(ref).solve()
        out = {"n_eff": np.asarray(mr.n_eff), "y": mr.y, "z": mr.z, "eps_r": mr.eps_r,
               "psi": mr.psi}
        for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
            out[c] = np.asarray(getattr(mr, c))
        np.savez_compressed(HERE / f"{key}.npz", **out)
        print("wrote", key, mr.n_eff, flush=True)


if __name__ == "__main__":
    # python make_golden.py [reference root] [--only scene,scene]
    args = [a for a in sys.argv[1:] if not a.startswith("--only")]
    only = [a.split("=", 1)[1].split(",") for a in sys.argv[1:] if a.startswith("--only=")]
    main(args[0] if args else os.environ.get("PHOTONFDTD_REF", "/root/reference"),
         only[0] if only else None)
