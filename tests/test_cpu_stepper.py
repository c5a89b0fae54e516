"""The OpenMP restatement of the reference's compiled CPU stepper
(``oracle/cpu_fdtd.cpp``; rust/src/kernel.rs, blocked.rs), the measured CPU
baseline. Checked like the reference checks its Rust backend against NumPy
(``tests/test_rust.py:37-83``, 1e-12 of the maximum): float64 runs against
the reference's own trajectories in the golden fixtures (ADE media, DFT
planes and sources included); temporal blocking bit for bit against plain
stepping, as ``blocked.rs`` claims for itself; float32 within 1e-4."""
import numpy as np
import pytest

import golden_io as G
from scenes import SCENES


@pytest.fixture(scope="module", autouse=True)
def built():
    from oracle import build_cpu
    build_cpu.build()


def _run(sim, npdt, **kw):
    from oracle.fdtd_cpu import CpuStepper
    eng = CpuStepper(sim, npdt, **kw)
    eng.advance(0, sim.n_steps)
    fields = {c: eng.get(c) for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
    dft = {m.name: eng.dft(m.name) for m in sim.monitors if type(m).__name__ == "DFTMonitor"}
    eng.close()
    return fields, dft


@pytest.mark.parametrize("key", ["box3d", "metals_xz", "lorentz1d", "tm2d_box", "te2d_poly"])
def test_float64_matches_reference_dft(key):
    """Scenes without flux or subpixel; DFT planes against the fixture."""
    import photonfdtd_amd as pf
    sim = SCENES[key](pf, "hip")
    gold = G.load(key)
    _, dft = _run(sim, np.float64, threads=4)
    assert dft, key
    for name, comps in dft.items():
        for c, v in comps.items():
            ref = gold[f"dft/{name}/{c}"]
            assert np.linalg.norm(v - ref) <= 1e-12 * np.linalg.norm(ref), (key, name, c)


def test_final_fields_match_oracle():
    import photonfdtd_amd as pf
    from oracle.fdtd_numpy import run_oracle
    sim = SCENES["metals_xz"](pf, "hip")
    _, state = run_oracle(SCENES["metals_xz"](pf, "hip"), return_state=True)
    fields, _ = _run(sim, np.float64, threads=3)
    for c in fields:
        ref = state[c]
        assert np.abs(fields[c] - ref).max() <= 1e-12 * max(np.abs(ref).max(), 1e-300), c


def _dielectric3d(pf):
    grid = pf.Grid(size=(64 * 40e-9, 20 * 40e-9, 22 * 40e-9), cell_size=40e-9, pml_layers=(6, 4, 5))
    box = pf.Box(center=(0.2e-6, 0, 0), size=(0.6e-6, 0.3e-6, 0.4e-6), medium=pf.Medium(6.0))
    w = pf.GaussianPulse(freq0=pf.C_0 / 1e-6, fwhm=4e-15)
    src = [pf.PointDipole(position=(-0.3e-6, 0.05e-6, 0.0), component="Ez", waveform=w),
           pf.PointDipole(position=(0.9e-6, -0.1e-6, 0.1e-6), component="Hy", waveform=w,
                          amplitude=1e-3)]
    return pf.Simulation(grid, structures=[box], sources=src, run_time=12e-15, backend="hip")


@pytest.mark.parametrize("t_block,rows", [(4, 8), (8, 10), (3, 16)])
def test_temporal_blocking_bitwise(t_block, rows):
    import photonfdtd_amd as pf
    sim = _dielectric3d(pf)
    plain, _ = _run(sim, np.float64, threads=4)
    blocked, _ = _run(sim, np.float64, threads=4, t_block=t_block, tile_rows=rows)
    for c in plain:
        assert np.array_equal(plain[c], blocked[c]), c
    assert np.abs(plain["Ez"]).max() > 0


def test_float32_within_1e4():
    import photonfdtd_amd as pf
    sim64 = SCENES["box3d"](pf, "hip")
    sim32 = SCENES["box3d"](pf, "hip", precision="float32")
    f64, d64 = _run(sim64, np.float64, threads=2)
    f32, d32 = _run(sim32, np.float32, threads=2)
    for name in d64:
        for c in d64[name]:
            a, b = d32[name][c], d64[name][c]
            assert np.linalg.norm(a - b) <= 1e-4 * np.linalg.norm(b)
