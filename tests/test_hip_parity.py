"""GPU parity: the HIP stepper against the reference's golden trajectories.

float64 runs use the reference operation order and must agree to
<= 1e-12 relative L2 (the fixtures were written by the reference NumPy path,
and the oracle reproduces them bit for bit). float32 runs in reference order
reproduce the reference's own float32 trajectories; float32 fused runs must
stay within 1e-4 relative L2 of the float64 golden.
"""
import os

import numpy as np
import pytest

import golden_io as G
from scenes import F32_SCENES, SCENES

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-4


def _run(key, precision, exact):
    import photonfdtd_amd as pf
    print(f"\n[case] {key} precision={precision} exact={exact}", flush=True)
    old = os.environ.get("PHOTONFDTD_HIP_EXACT")
    os.environ["PHOTONFDTD_HIP_EXACT"] = "1" if exact else "0"
    try:
        return SCENES[key](pf, "hip", precision=precision).run()
    finally:
        if old is None:
            os.environ.pop("PHOTONFDTD_HIP_EXACT", None)
        else:
            os.environ["PHOTONFDTD_HIP_EXACT"] = old


@pytest.mark.parametrize("key", list(SCENES))
def test_fp64_matches_reference(key):
    rep = {}
    worst = G.compare(_run(key, "float64", True), G.load(key), FP64_TOL, rep)
    print(f"[fp64] {key} worst rel L2 = {worst:.3e}", flush=True)


@pytest.mark.parametrize("key", list(F32_SCENES))
def test_fp32_reference_order_matches_reference_fp32(key):
    rep = {}
    worst = G.compare(_run(key, "float32", True), G.load(key + "_f32"), 1e-6, rep)
    print(f"[fp32 exact] {key} worst rel L2 vs reference float32 = {worst:.3e}", flush=True)


@pytest.mark.parametrize("key", list(SCENES))
def test_fp32_fast_within_1e4_of_fp64(key):
    rep = {}
    worst = G.compare(_run(key, "float32", False), G.load(key), FP32_TOL, rep)
    print(f"[fp32 fast] {key} worst rel L2 vs float64 = {worst:.3e}", flush=True)


def test_quickstart_2d_dipole_fp64_vs_oracle():
    """The reference README quick start (80x80 + 12-cell CPML, 200 fs)."""
    import photonfdtd_amd as pf
    from oracle.fdtd_numpy import run_oracle

    def sim():
        grid = pf.Grid(size=(4e-6, 4e-6), cell_size=50e-9, pml_layers=(12, 12, 0))
        src = pf.PointDipole(position=(0.0, 0.0), component="Ez",
                             waveform=pf.GaussianPulse(freq0=pf.C_0 / 1e-6, fwhm=10e-15))
        mon = pf.FieldMonitor(name="snap", components=("Ez",), interval=20)
        return pf.Simulation(grid, sources=[src], monitors=[mon], run_time=200e-15,
                             backend="hip")

    print("\n[case] quickstart fp64", flush=True)
    got = sim().run()
    ref = run_oracle(sim())
    a, b = got.fields["snap"]["Ez"], ref["fields"]["snap"]["Ez"]
    assert a.shape == b.shape == (105, 80, 80, 1)
    e = G.rel_l2(a, b)
    print(f"[fp64] quickstart rel L2 = {e:.3e}", flush=True)
    assert e <= FP64_TOL
    assert np.array_equal(got.monitor_times["snap"], ref["monitor_times"]["snap"])


def test_bad_inputs_raise():
    import photonfdtd_amd as pf
    grid = pf.Grid(size=(1e-6, 1e-6), cell_size=50e-9, pml_layers=(4, 4, 0))
    src = pf.PointDipole(position=(0, 0), component="Ez",
                         waveform=pf.GaussianPulse(freq0=3e14, fwhm=5e-15))
    sim = pf.Simulation(grid, sources=[src], run_time=2e-15, backend="hip")
    from photonfdtd_amd.hip_backend import HipEngine
    eng = HipEngine(sim, np.float64, True)
    with pytest.raises(ValueError):
        eng.advance(0, sim.n_steps + 5)          # beyond the waveform table
    with pytest.raises(ValueError):
        eng.read_box("Ez", [0, 0, 0], [30, 1, 1], [1, 1, 1])   # box outside grid
    eng.advance(0, 3)
    assert np.isfinite(eng.get("Ez")).all()
    eng.close()
