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


def _awkward_3d(pf, precision):
    """3-D grid whose extents are not multiples of the tile sizes, with
    CPML on every face, a dispersive box and an x-plane DFT."""
    grid = pf.Grid(size=(70 * 40e-9, 37 * 40e-9, 130 * 40e-9), cell_size=40e-9,
                   pml_layers=(6, 5, 7))
    drude = pf.DispersiveMedium.drude(1.0, f_plasma_Hz=80e12, f_collision_Hz=5e12)
    st = [pf.Box(center=(0, 0, 0), size=(1e-6, 0.6e-6, 2e-6), medium=pf.Medium(4.0)),
          pf.Box(center=(0.3e-6, 0.1e-6, -0.5e-6), size=(0.4e-6, 0.3e-6, 0.6e-6), medium=drude)]
    src = pf.PointDipole(position=(-0.4e-6, 0.05e-6, 0.3e-6), component="Ez",
                         waveform=pf.GaussianPulse(freq0=pf.C_0 / 1.2e-6, fwhm=4e-15))
    dft = pf.DFTMonitor(name="d", components=("Ex", "Ez", "Hy"), freqs=[pf.C_0 / 1.2e-6],
                        plane_axis="x", plane_position=0.5e-6)
    mon = pf.FieldMonitor(name="s", components=("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"),
                          times=[6e-15])
    return pf.Simulation(grid, structures=st, sources=[src], monitors=[dft, mon],
                         run_time=7e-15, precision=precision, backend="hip", courant=0.9)


@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("xchunk", ["64", "7"])
def test_fused_step_bitwise_equals_generic(precision, xchunk, monkeypatch):
    """The fused ping-pong kernel (tiles + recomputed H aprons + x-chunk
    prologues) must reproduce the generic in-place kernels bit for bit in
    reference operation order."""
    import photonfdtd_amd as pf
    print(f"\n[case] fused vs generic {precision} xchunk={xchunk}", flush=True)
    monkeypatch.setenv("PHOTONFDTD_HIP_EXACT", "1")
    monkeypatch.setenv("PHOTONFDTD_HIP_XCHUNK", xchunk)
    monkeypatch.setenv("PHOTONFDTD_HIP_FUSED", "1")
    a = _awkward_3d(pf, precision).run()
    monkeypatch.setenv("PHOTONFDTD_HIP_FUSED", "0")
    b = _awkward_3d(pf, precision).run()
    for c in a.fields["s"]:
        assert np.array_equal(a.fields["s"][c], b.fields["s"][c]), c
    for c in a.dft["d"]:
        assert np.array_equal(a.dft["d"][c], b.dft["d"][c]), c
    assert np.abs(a.fields["s"]["Ez"]).max() > 0


def test_fused_fp64_vs_oracle_awkward():
    import photonfdtd_amd as pf
    from oracle.fdtd_numpy import run_oracle
    print("\n[case] awkward 3-D fp64 vs oracle", flush=True)
    got = _awkward_3d(pf, "float64").run()
    ref = run_oracle(_awkward_3d(pf, "float64"))
    worst = 0.0
    for c in got.fields["s"]:
        worst = max(worst, G.rel_l2(got.fields["s"][c], ref["fields"]["s"][c]))
    for c in got.dft["d"]:
        worst = max(worst, G.rel_l2(got.dft["d"][c], ref["dft"]["d"][c]))
    print(f"[fp64] awkward 3-D worst rel L2 = {worst:.3e}", flush=True)
    assert worst <= FP64_TOL
