"""GPU: waveguide modes through the HIP stepper.

* ``unimode3d`` (a solved mode launched one way by electric and magnetic
  current sheets, so the generic kernels with H-side sources run) against the
  reference's float64 trajectory: within 1e-9 (the mode fields themselves
  agree with the reference's solver to ~1e-12).
* The reference's S-parameter acceptance tests (``tests/test_smatrix.py:91-160``,
  run there on its JAX backend), here on ``backend="hip"``: a mode launched
  down a straight lossless guide gives |S21| = 1 within 5 % between two
  downstream ports, and ``UniModeSource`` radiates backward at less than 0.3
  of its forward amplitude.
"""
import numpy as np
import pytest

import golden_io as G
import scenes

pytestmark = pytest.mark.gpu


def test_unimode3d_matches_reference():
    import photonfdtd_amd as pf
    print("\n[case] unimode3d float64 HIP vs reference", flush=True)
    res = scenes.FEATURE_SCENES["unimode3d"](pf, "hip").run()
    worst = G.compare(res, G.load("unimode3d"), 1e-9, {})
    print(f"[unimode3d] worst rel L2 = {worst:.3e}", flush=True)


def _wg_mode(pf, d, Ly, Lz):
    core = pf.Box(center=(0.0, 0.0), size=(0.45e-6, 0.22e-6), medium=pf.Medium.from_index(3.48))
    ms = pf.ModeSolver(size=(Ly, Lz), cell_size=d, structures=[core], wavelength=1.55e-6,
                       background_eps=1.0, num_modes=1)
    return ms, ms.solve()


def test_waveguide_s21_energy_conserving():
    import photonfdtd_amd as pf
    print("\n[case] straight-guide S21 (ModeSource)", flush=True)
    f0 = pf.C_0 / 1.55e-6
    d, Ly, Lz = 40e-9, 2.0e-6, 1.4e-6
    ms, mr = _wg_mode(pf, d, Ly, Lz)
    core = pf.Box(center=(0, 0, 0), size=(10e-6, 0.45e-6, 0.22e-6), medium=pf.Medium.from_index(3.48))
    grid = pf.Grid(size=(6e-6, Ly, Lz), cell_size=d, pml_layers=(10, 10, 10))
    src = pf.ModeSource(center=(-2e-6, 0, 0), size=(0, Ly, Lz), component="Ey",
                        waveform=pf.GaussianPulse(freq0=f0, fwhm=10e-15),
                        profile=np.real(mr.Ey[0]), profile_coords=(ms.y, ms.z))
    ports = [pf.DFTMonitor(name=nm, components=("Ey", "Ez", "Hy", "Hz"), freqs=[f0],
                           plane_axis="x", plane_position=x)
             for nm, x in (("p1", 0.0), ("p2", 2.0e-6))]
    res = pf.Simulation(grid, structures=[core], sources=[src], monitors=ports,
                        run_time=200e-15, precision="float32", backend="hip").run()
    a1, _ = pf.s_parameters(res, "p1", mr, 0, d * d)
    a2, _ = pf.s_parameters(res, "p2", mr, 0, d * d)
    print(f"[s21] |a2/a1| = {abs(a2[0] / a1[0]):.4f}", flush=True)
    assert abs(a2[0] / a1[0]) == pytest.approx(1.0, abs=0.05)


def test_unidirectional_source_extinction():
    import photonfdtd_amd as pf
    print("\n[case] UniModeSource extinction", flush=True)
    f0 = pf.C_0 / 1.55e-6
    d, Ly, Lz = 40e-9, 2.0e-6, 1.4e-6
    _, mr = _wg_mode(pf, d, Ly, Lz)
    core = pf.Box(center=(0, 0, 0), size=(12e-6, 0.45e-6, 0.22e-6), medium=pf.Medium.from_index(3.48))
    grid = pf.Grid(size=(8e-6, Ly, Lz), cell_size=d, pml_layers=(10, 10, 10))
    src = pf.UniModeSource(center=(0, 0, 0), size=(0, Ly, Lz),
                           waveform=pf.GaussianPulse(freq0=f0, fwhm=10e-15), mode=mr,
                           direction="+x")
    ports = [pf.DFTMonitor(name=nm, components=("Ey", "Ez", "Hy", "Hz"), freqs=[f0],
                           plane_axis="x", plane_position=x)
             for nm, x in (("up", -1.5e-6), ("dn", 1.5e-6))]
    res = pf.Simulation(grid, structures=[core], sources=[src], monitors=ports,
                        run_time=220e-15, precision="float32", backend="hip").run()
    fwd, _ = pf.s_parameters(res, "dn", mr, 0, d * d)
    _, back = pf.s_parameters(res, "up", mr, 0, d * d)
    print(f"[extinction] |back/fwd| = {abs(back[0]) / abs(fwd[0]):.4f}", flush=True)
    assert abs(back[0]) / abs(fwd[0]) < 0.3
