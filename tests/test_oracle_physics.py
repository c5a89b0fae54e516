"""Analytic checks of the oracle (CPU): propagation speed, CPML absorption,
ADE phase index, Drude stability, running DFT == post-hoc DFT, plane DFT ==
slice of the full DFT, and linearity in the source amplitude. Mirrors the
reference's physics tests (tests/test_physics.py:15-80,
tests/test_dispersion.py:53-97, tests/test_memory.py:30-54,
tests/test_smatrix.py:58-87)."""
import numpy as np
import pytest

import photonfdtd_amd as pf
from oracle.fdtd_numpy import run_oracle

C0 = pf.C_0


def test_1d_pulse_travels_at_c():
    lam0 = 1e-6
    grid = pf.Grid(size=(6e-6,), cell_size=lam0 / 30, pml_layers=(12,))
    src = pf.PointDipole(position=(-2e-6,), component="Ey",
                         waveform=pf.GaussianPulse(freq0=C0 / lam0, fwhm=4e-15))
    mon = pf.FieldMonitor(name="s", components=("Ey",), interval=5)
    sim = pf.Simulation(grid, sources=[src], monitors=[mon], run_time=22e-15)
    r = run_oracle(sim)
    Ey = r["fields"]["s"]["Ey"][..., 0, 0]
    t = r["monitor_times"]["s"]
    i1, i2 = int(len(t) * 0.55), int(len(t) * 0.80)
    corr = np.correlate(Ey[i2], Ey[i1], mode="full")
    lag = np.argmax(corr) - (Ey.shape[1] - 1)
    speed = lag * grid.dx / (t[i2] - t[i1])
    assert abs(speed - C0) / C0 < 0.03


def test_cpml_absorbs_20db_1d():
    lam0 = 1e-6
    grid = pf.Grid(size=(4e-6,), cell_size=lam0 / 25, pml_layers=(12,))
    src = pf.PointDipole(position=(0.0,), component="Ey",
                         waveform=pf.GaussianPulse(freq0=C0 / lam0, fwhm=5e-15))
    mon = pf.FieldMonitor(name="s", components=("Ey",), interval=10)
    sim = pf.Simulation(grid, sources=[src], monitors=[mon], run_time=120e-15)
    Ey = run_oracle(sim)["fields"]["s"]["Ey"][..., 0, 0]
    n = grid.shape[0]
    energy = np.sum(Ey[:, n // 4: 3 * n // 4] ** 2, axis=1)
    assert energy.max() > 0 and energy[-1] / energy.max() < 1e-2


def test_lorentz_phase_index_matches_analytic():
    med = pf.DispersiveMedium.lorentz(2.25, [(1.2, 500e12, 0.0)])
    L = 60e-6
    grid = pf.Grid(size=(L,), cell_size=50e-9, pml_layers=(30,))
    box = pf.Box(center=(0.0,), size=(2 * L,), medium=med)
    src = pf.PointDipole(position=(-20e-6,), component="Ez",
                         waveform=pf.GaussianPulse(freq0=200e12, fwhm=8e-15))
    freqs = np.array([150e12, 180e12, 210e12, 240e12, 270e12])
    mon = pf.DFTMonitor(name="d", components=("Ez",), freqs=freqs)
    sim = pf.Simulation(grid, structures=[box], sources=[src], monitors=[mon], run_time=400e-15)
    D = run_oracle(sim)["dft"]["d"]["Ez"]
    x = grid.coords[0]
    i1, i2 = int(np.argmin(np.abs(x))), int(np.argmin(np.abs(x - 15e-6)))
    for fi, f in enumerate(freqs):
        ph = np.unwrap(np.angle(D[fi, :, 0, 0]))
        n_fdtd = abs(ph[i2] - ph[i1]) / (x[i2] - x[i1]) * C0 / (2 * np.pi * f)
        assert n_fdtd == pytest.approx(np.sqrt(med.eps_model(f)).real, rel=0.025)


def test_drude_run_stays_finite():
    med = pf.DispersiveMedium.drude(1.0, f_plasma_Hz=150e12, f_collision_Hz=5e12)
    grid = pf.Grid(size=(8e-6, 8e-6), cell_size=40e-9, pml_layers=(10, 10, 0))
    box = pf.Box(center=(1e-6, 0.0), size=(2e-6, 2e-6), medium=med)
    src = pf.PointDipole(position=(-2e-6, 0.0), component="Ez",
                         waveform=pf.GaussianPulse(freq0=200e12, fwhm=8e-15))
    mon = pf.FieldMonitor(name="s", components=("Ez",), interval=50)
    sim = pf.Simulation(grid, structures=[box], sources=[src], monitors=[mon], run_time=60e-15)
    f = run_oracle(sim)["fields"]["s"]["Ez"]
    assert np.isfinite(f).all() and np.abs(f[-1]).max() < 5 * np.abs(f).max()


def test_running_dft_equals_post_hoc_dft():
    f0 = C0 / 1e-6
    freqs = [0.8 * f0, f0, 1.2 * f0]
    grid = pf.Grid(size=(3e-6, 3e-6), cell_size=50e-9, pml_layers=(10, 10, 0))
    src = pf.PointDipole(position=(0.0, 0.0), component="Ez",
                         waveform=pf.GaussianPulse(freq0=f0, fwhm=6e-15))
    td = pf.FieldMonitor(name="td", components=("Ez", "Hx"), interval=1)
    fd = pf.DFTMonitor(name="fd", components=("Ez", "Hx"), freqs=freqs)
    sim = pf.Simulation(grid, sources=[src], monitors=[td, fd], run_time=60e-15)
    r = run_oracle(sim)
    steps = np.round(r["monitor_times"]["td"] / sim.dt).astype(int)
    for comp in ("Ez", "Hx"):
        series = r["fields"]["td"][comp]
        got = r["dft"]["fd"][comp]
        tvec = (steps + (1.0 if comp[0] == "E" else 0.5)) * sim.dt
        manual = np.stack([(series * np.exp(1j * 2 * np.pi * f * tvec)[:, None, None, None]).sum(0)
                           for f in freqs]) * sim.dt
        assert np.abs(manual - got).max() / np.abs(got).max() < 1e-10


def test_plane_dft_equals_full_slice():
    f0 = C0 / 1e-6
    grid = pf.Grid(size=(3e-6, 3e-6, 2e-6), cell_size=1e-6 / 16, pml_layers=(6, 6, 6))
    src = pf.PointDipole(position=(-0.5e-6, 0, 0), component="Ez",
                         waveform=pf.GaussianPulse(freq0=f0, fwhm=6e-15))
    full = pf.DFTMonitor(name="full", components=("Ey", "Hz"), freqs=[f0])
    plane = pf.DFTMonitor(name="plane", components=("Ey", "Hz"), freqs=[f0],
                          plane_axis="x", plane_position=0.6e-6)
    sim = pf.Simulation(grid, sources=[src], monitors=[full, plane], run_time=12e-15)
    r = run_oracle(sim)
    xi = int(np.argmin(np.abs(grid.coords[0] - 0.6e-6)))
    for c in ("Ey", "Hz"):
        f = r["dft"]["full"][c][:, xi:xi + 1]
        assert np.array_equal(r["dft"]["plane"][c], f)


def test_linear_in_source_amplitude():
    """Maxwell + ADE are linear: doubling the amplitude doubles every output
    up to round-off (pins dropped or misordered source terms)."""
    def run(amp):
        med = pf.DispersiveMedium.lorentz(1.5, [(0.8, 300e12, 10e12)])
        grid = pf.Grid(size=(1.2e-6, 1.0e-6), cell_size=40e-9, pml_layers=(6, 6, 0))
        src = pf.PointDipole(position=(-0.2e-6, 0.0), component="Ez", amplitude=amp,
                             waveform=pf.GaussianPulse(freq0=250e12, fwhm=5e-15))
        mon = pf.FieldMonitor(name="s", components=("Ez", "Hy"), interval=40)
        box = pf.Box(center=(0.2e-6, 0), size=(0.4e-6, 0.4e-6), medium=med)
        return run_oracle(pf.Simulation(grid, structures=[box], sources=[src], monitors=[mon],
                                        run_time=15e-15))
    a, b = run(1.0), run(2.0)
    for c in ("Ez", "Hy"):
        x, y = a["fields"]["s"][c], b["fields"]["s"][c]
        assert np.abs(y - 2 * x).max() <= 1e-12 * np.abs(y).max()
