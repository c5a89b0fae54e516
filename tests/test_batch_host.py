"""Host-side checks of the batched stepping front end (no GPU needed)."""
import numpy as np
import pytest


def _sim(pf, comp="Ex", size=(1.6e-6, 1.6e-6), monitors=()):
    grid = pf.Grid(size=size, cell_size=40e-9, pml_layers=(6, 6, 0) if len(size) == 2 else 6)
    src = pf.PointDipole(position=(0.0, 0.0), component=comp,
                         waveform=pf.GaussianPulse(freq0=pf.C_0 / 1.55e-6, fwhm=10e-15))
    return pf.Simulation(grid, sources=[src], monitors=list(monitors), run_time=5e-15,
                         backend="hip")


def test_polarisation_from_sources():
    import photonfdtd_amd as pf
    from photonfdtd_amd.batch import polarisation
    assert polarisation(_sim(pf, "Ex")) == 0
    assert polarisation(_sim(pf, "Hz")) == 0
    assert polarisation(_sim(pf, "Ez")) == 1
    assert polarisation(_sim(pf, "Hy")) == 1
    s = _sim(pf, "Ex")
    s.sources.append(pf.PointDipole(position=(0.0, 0.0), component="Ez",
                                    waveform=s.sources[0].waveform))
    with pytest.raises(ValueError, match="mix"):
        polarisation(s)


def test_rejects_unsupported_scenes():
    import photonfdtd_amd as pf
    from photonfdtd_amd.batch import SimBatch
    cube = (0.8e-6, 0.8e-6, 0.8e-6)
    with pytest.raises(ValueError, match="2-D grids only or 3-D grids only"):
        SimBatch([_sim(pf), _sim(pf, size=cube)])
    with pytest.raises(ValueError, match="E-side sources only"):
        SimBatch([_sim(pf, "Hz", size=cube)])
    with pytest.raises(ValueError, match="DFTMonitor only"):
        SimBatch([_sim(pf, monitors=[pf.FieldMonitor(name="f", components=("Ex",))])])
    with pytest.raises(ValueError, match="empty"):
        SimBatch([])
    a = _sim(pf)
    b = _sim(pf)
    b2 = pf.Simulation(b.grid, sources=b.sources, run_time=5e-15, precision="float32",
                       backend="hip")
    with pytest.raises(ValueError, match="precision"):
        SimBatch([a, b2])


def test_dft_flux_closed_form():
    from photonfdtd_amd.batch import dft_flux
    rng = np.random.default_rng(0)
    ey = rng.normal(size=(3, 5)) + 1j * rng.normal(size=(3, 5))
    hz = rng.normal(size=(3, 5)) + 1j * rng.normal(size=(3, 5))
    got = dft_flux({"Ey": ey, "Hz": hz}, "x", 2.0)
    want = np.array([2.0 * sum((ey[f, j] * np.conj(hz[f, j])).real for j in range(5))
                     for f in range(3)])
    assert np.allclose(got, want, rtol=1e-14)
    # TM pair enters with the opposite sign: S_x = Ey Hz - Ez Hy
    got_tm = dft_flux({"Ez": ey, "Hy": hz}, "x", 2.0)
    assert np.allclose(got_tm, -want, rtol=1e-14)
    with pytest.raises(ValueError):
        dft_flux({"Ex": ey}, "x", 1.0)


def test_batched_scene_generator_is_seeded():
    import photonfdtd_amd as pf
    from bench import batched_2d_scenes
    a = batched_2d_scenes(pf, n_sims=3, lo=256, hi=1024, steps=10, seed=0, backend="hip")
    b = batched_2d_scenes(pf, n_sims=3, lo=256, hi=1024, steps=10, seed=0, backend="hip")
    assert [s.grid.shape for s in a] == [s.grid.shape for s in b]
    for s in a:
        nx, ny, nz = s.grid.shape
        assert nx == ny and 256 <= nx <= 1024 and nz == 1
        assert 1 <= len(s.structures) <= 4
        assert np.array_equal(s.eps_r, b[a.index(s)].eps_r)
