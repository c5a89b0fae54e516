"""Scene API checks against closed forms and published values (CPU only)."""
import math

import numpy as np
import pytest

import photonfdtd_amd as pf
from photonfdtd_amd import materials as m

C0 = 299_792_458.0


def test_grid_shape_coords_and_collapse():
    g = pf.Grid(size=(2e-6, 2e-6), cell_size=50e-9, pml_layers=(8, 8, 0))
    assert g.shape == (40, 40, 1) and g.ndim == 2
    assert np.allclose(g.coords[0], (np.arange(40) - 19.5) * 50e-9, rtol=0, atol=0)
    assert g.index_at((0.0, 0.0)) == (19, 19, 0)          # argmin tie -> first node
    with pytest.raises(ValueError):
        pf.Grid(size=(1e-7,), cell_size=50e-9)             # fewer than 4 cells


def test_courant_dt_closed_form():
    g = pf.Grid(size=(4e-6, 4e-6), cell_size=(50e-9, 50e-9), pml_layers=(12, 12, 0))
    sim = pf.Simulation(g, run_time=200e-15)
    assert sim.dt == 0.99 / (C0 * math.sqrt(2.0 / 50e-9 ** 2))
    assert sim.n_steps == math.ceil(200e-15 / sim.dt) == 1714
    # a scalar cell size keeps the collapsed axis' spacing (3-D Courant bound)
    g3 = pf.Grid(size=(4e-6, 4e-6), cell_size=50e-9, pml_layers=(12, 12, 0))
    assert pf.Simulation(g3, run_time=200e-15).n_steps == 2099


def test_cpml_coefficients_closed_form():
    """b = exp(-(sigma+alpha) dt / eps0), c = sigma/(sigma+alpha) (b - 1) with
    cubic sigma and linear alpha, checked node by node."""
    n, d, npml, dt = 30, 40e-9, 6, 5e-17
    be, ce, bh, ch = pf.cpml_coeffs_1axis(n, d, npml, dt)
    smax = 0.8 * 4 * math.log(1e7) / (2 * pf.ETA_0 * npml * d)
    for i in range(n):
        for arr_b, arr_c, x in ((be, ce, float(i)), (bh, ch, i + 0.5)):
            if i < npml:
                u = min(max((npml - 1 - x) / (npml - 1), 0.0), 1.0)
            elif i >= n - npml:
                u = min(max((x - (n - npml)) / (npml - 1), 0.0), 1.0)
            else:
                assert arr_b[i] == 0.0 and arr_c[i] == 0.0
                continue
            sig, alp = smax * u ** 3, 0.24 * (1 - u)
            b = math.exp(-(sig + alp) * dt / pf.EPS_0)
            assert arr_b[i] == pytest.approx(b, rel=1e-14)
            assert arr_c[i] == pytest.approx(sig / (sig + alp) * (b - 1), rel=1e-12, abs=1e-300)
    # c != 0 on 11 E nodes and 11/12 H nodes per slab (lower/upper)
    assert int((ce != 0).sum()) == 2 * (npml - 1)
    assert int((ch != 0).sum()) == 2 * npml - 1


def test_library_indices_match_published():
    n = lambda med, lam_um: med.index(C0 / (lam_um * 1e-6)).real
    assert n(m.silica(), 1.55) == pytest.approx(1.44402, abs=2e-4)
    assert n(m.silicon(), 1.55) == pytest.approx(3.4777, abs=2e-3)
    assert n(m.silicon_nitride(), 1.55) == pytest.approx(1.9963, abs=2e-3)
    assert n(m.lithium_niobate("o"), 1.55) == pytest.approx(2.2111, abs=2e-3)
    assert n(m.lithium_niobate("e"), 1.55) == pytest.approx(2.1376, abs=2e-3)
    eps_au = m.gold().eps_model(C0 / 633e-9)
    assert eps_au.real == pytest.approx(-9.81, abs=0.1)
    assert eps_au.imag == pytest.approx(1.96, abs=0.1)
    eps_ag = m.silver().eps_model(C0 / 633e-9)
    assert eps_ag.real == pytest.approx(-14.49, abs=0.2)
    assert eps_ag.imag == pytest.approx(1.10, abs=0.1)


def test_sellmeier_is_lossless_lorentz():
    terms = [(0.6961663, 0.0684043 ** 2), (0.4079426, 0.1162414 ** 2)]
    med = pf.DispersiveMedium.sellmeier(terms)
    for lam in (0.5, 1.0, 1.55, 3.0):
        n2 = 1.0 + sum(B * lam ** 2 / (lam ** 2 - C) for B, C in terms)
        assert med.eps_model(C0 / (lam * 1e-6)).real == pytest.approx(n2, rel=1e-9)


def test_ade_coefficients_closed_form_and_guard():
    med = pf.DispersiveMedium.lorentz(2.0, [(1.5, 300e12, 20e12)])
    grid = pf.Grid(size=(2e-6,), cell_size=50e-9, pml_layers=(6,))
    sim = pf.Simulation(grid, structures=[pf.Box((0,), (1e-6,), med)], run_time=1e-15)
    (ade,) = sim.ade_media()
    p = med.poles[0]
    dt = sim.dt
    den = 1 + p.gamma * dt
    assert ade.abc[0, 0] == (2 - p.omega0 ** 2 * dt ** 2) / den
    assert ade.abc[0, 1] == (p.gamma * dt - 1) / den
    assert ade.abc[0, 2] == pf.EPS_0 * p.strength * dt ** 2 / den
    assert ade.inv_eps == 1 / (pf.EPS_0 * 2.0)
    assert ade.cells.size == 20          # nodes at +-25 nm .. +-475 nm
    bad = pf.Simulation(pf.Grid(size=(4e-6,), cell_size=1e-7, pml_layers=(10,)),
                        structures=[pf.Box((0.0,), (2e-6,), m.silica())], run_time=5e-15)
    with pytest.raises(ValueError, match="unstable"):
        bad.ade_media()


def test_gaussian_pulse_closed_form():
    w = pf.GaussianPulse(freq0=2e14, fwhm=8e-15, delay=1e-15)
    sigma = 8e-15 / (2 * math.sqrt(2 * math.log(2)))
    t = np.linspace(0, 60e-15, 7)
    expect = np.exp(-((t - (4 * sigma + 1e-15)) / sigma) ** 2) * np.sin(
        2 * math.pi * 2e14 * (t - (4 * sigma + 1e-15)))
    assert np.allclose(w(t), expect, rtol=1e-13, atol=1e-300)


def test_box_and_polyslab_masks():
    g = pf.Grid(size=(1e-6, 1e-6), cell_size=100e-9)
    box = pf.Box(center=(0, 0), size=(0.4e-6, 0.2e-6), medium=pf.Medium(2.0))
    mask = box.region_mask(g)
    xs = np.flatnonzero(mask[:, 5, 0])
    assert g.coords[0][xs].min() >= -0.2e-6 and g.coords[0][xs].max() <= 0.2e-6
    sq = pf.PolySlab(vertices=((-0.26e-6, -0.26e-6), (0.26e-6, -0.26e-6),
                               (0.26e-6, 0.26e-6), (-0.26e-6, 0.26e-6)),
                     z_bounds=(-1, 1), medium=pf.Medium(3.0))
    pm = sq.region_mask(g)[:, :, 0]
    inside = (np.abs(g.coords[0])[:, None] < 0.26e-6) & (np.abs(g.coords[1])[None, :] < 0.26e-6)
    assert np.array_equal(pm, inside)


def test_mode_source_expansion_weights():
    g = pf.Grid(size=(1e-6, 1e-6, 1e-6), cell_size=100e-9)
    yy = np.linspace(-0.3e-6, 0.3e-6, 7)
    prof = np.outer(np.cos(yy / 0.4e-6), np.ones(3))
    src = pf.ModeSource(center=(0, 0, 0), size=(0, 0.6e-6, 0.2e-6), component="Ey",
                        waveform=pf.GaussianPulse(2e14, 5e-15), profile=prof,
                        profile_coords=(yy, np.array([-0.1e-6, 0.0, 0.1e-6])), amplitude=2.0)
    dips = src.expand(g)
    # plane nodes: y within +-0.3 um (6 nodes at +-50, 150, 250 nm), z within +-0.1 um (2)
    assert len(dips) == 12
    for d in dips:
        y = d.position[1]
        assert d.amplitude == pytest.approx(2.0 * np.interp(y, yy, prof[:, 0]), rel=1e-12)


def test_precision_resolution():
    from photonfdtd_amd.simulation import resolve_precision
    d = resolve_precision({"compute": "float32", "monitors": "float64"})
    assert d["Ex"] == np.float32 and d["psi"] == np.float32 and d["monitors"] == np.float64
    d = resolve_precision({"default": "float32", "Ez": "float64"})
    assert d["Ez"] == np.float64 and d["Ex"] == np.float32
    with pytest.raises(ValueError):
        resolve_precision({"bogus": "float32"})


def test_backend_gate():
    g = pf.Grid(size=(1e-6,), cell_size=50e-9)
    with pytest.raises(ValueError):
        pf.Simulation(g, backend="numpy")
    sim = pf.Simulation(g, run_time=1e-15, precision={"default": "float32", "Ez": "float64"})
    with pytest.raises(NotImplementedError):
        sim.run()
    # compression codecs resolve before any device work (storage.py:80-101)
    from photonfdtd_amd import storage
    if storage._zstd is None:
        sim = pf.Simulation(g, run_time=1e-15,
                            monitors=[pf.FieldMonitor("f", compression="zstd")])
        with pytest.raises(ImportError):
            sim.run()
