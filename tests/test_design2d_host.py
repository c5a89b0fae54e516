"""Host-side checks of the surrogate design problem (no GPU needed)."""
import numpy as np


def test_random_design_is_seeded_binary_and_balanced():
    from photonfdtd_amd.design2d import random_design
    a, b = random_design(7), random_design(7)
    assert np.array_equal(a, b) and a.shape == (128, 128)
    assert set(np.unique(a)) <= {0.0, 1.0}
    fills = [random_design(s).mean() for s in range(20)]
    assert 0.35 < np.mean(fills) < 0.65
    assert not np.array_equal(random_design(1), random_design(2))


def test_scene_and_design_embedding():
    import photonfdtd_amd as pf
    from photonfdtd_amd.design2d import (EPS_HI, EPS_LO, N_DESIGN, design_window, eps_with,
                                         flux_window, scene)
    sim = scene(pf, steps=50)
    assert sim.grid.shape == (512, 512, 1) and sim.n_steps == 50
    i0, j0 = design_window(sim)
    assert (i0, j0) == (192, 192)
    rho = np.zeros((N_DESIGN, N_DESIGN))
    rho[5, 7] = 1.0
    eps = eps_with(sim, rho)
    assert eps[i0 + 5, j0 + 7, 0] == EPS_HI and eps[i0, j0, 0] == EPS_LO
    outside = np.ones(sim.grid.shape, bool)
    outside[i0:i0 + N_DESIGN, j0:j0 + N_DESIGN] = False
    assert np.array_equal(eps[outside], np.asarray(sim.eps_r, dtype=np.float64)[outside])
    # input guide on the axis, output guide offset by 0.8 um
    y = np.asarray(sim.grid.coords[1])
    col = np.asarray(sim.eps_r)[50, :, 0]
    assert col[np.abs(y) < 0.1e-6].min() == EPS_HI and col[np.abs(y - 0.8e-6) < 0.1e-6].max() == EPS_LO
    col = np.asarray(sim.eps_r)[460, :, 0]
    assert col[np.abs(y - 0.8e-6) < 0.1e-6].min() == EPS_HI
    assert flux_window(sim).sum() == 80


def test_plane_flux_closed_form():
    from photonfdtd_amd.design2d import plane_flux
    rng = np.random.default_rng(0)
    ez = rng.normal(size=(3, 1, 10, 1)) + 1j * rng.normal(size=(3, 1, 10, 1))
    hy = rng.normal(size=(3, 1, 10, 1)) + 1j * rng.normal(size=(3, 1, 10, 1))
    mask = np.zeros(10, bool)
    mask[2:5] = True
    want = [-sum((ez[f, 0, j, 0] * np.conj(hy[f, 0, j, 0])).real for j in range(2, 5))
            for f in range(3)]
    assert np.allclose(plane_flux(ez, hy, mask), want, rtol=1e-14)


def test_network_size_and_resolution():
    from photonfdtd_amd.surrogate import ENCODER, layer_table, n_params
    assert 1.8e6 < n_params() < 2.2e6
    h = 128
    for _, _, _, s, _ in ENCODER:
        h = (h - 1) // s + 1
    assert h == 4
    assert layer_table()[-1][1] == 16
