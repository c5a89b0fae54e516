"""The torch restatement (used as the exact-gradient checker of the HIP
adjoint) reproduces the NumPy oracle's forward DFT outputs."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def small_3d(pf):
    grid = pf.Grid(size=(12 * 80e-9, 10 * 80e-9, 9 * 80e-9), cell_size=80e-9, pml_layers=(3, 3, 3))
    box = pf.Box(center=(0.1e-6, 0, 0), size=(0.3e-6, 0.3e-6, 0.3e-6), medium=pf.Medium(3.0))
    src = pf.PointDipole(position=(-0.2e-6, 0, 0), component="Ey",
                         waveform=pf.GaussianPulse(freq0=pf.C_0 / 1e-6, fwhm=4e-15))
    dft = pf.DFTMonitor(name="p", components=("Ey", "Ez", "Hy", "Hz"),
                        freqs=[pf.C_0 / 1.1e-6, pf.C_0 / 0.9e-6], plane_axis="x",
                        plane_position=0.25e-6)
    return pf.Simulation(grid, structures=[box], sources=[src], monitors=[dft], run_time=5e-15)


def test_torch_forward_matches_numpy_oracle():
    import photonfdtd_amd as pf
    from oracle.fdtd_numpy import run_oracle
    from oracle.fdtd_torch import run_torch
    sim = small_3d(pf)
    ref = run_oracle(sim)["dft"]["p"]
    got = run_torch(sim, torch.tensor(np.asarray(sim.eps_r, dtype=np.float64)))["p"]
    for c in ref:
        a = got[c].numpy()
        assert np.linalg.norm(a - ref[c]) <= 1e-14 * np.linalg.norm(ref[c])


def test_torch_gradient_matches_central_difference():
    import photonfdtd_amd as pf
    from oracle.fdtd_torch import run_torch
    sim = small_3d(pf)
    eps0 = np.asarray(sim.eps_r, dtype=np.float64)

    def loss(eps):
        d = run_torch(sim, eps)["p"]
        return sum((v.abs() ** 2).sum() for v in d.values())

    eps = torch.tensor(eps0, requires_grad=True)
    loss(eps).backward()
    g = eps.grad.numpy()
    for idx in [(6, 5, 4), (7, 4, 4), (5, 5, 5)]:
        h = 1e-4
        ep, em = eps0.copy(), eps0.copy()
        ep[idx] += h
        em[idx] -= h
        fd = (float(loss(torch.tensor(ep))) - float(loss(torch.tensor(em)))) / (2 * h)
        assert fd == pytest.approx(g[idx], rel=1e-5)
