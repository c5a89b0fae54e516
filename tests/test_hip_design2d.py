"""GPU: the surrogate design problem's FDTD labels and adjoint gradient.

Labels from the batched stepper must equal single runs of the same design bit
for bit; the adjoint gradient of the mean transmission w.r.t. the design
density must match central differences of the HIP forward (float64).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_labels_equal_single_runs():
    import photonfdtd_amd as pf
    from photonfdtd_amd.design2d import (Labeler, eps_with, flux_window, plane_flux,
                                         random_design, scene)
    steps = 400
    ref = np.ones(16)
    print("\n[case] design labels batch vs single", flush=True)
    lab = Labeler(pf, ref, batch=3, steps=steps)
    try:
        designs = np.stack([random_design(s) for s in (1, 2, 3)])
        got = lab(designs)
        single = lab(designs[1:2])          # the one-simulation batch
    finally:
        lab.close()
    for k, rho in enumerate(designs):
        sim = scene(pf, steps)
        sim.eps_r = eps_with(sim, rho).astype(sim.dtypes["eps_r"])
        r = sim.run()
        want = plane_flux(r.dft["out"]["Ez"], r.dft["out"]["Hy"], flux_window(sim))
        assert np.array_equal(got[k], want), k
    assert np.array_equal(single[0], got[1])


def test_design_gradient_matches_central_differences():
    import torch
    import photonfdtd_amd as pf
    from photonfdtd_amd.adjoint import value_and_grad_eps
    from photonfdtd_amd.design2d import (EPS_HI, EPS_LO, N_DESIGN, design_window, eps_with,
                                         flux_window, plane_flux, random_design, scene)
    print("\n[case] design gradient vs central differences (float64)", flush=True)
    steps = 3000            # the pulse crosses the design region and reaches the output plane
    sim0 = scene(pf, steps, precision="float64")
    mask = flux_window(sim0)
    i0, j0 = design_window(sim0)
    rho = 0.5 * random_design(4) + 0.25

    def run(r):
        s = scene(pf, steps, precision="float64")
        s.eps_r = eps_with(s, r)
        return s

    def loss(out):
        ez, hy = out["dft"]["out"]["Ez"], out["dft"]["out"]["Hy"]
        m = torch.tensor(mask)
        e = ez.reshape(ez.shape[0], -1)[:, m]
        h = hy.reshape(hy.shape[0], -1)[:, m]
        return -(e * torch.conj(h)).real.sum(dim=1).mean()

    _, g = value_and_grad_eps(run(rho), loss)
    grad = np.asarray(g)[i0:i0 + N_DESIGN, j0:j0 + N_DESIGN, 0] * (EPS_HI - EPS_LO)

    def value(r):
        res = run(r).run()
        return float(plane_flux(res.dft["out"]["Ez"], res.dft["out"]["Hy"], mask).mean())
    # design pixels on the input guide, in the middle, and at the output guide
    for idx in [(10, 64), (64, 80), (120, 104)]:
        h = 1e-3
        rp, rm = rho.copy(), rho.copy()
        rp[idx] += h
        rm[idx] -= h
        fd = (value(rp) - value(rm)) / (2 * h)
        print(f"[fd] {idx}: adjoint {grad[idx]:.6e} fd {fd:.6e}", flush=True)
        assert fd == pytest.approx(grad[idx], rel=1e-3, abs=1e-5 * np.abs(grad).max())


def test_trained_network_improves_on_its_best_designs():
    """What the trained surrogate gains over the method without it. From the
    16 best of 2048 FDTD-labelled designs, 30 steps of gradient ascent on the
    trained network's predicted transmission give designs whose FDTD-verified
    transmission beats their starting designs for most starts and beats the
    same ascent on an untrained network (same architecture and target
    scaling): the network carries design knowledge that the verification
    budget alone (re-checking the best known designs) does not."""
    import torch
    import photonfdtd_amd as pf
    from photonfdtd_amd import design2d as D
    from photonfdtd_amd.surrogate import SpectralSurrogate
    steps = 4000
    print("\n[case] surrogate gain over its training designs", flush=True)
    ref = D.reference_flux(pf, steps)
    lab = D.Labeler(pf, ref, batch=256, steps=steps)
    try:
        X = D.training_designs(2048)
        Y = lab(X)
        J = D.objective(Y)
        net = SpectralSurrogate(seed=0)
        D.train(net, X, Y, epochs=15, batch=64, lr=1e-3)
        torch.cuda.synchronize()
        untrained = SpectralSurrogate(seed=1)
        untrained.y_mu, untrained.y_sd = net.y_mu, net.y_sd
        top = np.argsort(J)[::-1][:16]
        got = {}
        for name, m in (("trained", net), ("untrained", untrained)):
            cand = D.propose_from(m, X[top], iters=30, step=0.02, snapshots=1)
            got[name] = D.objective(lab(cand))
    finally:
        lab.close()
    start = J[top]
    gain_t = int((got["trained"] > start).sum())
    gain_u = int((got["untrained"] > start).sum())
    print(f"[surrogate] start mean {start.mean():.4f}; trained ascent mean "
          f"{got['trained'].mean():.4f} ({gain_t}/16 improved), untrained "
          f"{got['untrained'].mean():.4f} ({gain_u}/16 improved)", flush=True)
    assert gain_t >= 12 and gain_t > gain_u
    assert got["trained"].mean() > start.mean()
    assert got["trained"].mean() > got["untrained"].mean() + 0.005
