"""GPU: the hand-written discrete adjoint against exact autograd gradients.

The checker is torch float64 autograd through ``oracle/fdtd_torch.py`` (the
same discrete forward map; its forward equals the NumPy oracle), plus central
differences of the HIP forward itself.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def small_3d(pf, precision="float64"):
    grid = pf.Grid(size=(12 * 80e-9, 10 * 80e-9, 9 * 80e-9), cell_size=80e-9, pml_layers=(3, 3, 3))
    box = pf.Box(center=(0.1e-6, 0, 0), size=(0.3e-6, 0.3e-6, 0.3e-6), medium=pf.Medium(3.0))
    src = pf.PointDipole(position=(-0.2e-6, 0, 0), component="Ey",
                         waveform=pf.GaussianPulse(freq0=pf.C_0 / 1e-6, fwhm=4e-15))
    dft = pf.DFTMonitor(name="p", components=("Ey", "Ez", "Hy", "Hz"),
                        freqs=[pf.C_0 / 1.1e-6, pf.C_0 / 0.9e-6], plane_axis="x",
                        plane_position=0.25e-6)
    return pf.Simulation(grid, structures=[box], sources=[src], monitors=[dft], run_time=5e-15,
                         precision=precision, backend="hip")


def tall_3d(pf, precision="float64"):
    """small_3d with 64 z cells: the adjoint's forward run takes the 64-lane
    march on dense rows (PF_FLAG_DENSE64)."""
    grid = pf.Grid(size=(12 * 80e-9, 10 * 80e-9, 64 * 80e-9), cell_size=80e-9, pml_layers=(3, 3, 3))
    box = pf.Box(center=(0.1e-6, 0, 0.4e-6), size=(0.3e-6, 0.3e-6, 0.5e-6), medium=pf.Medium(3.0))
    src = pf.PointDipole(position=(-0.2e-6, 0, 0.2e-6), component="Ey",
                         waveform=pf.GaussianPulse(freq0=pf.C_0 / 1e-6, fwhm=4e-15))
    dft = pf.DFTMonitor(name="p", components=("Ey", "Ez", "Hy", "Hz"),
                        freqs=[pf.C_0 / 1.1e-6, pf.C_0 / 0.9e-6], plane_axis="x",
                        plane_position=0.25e-6)
    return pf.Simulation(grid, structures=[box], sources=[src], monitors=[dft], run_time=5e-15,
                         precision=precision, backend="hip")


def te_2d(pf):
    grid = pf.Grid(size=(1.2e-6, 1.0e-6), cell_size=(40e-9, 40e-9), pml_layers=(6, 6, 0))
    poly = pf.PolySlab(vertices=((-0.2e-6, -0.15e-6), (0.25e-6, -0.1e-6), (0.1e-6, 0.2e-6)),
                       z_bounds=(-1, 1), medium=pf.Medium(5.0))
    src = pf.PointDipole(position=(-0.35e-6, 0.05e-6), component="Ex",
                         waveform=pf.GaussianPulse(freq0=pf.C_0 / 1e-6, fwhm=4e-15))
    hsrc = pf.PointDipole(position=(-0.3e-6, -0.1e-6), component="Hz", amplitude=1e-3,
                          waveform=pf.GaussianPulse(freq0=pf.C_0 / 1e-6, fwhm=4e-15))
    dft = pf.DFTMonitor(name="p", components=("Ey", "Hz"), freqs=[pf.C_0 / 1e-6],
                        plane_axis="x", plane_position=0.3e-6, interval=2)
    full = pf.DFTMonitor(name="f", components=("Ex",), freqs=[pf.C_0 / 0.95e-6], downsample=2)
    return pf.Simulation(grid, structures=[poly], sources=[src, hsrc], monitors=[dft, full],
                         run_time=12e-15, backend="hip")


def loss_of(out):
    tot = 0
    for name, comps in out["dft"].items():
        for c, v in comps.items():
            w = 1.0 + 0.5 * (c[0] == "H")
            tot = tot + w * (v.abs() ** 2).sum() + (v.real * v.imag).sum()
    return tot


def reference_grad(sim):
    from oracle.fdtd_torch import run_torch
    eps = torch.tensor(np.asarray(sim.eps_r, dtype=np.float64), requires_grad=True)
    val = loss_of({"dft": run_torch(sim, eps)})
    val.backward()
    return float(val.detach()), eps.grad.numpy()


@pytest.mark.parametrize("scene", ["small_3d", "te_2d", "tall_3d"])
def test_adjoint_matches_exact_autograd(scene):
    import photonfdtd_amd as pf
    from photonfdtd_amd.adjoint import value_and_grad_eps
    print(f"\n[case] adjoint vs torch autograd {scene}", flush=True)
    build = {"small_3d": small_3d, "te_2d": te_2d, "tall_3d": tall_3d}[scene]
    v_ref, g_ref = reference_grad(build(pf))
    v, g = value_and_grad_eps(build(pf), loss_of, checkpoint_interval=7)
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print(f"[adjoint] {scene}: value rel {abs(v - v_ref) / abs(v_ref):.2e}, grad rel L2 {rel:.2e}",
          flush=True)
    assert abs(v - v_ref) <= 1e-12 * abs(v_ref)
    assert rel <= 1e-9


def test_checkpoint_interval_does_not_change_the_gradient():
    import photonfdtd_amd as pf
    from photonfdtd_amd.adjoint import value_and_grad_eps
    print("\n[case] adjoint checkpoint intervals", flush=True)
    _, g1 = value_and_grad_eps(small_3d(pf), loss_of, checkpoint_interval=1)
    _, g2 = value_and_grad_eps(small_3d(pf), loss_of, checkpoint_interval=1000)
    assert np.array_equal(g1, g2)


def test_adjoint_float32_close_to_float64():
    """float32 forward + adjoint against float64: the gradient is a sum over
    steps of lambda * curl products with cancellation, so it carries ~1e-3
    relative float32 round-off on this scene (finite-difference checks of
    inverse-design gradients run in float64)."""
    import photonfdtd_amd as pf
    from photonfdtd_amd.adjoint import value_and_grad_eps
    print("\n[case] adjoint float32", flush=True)
    _, g64 = value_and_grad_eps(small_3d(pf), loss_of)
    _, g32 = value_and_grad_eps(small_3d(pf, "float32"), loss_of)
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"[adjoint] float32 grad rel L2 vs float64 {rel:.2e}", flush=True)
    assert rel <= 5e-3


def test_adjoint_central_difference_on_hip_forward():
    import photonfdtd_amd as pf
    from photonfdtd_amd.adjoint import value_and_grad_eps
    print("\n[case] adjoint vs central differences", flush=True)
    sim = small_3d(pf)
    eps0 = np.asarray(sim.eps_r, dtype=np.float64)
    _, g = value_and_grad_eps(sim, loss_of)

    def value(eps):
        s = small_3d(pf)
        s.eps_r = eps
        r = s.run()
        return float(loss_of({"dft": {n: {c: torch.from_numpy(v) for c, v in d.items()}
                                      for n, d in r.dft.items()}}))
    for idx in [(6, 5, 4), (7, 4, 4), (5, 5, 5)]:
        h = 1e-4
        ep, em = eps0.copy(), eps0.copy()
        ep[idx] += h
        em[idx] -= h
        fd = (value(ep) - value(em)) / (2 * h)
        assert fd == pytest.approx(g[idx], rel=1e-5)


def test_dispersive_scene_is_rejected():
    import photonfdtd_amd as pf
    from photonfdtd_amd.adjoint import value_and_grad_eps
    sim = small_3d(pf)
    sim.structures.append(pf.Box(center=(0, 0, 0), size=(0.1e-6,) * 3,
                                 medium=pf.DispersiveMedium.drude(1.0, 50e12, 5e12)))
    sim._has_dispersion = True
    with pytest.raises(NotImplementedError):
        value_and_grad_eps(sim, loss_of)


def test_dense64_forward_runs_the_march_and_matches_generic():
    """With nz a multiple of 64 the checkpointed forward run and the
    reverse sweep's segment recompute are the 64-lane march (fused path,
    dense rows of nz slots, trajectory slots as its levels); the float64
    gradient has the bits of the generic-kernel run."""
    import photonfdtd_amd as pf
    from photonfdtd_amd import _native as N
    from photonfdtd_amd.adjoint import _checkpointed_engine, value_and_grad_eps
    from photonfdtd_amd.hip_backend import HipEngine, exact_mode, validate
    print("\n[case] adjoint dense-64 march forward vs generic", flush=True)
    sim = tall_3d(pf)
    eng = _checkpointed_engine(sim, validate(sim), 7)
    info = eng.info()
    eng.close()
    assert info["path"] == 2 and info["zpad"] == 64 and info["zrow"] == 64, info
    _, g_march = value_and_grad_eps(tall_3d(pf), loss_of, checkpoint_interval=7)

    import photonfdtd_amd.adjoint as adj

    def generic_only(run, npdt, interval):
        e = HipEngine(run, npdt, exact_mode(npdt), flags=N.PF_FLAG_GENERIC)
        e.enable_checkpoints(interval)
        return e
    orig = adj._checkpointed_engine
    adj._checkpointed_engine = generic_only
    try:
        _, g_gen = value_and_grad_eps(tall_3d(pf), loss_of, checkpoint_interval=7)
    finally:
        adj._checkpointed_engine = orig
    assert np.array_equal(g_march, g_gen)


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_march_recompute_same_gradient_as_generic_recompute(precision, monkeypatch):
    """On the dense 64-slot rows the reverse sweep recomputes each checkpoint
    segment with the forward's march (trajectory slots as its levels, the
    E-side psi copied into the output slot first); with
    PHOTONFDTD_HIP_ADJ_MARCH=0 it takes the generic kernels from the same
    checkpoints. Both give the same gradient bits (float64 in reference order
    and float32 in the fast order, checkpoint interval 7: a short last
    segment)."""
    import photonfdtd_amd as pf
    from photonfdtd_amd.adjoint import value_and_grad_eps
    print(f"\n[case] adjoint march vs generic recompute {precision}", flush=True)
    monkeypatch.delenv("PHOTONFDTD_HIP_ADJ_MARCH", raising=False)
    _, g_march = value_and_grad_eps(tall_3d(pf, precision), loss_of, checkpoint_interval=7)
    monkeypatch.setenv("PHOTONFDTD_HIP_ADJ_MARCH", "0")
    _, g_gen = value_and_grad_eps(tall_3d(pf, precision), loss_of, checkpoint_interval=7)
    g_march, g_gen = np.asarray(g_march), np.asarray(g_gen)
    assert np.abs(g_march).max() > 0
    assert np.array_equal(g_march, g_gen)
