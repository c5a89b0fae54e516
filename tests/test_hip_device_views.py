"""GPU: zero-copy device handoff (pf_field_ptr / pf_dft_ptr) and the adjoint
with its DFT outputs, cotangents and gradient on the device.

The device views alias the stepper's memory: a field view equals the host
read of the same level bit for bit (on the 64-lane march's rows of 64-slot
multiples, on dense rows and on the generic path), a DFT view equals
pf_dft_read, and the padded 62-node layout refuses a strided view. The
adjoint fed from device tensors gives the host path's value and gradient.
"""
import numpy as np
import pytest

from test_hip_adjoint import loss_of, small_3d, te_2d

pytestmark = pytest.mark.gpu

KEYS = ("PHOTONFDTD_HIP_ZPAD", "PHOTONFDTD_HIP_SPLIT", "PHOTONFDTD_HIP_FUSED")


def _engine(pf, precision, env, monkeypatch, flags=0):
    from photonfdtd_amd.hip_backend import HipEngine, exact_mode, validate
    for k in KEYS:
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    sim = small_3d(pf, precision)
    npdt = validate(sim)
    return sim, HipEngine(sim, npdt, exact_mode(npdt), flags=flags)


@pytest.mark.parametrize("layout", ["march64", "dense", "generic"])
@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_field_and_dft_views_equal_host_reads(layout, precision, monkeypatch):
    import torch
    import photonfdtd_amd as pf
    from photonfdtd_amd import _native as N
    print(f"\n[case] device views {layout} {precision}", flush=True)
    env = {"dense": {"PHOTONFDTD_HIP_ZPAD": "0"}, "march64": {"PHOTONFDTD_HIP_ZPAD": "64"}}.get(layout, {})
    flags = N.PF_FLAG_GENERIC if layout == "generic" else 0
    sim, eng = _engine(pf, precision, env, monkeypatch, flags)
    try:
        eng.advance(0, min(37, sim.n_steps))
        info = eng.info()
        assert (info["path"] == 0) == (layout == "generic"), info
        for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
            t = eng.field_tensor(c)
            assert t.is_cuda and tuple(t.shape) == sim.grid.shape
            assert t.dtype == (torch.float32 if precision == "float32" else torch.float64)
            eng._cache.clear()
            assert np.array_equal(t.cpu().numpy(), eng.get(c)), c
        host = eng.dft("p")
        dev = eng.dft_tensors("p")
        assert set(dev) == set(host)
        for c in host:
            assert dev[c].is_cuda and dev[c].dtype == torch.complex128
            assert np.array_equal(dev[c].cpu().numpy(), host[c]), c
        assert np.abs(host["Ey"]).max() > 0
    finally:
        eng.close()


def test_padded_blocks_refuse_field_view(monkeypatch):
    import photonfdtd_amd as pf
    print("\n[case] device view refused on padded 62-node rows", flush=True)
    sim, eng = _engine(pf, "float32", {"PHOTONFDTD_HIP_ZPAD": "62"}, monkeypatch)
    try:
        eng.advance(0, min(3, sim.n_steps))
        assert eng.info()["zpad"] == 62
        with pytest.raises(RuntimeError, match="padded"):
            eng.field_tensor("Ey")
        assert eng.dft_tensors("p")["Ey"].is_cuda   # DFT accumulators are dense
    finally:
        eng.close()


@pytest.mark.parametrize("scene", ["small_3d", "te_2d"])
def test_adjoint_on_device_matches_host_path(scene):
    import photonfdtd_amd as pf
    from photonfdtd_amd.adjoint import value_and_grad_eps
    print(f"\n[case] adjoint device handoff {scene}", flush=True)
    build = small_3d if scene == "small_3d" else te_2d
    v_h, g_h = value_and_grad_eps(build(pf), loss_of, checkpoint_interval=7)
    v_d, g_d = value_and_grad_eps(build(pf), loss_of, checkpoint_interval=7, device="cuda")
    assert abs(v_d - v_h) <= 1e-12 * abs(v_h)
    rel = np.linalg.norm(g_d - g_h) / np.linalg.norm(g_h)
    print(f"[device] {scene}: value {v_d:.6e} vs {v_h:.6e}, grad rel L2 {rel:.2e}", flush=True)
    assert rel <= 1e-12
