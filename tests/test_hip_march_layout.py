"""GPU: the 3-D march gives the same bits under every layout and launch shape.

On grids below 2^25 nodes (this scene) the march keeps padded rows (node k
in slot k + 2 * (k // 62)) with a psi-free and a CPML launch; from 2^25
nodes it runs 64 lanes per z tile on rows of a multiple of 64 slots (node k
in slot k) with one launch per CPML axis set. Every layout (those two and
dense rows with 62-node windows), a single launch, no overlap and other
tile shapes must give identical fields, DFT planes, flux totals and field
snapshots; snapshots and DFT planes read the rows through the same slot
mapping, and write_field / read_field round-trip through it.
"""
import numpy as np
import pytest

from test_hip_pair import _scene

pytestmark = pytest.mark.gpu

CONFIGS = {
    "dense single launch": {"PHOTONFDTD_HIP_ZPAD": "0", "PHOTONFDTD_HIP_SPLIT": "0"},
    "64-lane split": {"PHOTONFDTD_HIP_ZPAD": "64"},
    "64-lane single launch": {"PHOTONFDTD_HIP_ZPAD": "64", "PHOTONFDTD_HIP_SPLIT": "0"},
    "64-lane split, no overlap": {"PHOTONFDTD_HIP_ZPAD": "64", "PHOTONFDTD_HIP_OVERLAP": "0"},
    "64-lane split, 5-plane CPML chunks, 10 waves": {"PHOTONFDTD_HIP_ZPAD": "64",
                                                     "PHOTONFDTD_HIP_PML_XCHUNK": "5",
                                                     "PHOTONFDTD_HIP_NW": "10",
                                                     "PHOTONFDTD_HIP_PML_NW": "10"},
    "64-lane split, 16 waves, 7-plane chunks": {"PHOTONFDTD_HIP_ZPAD": "64",
                                                "PHOTONFDTD_HIP_NW": "16",
                                                "PHOTONFDTD_HIP_XCHUNK": "7"},
    "dense split": {"PHOTONFDTD_HIP_ZPAD": "0", "PHOTONFDTD_HIP_SPLIT": "1"},
    "padded split, no overlap": {"PHOTONFDTD_HIP_ZPAD": "62", "PHOTONFDTD_HIP_OVERLAP": "0"},
    "padded single launch": {"PHOTONFDTD_HIP_ZPAD": "62", "PHOTONFDTD_HIP_SPLIT": "0"},
}
KEYS = ("PHOTONFDTD_HIP_ZPAD", "PHOTONFDTD_HIP_SPLIT", "PHOTONFDTD_HIP_OVERLAP",
        "PHOTONFDTD_HIP_PML_XCHUNK", "PHOTONFDTD_HIP_NW", "PHOTONFDTD_HIP_PML_NW",
        "PHOTONFDTD_HIP_XCHUNK", "PHOTONFDTD_HIP_PAIR")


def _run(pf, precision, env, monkeypatch):
    for k in KEYS:
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    sim = _scene(pf, precision)
    sim.monitors.append(pf.FluxMonitor(name="fx", plane_axis="x", plane_position=0.3e-6))
    sim.monitors.append(pf.FluxMonitor(name="fz", plane_axis="z", plane_position=-1.1e-6))
    return sim.run()


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_layouts_and_launch_shapes_agree(precision, monkeypatch):
    import photonfdtd_amd as pf
    monkeypatch.setenv("PHOTONFDTD_HIP_EXACT", "1")
    print(f"\n[case] march layouts {precision} default", flush=True)
    ref = _run(pf, precision, {}, monkeypatch)
    assert np.abs(ref.fields["s"]["Ez"]).max() > 0
    assert all(np.isfinite(np.asarray(v)).all() for v in ref.flux.values())
    for label, env in CONFIGS.items():
        print(f"[case] march layouts {precision} {label}", flush=True)
        got = _run(pf, precision, env, monkeypatch)
        for c in ref.fields["s"]:
            assert np.array_equal(got.fields["s"][c], ref.fields["s"][c]), (label, c)
        for m in ref.dft:
            for c in ref.dft[m]:
                assert np.array_equal(got.dft[m][c], ref.dft[m][c]), (label, m, c)
        for m in ref.flux:
            assert np.array_equal(np.asarray(got.flux[m]), np.asarray(ref.flux[m])), (label, m)


@pytest.mark.parametrize("zpad", [None, "64"])
def test_write_and_read_field_through_padded_rows(zpad, monkeypatch):
    """pf_write_field / pf_read_box on a stepper whose rows carry slots past
    nz (padded 62-node blocks, the default here; 64-slot rows): a written
    field reads back unchanged."""
    import photonfdtd_amd as pf
    from photonfdtd_amd.hip_backend import HipEngine
    for k in KEYS:
        monkeypatch.delenv(k, raising=False)
    if zpad:
        monkeypatch.setenv("PHOTONFDTD_HIP_ZPAD", zpad)
    print(f"\n[case] write/read round trip, zpad {zpad or 'default'}", flush=True)
    sim = _scene(pf, "float32")
    eng = HipEngine(sim, np.float32, True)
    try:
        rng = np.random.default_rng(3)
        v = rng.standard_normal(sim.grid.shape).astype(np.float32)
        eng.set("Hy", v)
        assert np.array_equal(eng.get("Hy"), v)
    finally:
        eng.close()
