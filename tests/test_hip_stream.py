"""GPU: the beyond-HBM streaming stepper (pinned host domain, x-slab tiles
visiting the device T steps at a time) against the reference's fixtures and
against the device-resident stepper.

The reference checks its streaming CUDA stepper bit for bit against the
resident one (tests/test_rust.py:191-203, PHOTONFDTD_CUDA_STREAM=1 forced);
here float64 runs in reference order must reproduce the fixtures bit for bit
(<= 1e-12 asserted, 0.0 expected) for several tile shapes: tiles smaller than
the halo, a halo wider than the grid, chunks cut short by monitor steps.
float32 must give the resident stepper's bits. Features the streaming
stepper does not step are refused with NotImplementedError.
"""
import os

import numpy as np
import pytest

import golden_io as G
from scenes import FEATURE_SCENES, SCENES

pytestmark = pytest.mark.gpu

STREAM_SCENES = {k: SCENES[k] for k in ("tm2d_box", "te2d_poly", "box3d", "mode3d")}
STREAM_SCENES.update(FEATURE_SCENES)
FIXTURE_BITWISE = ("tm2d_box", "te2d_poly", "box3d", "mode3d")
KEYS = ("PHOTONFDTD_HIP_STREAM", "PHOTONFDTD_HIP_TB", "PHOTONFDTD_HIP_TILE", "PHOTONFDTD_HIP_EXACT")


def _run(key, precision, env, monkeypatch, sim=None):
    import photonfdtd_amd as pf
    for k in KEYS:
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    return (sim or STREAM_SCENES[key](pf, "hip", precision=precision)).run()


def _same(a, b):
    for name in a.fields:
        for c in a.fields[name]:
            assert np.array_equal(np.asarray(a.fields[name][c]),
                                  np.asarray(b.fields[name][c])), (name, c)
    for name in a.dft:
        for c in a.dft[name]:
            assert np.array_equal(a.dft[name][c], b.dft[name][c]), (name, c)


@pytest.mark.parametrize("tb,tile", [("3", "5"), ("7", "4"), ("40", "9"), ("32", "0")])
@pytest.mark.parametrize("key", FIXTURE_BITWISE)
def test_fp64_streaming_matches_reference(key, tb, tile, monkeypatch):
    print(f"\n[case] streaming {key} fp64 T={tb} tile={tile}", flush=True)
    got = _run(key, "float64", {"PHOTONFDTD_HIP_STREAM": "1", "PHOTONFDTD_HIP_TB": tb,
                                "PHOTONFDTD_HIP_TILE": tile}, monkeypatch)
    rep = {}
    worst = G.compare(got, G.load(key), 1e-12, rep)
    print(f"[stream] {key} T={tb} tile={tile} worst rel L2 = {worst:.3e}", flush=True)
    assert worst == 0.0


@pytest.mark.parametrize("key,precision", [("box3d", "float32"), ("tm2d_box", "float32"),
                                           ("unimode3d", "float64"), ("compressed2d", "float64")])
def test_streaming_same_bits_as_resident(key, precision, monkeypatch):
    import photonfdtd_amd as pf
    print(f"\n[case] streaming {key} {precision} vs resident", flush=True)
    # one scene object for both runs (the mode solver's Arnoldi start vector
    # is random, so a second solve moves the source amplitudes in the last bits)
    sim = STREAM_SCENES[key](pf, "hip", precision=precision)
    res = _run(key, precision, {"PHOTONFDTD_HIP_STREAM": "0"}, monkeypatch, sim)
    got = _run(key, precision, {"PHOTONFDTD_HIP_STREAM": "1", "PHOTONFDTD_HIP_TB": "5",
                                "PHOTONFDTD_HIP_TILE": "6"}, monkeypatch, sim)
    _same(got, res)


def test_stream_info_and_refusals(monkeypatch):
    import photonfdtd_amd as pf
    from photonfdtd_amd.hip_backend import HipEngine, exact_mode, validate
    for k in KEYS:
        monkeypatch.delenv(k, raising=False)
    monkeypatch.setenv("PHOTONFDTD_HIP_TILE", "6")
    print("\n[case] streaming info and refusals", flush=True)
    sim = SCENES["box3d"](pf, "hip")
    npdt = validate(sim)
    eng = HipEngine(sim, npdt, exact_mode(npdt), streaming=True)
    try:
        info = eng.info()
        assert info["path"] == 4 and info["tiles_bulk"] == (sim.grid.shape[0] + 5) // 6, info
        with pytest.raises(RuntimeError, match="host memory"):
            eng.field_tensor("Ey")
    finally:
        eng.close()
    for key in ("metals_xz", "subpixel2d", "flux2d"):
        s = SCENES[key](pf, "hip")
        with pytest.raises(NotImplementedError, match="device-resident"):
            HipEngine(s, validate(s), True, streaming=True)
