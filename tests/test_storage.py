"""Compressed FieldMonitor storage (reference ``storage.py:104-219``) against
the fixture the reference's NumPy path wrote for ``compressed2d``: the
oracle's uncompressed float64 snapshots, quantised as the fixture's frames
say (int8 ``rint(x / max|x| * 127)``, float16 ``x / max|x|``), go through
``CompressedFieldSeries`` and come back as the fixture's frames bit for bit,
with the same compressed byte counts (same zlib stream)."""
import numpy as np
import pytest

import golden_io as G
from scenes import compressed2d


def _quantise(snap, bits):
    scale = float(np.abs(snap).max())
    if scale == 0.0 or not np.isfinite(scale):
        scale = 1.0
    q = np.rint(snap / scale * 127.0).astype(np.int8) if bits == 8 else (snap / scale).astype(np.float16)
    return q, scale


@pytest.fixture(scope="module")
def oracle_and_gold():
    import photonfdtd_amd as pf
    from oracle.fdtd_numpy import run_oracle
    return run_oracle(compressed2d(pf, "hip")), G.load("compressed2d")


@pytest.mark.parametrize("name,bits", [("q8", 8), ("q16", 16)])
def test_series_reproduces_reference_frames(oracle_and_gold, name, bits):
    from photonfdtd_amd.storage import CompressedFieldSeries, get_codec
    ref, gold = oracle_and_gold
    for c, snaps in ref["fields"][name].items():
        want = gold[f"field/{name}/{c}"]
        assert snaps.shape == want.shape
        series = CompressedFieldSeries(snaps.shape[1:], np.float64, get_codec("zlib"), bits=bits)
        for snap in snaps:
            series.append_quantized(*_quantise(snap, bits))
        series.finalize()
        got = np.asarray(series)
        assert np.array_equal(got, want), (name, c)
        assert series.nbytes == int(gold[f"zbytes/{name}/{c}"])
        assert len(series) == want.shape[0] and series.shape == want.shape
        assert np.array_equal(series[-1], want[-1])
        # quantisation error bound of the stored frames
        tol = (1 / 254 if bits == 8 else 2.0 ** -11) * np.abs(snaps).max(axis=(1, 2, 3))
        assert np.all(np.abs(got - snaps).max(axis=(1, 2, 3)) <= tol * (1 + 1e-12))
        series.close()


def test_codec_selection():
    from photonfdtd_amd import storage
    assert storage.get_codec("zlib").name == "zlib"
    assert storage.get_codec(True).name in ("zlib", "zstd")
    with pytest.raises(ValueError):
        storage.get_codec("lz4")
    if storage._zstd is None:
        with pytest.raises(ImportError):
            storage.get_codec("zstd")


def test_series_rejects_wrong_frames():
    from photonfdtd_amd.storage import CompressedFieldSeries, get_codec
    s = CompressedFieldSeries((4, 3, 1), np.float32, get_codec("zlib"), bits=8)
    with pytest.raises(ValueError):
        s.append_quantized(np.zeros((4, 3, 1), dtype=np.float16), 1.0)
    with pytest.raises(ValueError):
        CompressedFieldSeries((4,), np.float32, get_codec("zlib"), bits=4)
    s.append_quantized(np.full((4, 3, 1), -127, dtype=np.int8), 2.0)
    assert s.dtype == np.float32 and np.all(s[0] == -2.0)
    with pytest.raises(IndexError):
        s[1]
    s.close()
