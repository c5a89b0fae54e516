"""Disk-backed, lossily compressed FieldMonitor snapshots
(``FieldMonitor(compression=...)``; reference ``src/photonfdtd/storage.py``).

A frame is stored as one float64 ``scale`` (max |snapshot|, 1 when that is 0
or not finite) and the snapshot divided by it, quantised to int8 (``bits=8``:
``rint(x / scale * 127)``) or float16 (``bits=16``), then byte-compressed
(zstd when the ``zstandard`` package imports, else zlib level 6). On the HIP
backend the quantisation runs on the GPU (``pf_read_box_q``); this module
receives the quantised bytes and only runs the codec and the file.

The series reads like an array of frames: ``len``, ``shape``, ``dtype``,
``series[i]`` (one frame, decompressed on demand), ``np.asarray(series)``,
iteration, ``nbytes`` (compressed) and ``nbytes_uncompressed``. Frames are
appended to a temporary file that is deleted on ``close()`` or collection.
"""
from __future__ import annotations

import os
import tempfile
import zlib
from typing import List, Tuple

import numpy as np

try:
    import zstandard as _zstd
except ImportError:
    _zstd = None


class ZlibCodec:
    name = "zlib"

    def __init__(self, level: int = 6):
        self.level = level

    def compress(self, data: bytes) -> bytes:
        return zlib.compress(data, self.level)

    def decompress(self, data: bytes) -> bytes:
        return zlib.decompress(data)


class ZstdCodec:
    name = "zstd"

    def __init__(self, level: int = 6):
        self._enc = _zstd.ZstdCompressor(level=level)
        self._dec = _zstd.ZstdDecompressor()

    def compress(self, data: bytes) -> bytes:
        return self._enc.compress(data)

    def decompress(self, data: bytes) -> bytes:
        return self._dec.decompress(data)


def get_codec(compression):
    """Codec for a ``FieldMonitor.compression`` value: ``True``/``"auto"``
    (zstd if available, else zlib), ``"zstd"`` or ``"zlib"``."""
    if compression in (True, "auto"):
        return ZstdCodec() if _zstd is not None else ZlibCodec()
    if compression == "zstd":
        if _zstd is None:
            raise ImportError("compression='zstd' needs the 'zstandard' package; use "
                              "compression='zlib', or compression=True to pick automatically")
        return ZstdCodec()
    if compression == "zlib":
        return ZlibCodec()
    raise ValueError(f"unknown compression {compression!r}; use True/'auto', 'zstd' or 'zlib'")


class CompressedFieldSeries:
    """Append-only stack of quantised, compressed snapshots in a temp file."""

    def __init__(self, snap_shape, out_dtype, codec, bits: int = 8, tmpdir=None):
        if bits not in (8, 16):
            raise ValueError(f"bits must be 8 or 16, got {bits!r}")
        self._shape = tuple(int(n) for n in snap_shape)
        self._dtype = np.dtype(out_dtype)
        self._codec = codec
        self._bits = int(bits)
        self._qtype = np.dtype(np.int8) if bits == 8 else np.dtype(np.float16)
        fd, self._path = tempfile.mkstemp(prefix="pfhip_mon_", suffix=".zfld", dir=tmpdir)
        self._file = os.fdopen(fd, "w+b")
        self._frames: List[Tuple[int, int, float]] = []    # (offset, length, scale)

    # -- writing ---------------------------------------------------------- #
    def append_quantized(self, q: np.ndarray, scale: float) -> None:
        """Append one frame already quantised (int8 or float16 values of the
        snapshot divided by ``scale``, in C order)."""
        q = np.ascontiguousarray(q)
        if q.dtype != self._qtype or q.shape != self._shape:
            raise ValueError(f"frame must be {self._qtype} of shape {self._shape}")
        blob = self._codec.compress(q.tobytes())
        offset = self._file.seek(0, os.SEEK_END)
        self._file.write(blob)
        self._frames.append((offset, len(blob), float(scale)))

    def finalize(self) -> None:
        self._file.flush()

    # -- reading ---------------------------------------------------------- #
    def _frame(self, i: int) -> np.ndarray:
        offset, length, scale = self._frames[i]
        self._file.seek(offset)
        q = np.frombuffer(self._codec.decompress(self._file.read(length)), dtype=self._qtype)
        unit = scale / 127.0 if self._bits == 8 else scale
        return q.reshape(self._shape).astype(self._dtype) * self._dtype.type(unit)

    def __len__(self) -> int:
        return len(self._frames)

    @property
    def shape(self) -> Tuple[int, ...]:
        return (len(self._frames),) + self._shape

    @property
    def ndim(self) -> int:
        return 1 + len(self._shape)

    @property
    def dtype(self) -> np.dtype:
        return self._dtype

    @property
    def nbytes(self) -> int:
        """Compressed size on disk."""
        return sum(length for _, length, _ in self._frames)

    @property
    def nbytes_uncompressed(self) -> int:
        return len(self._frames) * int(np.prod(self._shape)) * self._dtype.itemsize

    def __array__(self, dtype=None, copy=None) -> np.ndarray:
        out = np.empty(self.shape, dtype=self._dtype)
        for i in range(len(self._frames)):
            out[i] = self._frame(i)
        return out if dtype is None else out.astype(dtype)

    def __getitem__(self, idx):
        if isinstance(idx, (int, np.integer)):
            n = len(self._frames)
            i = int(idx) + (n if idx < 0 else 0)
            if not 0 <= i < n:
                raise IndexError(f"frame index {idx} out of range [0, {n})")
            return self._frame(i)
        return np.asarray(self)[idx]

    def __iter__(self):
        for i in range(len(self._frames)):
            yield self._frame(i)

    def close(self) -> None:
        if self._file is not None:
            try:
                self._file.close()
            finally:
                self._file = None
                try:
                    os.unlink(self._path)
                except OSError:
                    pass

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
