"""Monitor descriptions (``src/photonfdtd/monitors.py:17-191``).

* :class:`FieldMonitor` records snapshots every ``interval`` steps or at the
  steps nearest ``times``; optional spatial ``downsample`` and ``plane_z``.
* :class:`DFTMonitor` keeps a running complex128 transform
  ``F(w) = sum_n f(t_n) e^{+i w t_n} dt`` with E sampled at ``(n+1) dt`` and
  H at ``(n+1/2) dt``; optional plane restriction on any axis.
* :class:`FluxMonitor` integrates ``(E x H) . n dA dt`` through a plane.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

_AXES = {"x": 0, "y": 1, "z": 2}


@dataclass
class FieldMonitor:
    name: str
    components: Tuple[str, ...] = ("Ez",)
    interval: Optional[int] = None
    times: Optional[Sequence[float]] = None
    downsample: int = 1
    plane_z: Optional[float] = None
    compression: object = None
    compression_bits: int = 8

    def __post_init__(self):
        if self.interval is None and self.times is None:
            self.interval = 1
        if self.interval is not None and self.times is not None:
            raise ValueError("Specify only one of interval or times")
        if int(self.downsample) != self.downsample or self.downsample < 1:
            raise ValueError("downsample must be a positive integer")
        self.downsample = int(self.downsample)
        if self.compression is not None and self.compression not in (True, "auto", "zstd", "zlib"):
            raise ValueError("compression must be None, True, 'zstd', or 'zlib', "
                             f"got {self.compression!r}")
        if self.compression_bits not in (8, 16):
            raise ValueError("compression_bits must be 8 or 16")


@dataclass
class DFTMonitor:
    name: str
    components: Tuple[str, ...] = ("Ez",)
    freqs: Sequence[float] = ()
    interval: int = 1
    downsample: int = 1
    plane_z: Optional[float] = None
    plane_axis: Optional[str] = None
    plane_position: Optional[float] = None

    def __post_init__(self):
        self.freqs = tuple(float(f) for f in self.freqs)
        if not self.freqs:
            raise ValueError("DFTMonitor requires at least one frequency in `freqs`")
        if int(self.interval) != self.interval or self.interval < 1:
            raise ValueError("interval must be a positive integer")
        self.interval = int(self.interval)
        if int(self.downsample) != self.downsample or self.downsample < 1:
            raise ValueError("downsample must be a positive integer")
        self.downsample = int(self.downsample)
        if self.plane_z is not None and self.plane_axis is None:
            self.plane_axis, self.plane_position = "z", self.plane_z
        if self.plane_axis is not None:
            if self.plane_axis not in _AXES:
                raise ValueError("plane_axis must be 'x', 'y', or 'z'")
            if self.plane_position is None:
                raise ValueError("plane_axis requires plane_position")

    def plane(self):
        """``(axis_index, position)`` of the plane restriction, or None."""
        if self.plane_axis is None:
            return None
        return (_AXES[self.plane_axis], float(self.plane_position))


@dataclass
class FluxMonitor:
    name: str
    plane_axis: str
    plane_position: float
