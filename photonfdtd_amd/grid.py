"""Uniform Cartesian Yee grid (scene description, no field storage).

Mirrors ``src/photonfdtd/grid.py:17-111``: the PML sits *inside* the stated
extent, each active axis holds ``round(size/cell)`` E-nodes centred on the
origin, and a zero/None extent collapses that axis to one node at 0.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np


def _triple(value, fill=None) -> Tuple:
    """Normalise a scalar / 1- / 2- / 3-sequence to a 3-tuple (``grid.py:99-111``)."""
    if value is None:
        return (fill, fill, fill)
    if isinstance(value, (int, float)):
        return (value, value, value)
    vals = tuple(value)
    if not 1 <= len(vals) <= 3:
        raise ValueError(f"Expected 1/2/3-tuple or scalar, got length {len(vals)}")
    return vals + (fill,) * (3 - len(vals))


@dataclass
class Grid:
    """Yee grid over ``size`` (m) with spacing ``cell_size`` (m).

    ``shape`` is always a 3-tuple; collapsed axes have one node. ``coords[a]``
    are the E-node coordinates ``(arange(n) - (n-1)/2) * d`` (``grid.py:55``).
    """
    size: Tuple[float, ...]
    cell_size: Tuple[float, ...]
    pml_layers: Tuple[int, ...] = (0, 0, 0)

    shape: Tuple[int, ...] = field(init=False)
    coords: Tuple[np.ndarray, ...] = field(init=False)
    ndim: int = field(init=False)

    def __post_init__(self) -> None:
        sizes = _triple(self.size)
        cells = _triple(self.cell_size)
        pmls = _triple(self.pml_layers, fill=0)
        active = [s is not None and s > 0 for s in sizes]
        if not any(active):
            raise ValueError("Grid must have at least one non-zero dimension")
        shape, coords, pml = [], [], []
        for axis in range(3):
            if not active[axis]:
                shape.append(1)
                coords.append(np.array([0.0]))
                pml.append(0)
                continue
            d = cells[axis]
            if d is None or d <= 0:
                raise ValueError(f"axis {axis}: cell_size must be > 0")
            n = int(round(sizes[axis] / d))
            if n < 4:
                raise ValueError(f"axis {axis}: need at least 4 cells, got {n}")
            shape.append(n)
            coords.append((np.arange(n) - (n - 1) / 2) * d)
            pml.append(int(pmls[axis]))
        self.size = tuple(0.0 if s is None else float(s) for s in sizes)
        self.cell_size = tuple(0.0 if d is None else float(d) for d in cells)
        self.pml_layers = tuple(pml)
        self.shape = tuple(shape)
        self.coords = tuple(coords)
        self.ndim = sum(active)

    @property
    def dx(self) -> float:
        return float(self.cell_size[0])

    @property
    def dy(self) -> float:
        return float(self.cell_size[1])

    @property
    def dz(self) -> float:
        return float(self.cell_size[2])

    @property
    def min_cell(self) -> float:
        return min(d for d in self.cell_size if d > 0)

    @property
    def n_cells(self) -> int:
        return int(self.shape[0] * self.shape[1] * self.shape[2])

    def index_at(self, point: Sequence[float]) -> Tuple[int, int, int]:
        """Nearest E-node index per axis; first node wins ties (``grid.py:82-96``)."""
        pt = list(point) + [0.0] * (3 - len(point))
        out = []
        for axis in range(3):
            c = self.coords[axis]
            out.append(0 if c.size == 1 else int(np.argmin(np.abs(c - pt[axis]))))
        return tuple(out)
