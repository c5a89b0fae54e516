"""Structures rasterised onto E-node centres; later structures win.

Semantics of ``src/photonfdtd/geometry.py:22-144``:

* :class:`Box` selects nodes with ``lo <= coord <= hi`` on every active
  axis (inclusive); a collapsed axis is inside iff ``lo <= 0 <= hi``.
* :class:`PolySlab` is an xy polygon extruded over ``z_bounds``; the xy
  point-in-polygon test is matplotlib's ``Path.contains_points`` on the
  node centres, exactly as the reference does, so boundary nodes agree.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Sequence, Tuple

import numpy as np

from .grid import Grid


def _pad3(x: Sequence[float]) -> Tuple[float, float, float]:
    vals = tuple(float(v) for v in x)
    if not 1 <= len(vals) <= 3:
        raise ValueError(f"Expected 1/2/3-tuple, got length {len(vals)}")
    return vals + (0.0,) * (3 - len(vals))


@dataclass(frozen=True)
class Box:
    center: Tuple[float, ...]
    size: Tuple[float, ...]
    medium: object

    def bounds(self) -> Tuple[Tuple[float, float], ...]:
        c, s = _pad3(self.center), _pad3(self.size)
        return tuple((ci - si / 2, ci + si / 2) for ci, si in zip(c, s))

    def region_mask(self, grid: Grid) -> np.ndarray:
        mask = np.ones(grid.shape, dtype=bool)
        for axis, (lo, hi) in enumerate(self.bounds()):
            coord = grid.coords[axis]
            if coord.size == 1:
                if not (lo <= 0.0 <= hi):
                    return np.zeros(grid.shape, dtype=bool)
                continue
            sel = (coord >= lo) & (coord <= hi)
            shape = [1, 1, 1]
            shape[axis] = sel.size
            mask &= sel.reshape(shape)
        return mask

    def stamp(self, grid: Grid, eps_r: np.ndarray) -> None:
        eps_r[self.region_mask(grid)] = self.medium.eps_r


@dataclass(frozen=True)
class PolySlab:
    vertices: Tuple[Tuple[float, float], ...]
    z_bounds: Tuple[float, float]
    medium: object

    def _z_layers(self, grid: Grid):
        verts = np.asarray(self.vertices, dtype=float)
        if verts.ndim != 2 or verts.shape[1] != 2 or verts.shape[0] < 3:
            raise ValueError("PolySlab.vertices must be a (N>=3, 2) array")
        z_lo, z_hi = sorted((float(self.z_bounds[0]), float(self.z_bounds[1])))
        zc = grid.coords[2]
        if zc.size == 1:
            return verts, (np.array([0]) if z_lo <= 0.0 <= z_hi else np.array([], dtype=int))
        return verts, np.flatnonzero((zc >= z_lo) & (zc <= z_hi))

    def _inside_xy(self, grid: Grid, verts: np.ndarray) -> np.ndarray:
        from matplotlib.path import Path
        xc, yc = grid.coords[0], grid.coords[1]
        XX, YY = np.meshgrid(xc, yc, indexing="ij")
        pts = np.column_stack([XX.ravel(), YY.ravel()])
        return Path(verts).contains_points(pts).reshape(xc.size, yc.size)

    def region_mask(self, grid: Grid) -> np.ndarray:
        mask = np.zeros(grid.shape, dtype=bool)
        verts, ks = self._z_layers(grid)
        if ks.size == 0:
            return mask
        inside = self._inside_xy(grid, verts)
        mask[:, :, ks] = inside[:, :, None]
        return mask

    def stamp(self, grid: Grid, eps_r: np.ndarray) -> None:
        mask = self.region_mask(grid)
        eps_r[mask] = self.medium.eps_r
