"""Anisotropic subpixel smoothing of material interfaces
(``Simulation(subpixel=True)``; reference ``src/photonfdtd/smoothing.py:119-172``).

A cell cut by an interface gets an effective diagonal permittivity tensor
instead of the value at its centre (Farjadpour et al., Opt. Lett. 31, 2972
(2006)): with ``eps_a`` the arithmetic and ``eps_h`` the harmonic mean of the
permittivity over the cell and ``n`` the unit interface normal,

    eps_ii = eps_a + (eps_h - eps_a) n_i^2        (i = x, y, z),

and ``eps_xx / eps_yy / eps_zz`` drive the ``Ex / Ey / Ez`` updates. The means
come from supersampling each cell ``factor`` times per active axis (the
structures are stamped on the fine nodes, one native x-slab at a time so
memory stays at one refined slab); ``n`` is the normalised gradient of
``eps_a`` (zero where that gradient vanishes). Cells inside one material get
exactly that material's permittivity. The reductions are NumPy's, in the
reference's order, so the tensors are the reference's bits.
"""
from __future__ import annotations

from types import SimpleNamespace
from typing import Iterator, Sequence, Tuple

import numpy as np

from .grid import Grid


def _refined(grid: Grid, factor: int) -> Grid:
    """The same domain with ``factor`` nodes per native cell on every active
    axis (collapsed axes stay collapsed; no PML)."""
    cells = tuple(d / factor if d > 0 else 0.0 for d in grid.cell_size)
    return Grid(size=grid.size, cell_size=cells, pml_layers=(0, 0, 0))


def _fine_slabs(grid: Grid, structures, background: float,
                factor: int) -> Iterator[Tuple[int, np.ndarray]]:
    """``(i, eps)`` per native x index: the permittivity stamped on the fine
    nodes of that slab, shaped ``(1, fx, ny, fy, nz, fz)``."""
    fine = _refined(grid, factor)
    xf, yf, zf = fine.coords
    nx, ny, nz = grid.shape
    fx, fy, fz = (factor if n > 1 else 1 for n in grid.shape)
    for i in range(nx):
        xs = xf[i * fx:(i + 1) * fx]
        view = SimpleNamespace(shape=(xs.size, yf.size, zf.size), coords=(xs, yf, zf),
                               cell_size=fine.cell_size)
        eps = np.full(view.shape, float(background), dtype=np.float64)
        for st in structures:
            eps[st.region_mask(view)] = st.medium.eps_r
        yield i, eps.reshape(1, fx, ny, fy, nz, fz)


def cell_means(grid: Grid, structures, background: float = 1.0,
               factor: int = 3) -> Tuple[np.ndarray, np.ndarray]:
    """Arithmetic and harmonic mean permittivity of every native cell."""
    arith = np.empty(grid.shape, dtype=np.float64)
    harm = np.empty(grid.shape, dtype=np.float64)
    sub_axes = (1, 3, 5)
    for i, eps in _fine_slabs(grid, structures, background, factor):
        arith[i] = eps.mean(axis=sub_axes)[0]
        harm[i] = (1.0 / (1.0 / eps).mean(axis=sub_axes))[0]
    return arith, harm


def smooth_permittivity(grid: Grid, structures: Sequence, background_eps: float = 1.0,
                        factor: int = 3, dtype=np.float64):
    """``(eps_xx, eps_yy, eps_zz)`` on ``grid.shape`` for the Ex / Ey / Ez
    updates. ``factor`` (>= 1) is the supersampling per active axis; 1
    reproduces plain stamping."""
    if factor < 1:
        raise ValueError("factor must be >= 1")
    arith, harm = cell_means(grid, structures, background_eps, factor)
    shape = grid.shape
    grads = []
    for a in range(3):
        if shape[a] > 1:
            grads.append(np.gradient(arith, grid.cell_size[a], axis=a))
        else:
            grads.append(np.zeros_like(arith))
    norm = np.sqrt(grads[0] ** 2 + grads[1] ** 2 + grads[2] ** 2)
    floor = max(norm.max() * 1e-9 if norm.size else 0.0, np.finfo(np.float64).tiny)
    has_normal = norm > floor
    inv = np.where(has_normal, 1.0 / np.where(has_normal, norm, 1.0), 0.0)
    jump = harm - arith
    return tuple((arith + jump * (g * inv) ** 2).astype(dtype, copy=False) for g in grads)
