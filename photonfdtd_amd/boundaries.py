"""Convolutional PML (Roden & Gedney 2000, kappa = 1) coefficient tables.

Restates ``src/photonfdtd/boundaries.py:23-97``: cubic sigma grading with
``sigma_max = sigma_factor * (m+1) ln(1/R) / (2 eta0 n d)``, linear alpha
grading, ``b = exp(-(sigma+alpha) dt/eps0)`` and
``c = sigma/(sigma+alpha) (b-1)``. E coefficients sit on integer nodes, H
coefficients on the half-integer nodes; both are zero outside the slabs,
which is what lets the stepper keep psi only where ``c != 0``.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Tuple

import numpy as np

from .constants import EPS_0, ETA_0


@dataclass(frozen=True)
class CPMLParams:
    poly_order: int = 3
    sigma_factor: float = 0.8
    alpha_max: float = 0.24
    R_target: float = 1e-7


def _graded_bc(u: np.ndarray, sigma_max: float, params: CPMLParams, dt: float):
    """(b, c) for normalised depths ``u`` in [0, 1] (0 = inner PML edge)."""
    sigma = sigma_max * u ** params.poly_order
    alpha = params.alpha_max * (1.0 - u)
    loss = sigma + alpha
    b = np.exp(-loss * dt / EPS_0)
    c = np.zeros_like(sigma)
    ok = loss > 0
    c[ok] = sigma[ok] / loss[ok] * (b[ok] - 1.0)
    return b, c


def cpml_coeffs_1axis(n_cells: int, cell_size: float, n_pml: int, dt: float,
                      params: CPMLParams = CPMLParams()
                      ) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """Per-axis ``(b_e, c_e, b_h, c_h)`` float64 arrays of length ``n_cells``."""
    be = np.zeros(n_cells)
    ce = np.zeros(n_cells)
    bh = np.zeros(n_cells)
    ch = np.zeros(n_cells)
    if n_pml == 0:
        return be, ce, bh, ch
    m = params.poly_order
    sigma_max = params.sigma_factor * (-(m + 1) * np.log(params.R_target)
                                       / (2.0 * ETA_0 * n_pml * cell_size))
    span = max(n_pml - 1, 1)
    node = np.arange(n_cells, dtype=np.float64)
    for offset, b_out, c_out in ((0.0, be, ce), (0.5, bh, ch)):
        pos = node + offset
        lo = np.clip((n_pml - 1 - pos[:n_pml]) / span, 0.0, 1.0)
        hi = np.clip((pos[-n_pml:] - (n_cells - n_pml)) / span, 0.0, 1.0)
        b_out[:n_pml], c_out[:n_pml] = _graded_bc(lo, sigma_max, params, dt)
        b_out[-n_pml:], c_out[-n_pml:] = _graded_bc(hi, sigma_max, params, dt)
    return be, ce, bh, ch
