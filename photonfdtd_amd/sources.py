"""Soft additive sources (``src/photonfdtd/sources.py:39-337``).

* :class:`GaussianPulse` ``w(t) = exp(-((t-t0)/sigma)^2) sin(2 pi f0 (t-t0))``
  with ``sigma = fwhm / (2 sqrt(2 ln 2))`` and ``t0 = 4 sigma + delay``.
* :class:`PointDipole` adds ``amplitude * w(t)`` to one component at the
  nearest node, H at ``(n+1/2) dt`` and E at ``(n+1) dt``.
* :class:`ModeSource` samples a transverse profile onto the injection plane
  and becomes one dipole per plane node with a non-zero weight.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np


@dataclass(frozen=True)
class GaussianPulse:
    freq0: float
    fwhm: float
    delay: float = 0.0

    @property
    def sigma(self) -> float:
        return float(self.fwhm) / (2.0 * math.sqrt(2.0 * math.log(2.0)))

    @property
    def t0(self) -> float:
        return 4.0 * self.sigma + float(self.delay)

    @property
    def envelope_l2(self) -> float:
        return float(self.sigma) * math.sqrt(math.pi / 2.0)

    def __call__(self, t):
        t = np.asarray(t)
        shifted = t - self.t0
        return np.exp(-(shifted / self.sigma) ** 2) * np.sin(2.0 * math.pi * self.freq0 * shifted)


@dataclass(frozen=True)
class PointDipole:
    position: Tuple[float, ...]
    component: str
    waveform: GaussianPulse
    amplitude: float = 1.0


@dataclass(frozen=True)
class ModeSource:
    """Profile-weighted soft source on a plane; the zero-size axis propagates.

    ``profile_coords`` are relative to ``center``; the profile is linearly
    interpolated (zero outside its support) onto the plane nodes, one axis at
    a time for a 2-D profile (``sources.py:279-307``).
    """
    center: Tuple[float, float, float]
    size: Tuple[float, float, float]
    component: str
    waveform: GaussianPulse
    profile: np.ndarray
    profile_coords: Tuple[np.ndarray, ...]
    amplitude: float = 1.0

    def _propagation_axis(self) -> int:
        sizes = list(self.size) + [0.0] * (3 - len(self.size))
        zeros = [a for a, s in enumerate(sizes) if s == 0]
        return zeros[0] if zeros else int(np.argmin(sizes))

    def expand(self, grid) -> List[PointDipole]:
        center = list(self.center) + [0.0] * (3 - len(self.center))
        size = list(self.size) + [0.0] * (3 - len(self.size))
        prop = self._propagation_axis()
        tan = [a for a in range(3) if a != prop]

        nodes, node_pos = [], []
        for ax in tan:
            c = grid.coords[ax]
            if c.size == 1:
                nodes.append(np.array([0]))
                node_pos.append(np.array([0.0]))
                continue
            sel = (c >= center[ax] - size[ax] / 2.0) & (c <= center[ax] + size[ax] / 2.0)
            if not sel.any():
                raise ValueError(f"ModeSource injection plane has zero overlap on axis {ax}")
            nodes.append(np.flatnonzero(sel))
            node_pos.append(c[sel])

        c_prop = grid.coords[prop]
        i_prop = 0 if c_prop.size == 1 else int(np.argmin(np.abs(c_prop - center[prop])))

        prof = np.asarray(self.profile)
        rel0 = node_pos[0] - center[tan[0]]
        rel1 = node_pos[1] - center[tan[1]]
        if prof.ndim == 1:
            if len(self.profile_coords) != 1:
                raise ValueError("1-D profile requires one profile_coords array")
            if node_pos[1].size != 1:
                raise ValueError("a 1-D profile needs a line-shaped injection plane "
                                 "(one node on the second tangent axis)")
            weights = np.interp(rel0, self.profile_coords[0], prof,
                                left=0.0, right=0.0)[:, None]
        elif prof.ndim == 2:
            if len(self.profile_coords) != 2:
                raise ValueError("2-D profile requires two profile_coords arrays")
            along1 = np.empty((prof.shape[0], rel1.size))
            for r in range(prof.shape[0]):
                along1[r] = np.interp(rel1, self.profile_coords[1], prof[r], left=0.0, right=0.0)
            weights = np.empty((rel0.size, rel1.size))
            for col in range(rel1.size):
                weights[:, col] = np.interp(rel0, self.profile_coords[0], along1[:, col],
                                            left=0.0, right=0.0)
        else:
            raise ValueError("profile must be 1-D or 2-D")

        out: List[PointDipole] = []
        for a, i_t0 in enumerate(nodes[0]):
            for b, i_t1 in enumerate(nodes[1]):
                w = float(weights[a, b]) * float(self.amplitude)
                if w == 0.0:
                    continue
                pos = [0.0, 0.0, 0.0]
                pos[prop] = float(c_prop[i_prop] if c_prop.size > 1 else center[prop])
                c0 = grid.coords[tan[0]]
                pos[tan[0]] = float(c0[i_t0] if c0.size > 1 else center[tan[0]])
                c1 = grid.coords[tan[1]]
                pos[tan[1]] = float(c1[i_t1] if c1.size > 1 else center[tan[1]])
                out.append(PointDipole(position=tuple(pos), component=self.component,
                                       waveform=self.waveform, amplitude=w))
        return out


def _nearest(coords: np.ndarray, x: float) -> int:
    return 0 if coords.size == 1 else int(np.argmin(np.abs(coords - x)))


@dataclass
class UniModeSource:
    """One-way launch of a solved waveguide mode along +x or -x
    (reference ``sources.py:340-433``).

    Love's equivalence: an electric sheet ``J = x_hat x H_m`` and a magnetic
    sheet ``M = -x_hat x E_m`` on the injection plane radiate the mode forward
    and cancel it backward. As soft sources on the plane's nodes (E at the
    E time, H half a step later, as the stepper staggers them):

        Ey += C Hz_m / (eps0 eps_r)      Ez += -C Hy_m / (eps0 eps_r)
        Hy += -s C Ez_m / mu0            Hz +=  s C Ey_m / mu0

    with ``s = +1`` for "+x", ``-1`` for "-x". The mode's transverse fields
    are first rotated by one phase so that Ey + Ez is real at the node where
    |Ey| + |Ez| peaks, then their real parts are used. Each node of the mode
    grid maps to the nearest node of the simulation plane through ``center``;
    per mode node the four components are emitted in the order above, skipping
    zero amplitudes. ``mode`` is a ``ModeResult`` solved on the plane's
    (y, z) grid.
    """
    center: Tuple[float, float, float]
    size: Tuple[float, float, float]
    waveform: GaussianPulse
    mode: object
    mode_index: int = 0
    direction: str = "+x"
    amplitude: float = 1.0

    def expand(self, grid) -> List[PointDipole]:
        from .constants import EPS_0, MU_0
        if self.direction not in ("+x", "-x"):
            raise ValueError("UniModeSource supports direction '+x' or '-x'")
        s = 1.0 if self.direction == "+x" else -1.0
        c = list(self.center) + [0.0] * (3 - len(self.center))
        m, i = self.mode, self.mode_index
        ey, ez = np.asarray(m.Ey[i]), np.asarray(m.Ez[i])
        hy, hz = np.asarray(m.Hy[i]), np.asarray(m.Hz[i])
        peak = int(np.argmax(np.abs(ey) + np.abs(ez)))
        rot = np.exp(-1j * np.angle((ey.ravel() + ez.ravel())[peak]))
        ey, ez, hy, hz = (np.real(f * rot) for f in (ey, ez, hy, hz))
        eps = np.asarray(m.eps_r)
        C = float(self.amplitude)
        sheets = (("Ey", C / (EPS_0 * eps) * hz), ("Ez", -C / (EPS_0 * eps) * hy),
                  ("Hy", -s * C / MU_0 * ez), ("Hz", s * C / MU_0 * ey))
        xc, yc, zc = grid.coords
        ix = _nearest(xc, c[0])
        x = float(xc[ix] if xc.size > 1 else c[0])
        out: List[PointDipole] = []
        for jy, y_m in enumerate(np.asarray(m.y)):
            iy = _nearest(yc, c[1] + y_m)
            y = float(yc[iy] if yc.size > 1 else c[1])
            for jz, z_m in enumerate(np.asarray(m.z)):
                iz = _nearest(zc, c[2] + z_m)
                z = float(zc[iz] if zc.size > 1 else c[2])
                for comp, amp in sheets:
                    a = float(amp[jy, jz])
                    if a != 0.0:
                        out.append(PointDipole(position=(x, y, z), component=comp,
                                               waveform=self.waveform, amplitude=a))
        return out
