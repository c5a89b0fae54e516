"""Port-mode amplitudes and S-parameters from x-normal DFT planes.

Interface of the reference's ``src/photonfdtd/smatrix.py:41-118``
(``mode_amplitudes``, ``port_fields``, ``mode_port_fields``,
``s_parameters``), written from the formula:

* the x component of a cross product is ``(E x H)_x = E_y H_z - E_z H_y``;
* the reciprocity inner product of two port fields is
  ``<a, b> = dA * sum_plane [ (E_a^* x H_b)_x + (E_b x H_a^*)_x ]``;
* a mode travelling in -x has the same transverse E and the opposite
  transverse H, and the two directions are orthogonal under ``<.,.>``;
* the amplitude of mode m in direction s is ``<m_s, f> / <m_s, m_s>``.

Only ``*``, ``-``, ``.conj()``, ``.reshape`` and ``.sum`` are used, so NumPy
arrays and torch tensors both work (torch makes the amplitude differentiable
for the adjoint objective; the mode is a constant).
"""
from __future__ import annotations

from typing import Dict, NamedTuple, Tuple

import numpy as np

PORT_COMPONENTS = ("Ey", "Ez", "Hy", "Hz")   # tangential to an x-normal plane


class PortField(NamedTuple):
    """Tangential fields on an x-normal plane; the plane is the last axis."""
    ey: object
    ez: object
    hy: object
    hz: object

    @classmethod
    def from_dict(cls, d: Dict, keep_leading: bool) -> "PortField":
        """Flatten each component's plane to one trailing axis, keeping a
        leading (frequency) axis when ``keep_leading``."""
        def flat(a):
            return a.reshape(a.shape[0], -1) if keep_leading else a.reshape(-1)
        return cls(*(flat(d[c]) for c in PORT_COMPONENTS))

    def reversed(self) -> "PortField":
        """The same profile travelling the other way (transverse H negated)."""
        return PortField(self.ey, self.ez, -self.hy, -self.hz)


def _cross_x(ey, ez, hy, hz):
    return ey * hz - ez * hy


def inner(a: PortField, b: PortField, dA: float):
    """``<a, b>``, summed over the trailing plane axis."""
    left = _cross_x(a.ey.conj(), a.ez.conj(), b.hy, b.hz)
    right = _cross_x(b.ey, b.ez, a.hy.conj(), a.hz.conj())
    return (left + right).sum(-1) * dA


def mode_amplitudes(sim_fields: Dict, mode_fields: Dict, dA: float):
    """``(alpha_plus, alpha_minus)``: the projection of the simulated port
    fields (each ``(n_freq, ...)``) on the mode travelling in +x and in -x."""
    field = PortField.from_dict(sim_fields, keep_leading=True)
    mode = PortField.from_dict(mode_fields, keep_leading=False)
    out = []
    for m in (mode, mode.reversed()):
        out.append(inner(m, field, dA) / inner(m, m, dA))
    return out[0], out[1]


def port_fields(result, monitor_name: str) -> Dict:
    """The four tangential components of a port DFT monitor of ``result``."""
    dft = result.dft[monitor_name]
    absent = [c for c in PORT_COMPONENTS if c not in dft]
    if absent:
        raise ValueError(f"port monitor {monitor_name!r} is missing components {absent}; "
                         f"record components={PORT_COMPONENTS} on the port plane")
    return {c: dft[c] for c in PORT_COMPONENTS}


def mode_port_fields(mode_result, mode_index: int = 0) -> Dict:
    """Tangential components of solved mode ``mode_index`` (a ``ModeResult``)."""
    return {c: np.asarray(getattr(mode_result, c)[mode_index]) for c in PORT_COMPONENTS}


def s_parameters(result, out_monitor: str, mode_result, mode_index: int = 0,
                 dA: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
    """Forward / backward amplitudes of one solved mode at a port monitor; with
    a unit-power one-way source elsewhere, ``alpha_plus`` is the S-parameter
    into this port."""
    return mode_amplitudes(port_fields(result, out_monitor),
                           mode_port_fields(mode_result, mode_index), dA)
