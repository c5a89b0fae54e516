"""Modal amplitudes from port-plane DFT fields (``src/photonfdtd/smatrix.py:41-85``).

``<a, b> = sum (Ea* x Hb + Eb x Ha*) . x_hat dA`` with tangential
``(Ey, Ez, Hy, Hz)``; the forward amplitude of mode m is ``<m+, sim>/<m+, m+>``
and the backward one uses the mode with transverse H flipped. Written with
plain array operations, so NumPy arrays and torch tensors both work (the
adjoint objective of an inverse-design loop).
"""
from __future__ import annotations

_TANGENTIAL_X = ("Ey", "Ez", "Hy", "Hz")


def _overlap(a, b, dA):
    Eay, Eaz, Hay, Haz = a
    Eby, Ebz, Hby, Hbz = b
    integrand = (Eay.conj() * Hbz - Eaz.conj() * Hby) + (Eby * Haz.conj() - Ebz * Hay.conj())
    return integrand.sum(-1) * dA


def mode_amplitudes(sim_fields, mode_fields, dA: float):
    """``(alpha_plus, alpha_minus)`` per frequency on an x-normal port plane."""
    sim = tuple(sim_fields[k].reshape(sim_fields[k].shape[0], -1) for k in _TANGENTIAL_X)
    Emy, Emz, Hmy, Hmz = (mode_fields[k].reshape(-1) for k in _TANGENTIAL_X)
    fwd = (Emy, Emz, Hmy, Hmz)
    bwd = (Emy, Emz, -Hmy, -Hmz)
    return (_overlap(fwd, sim, dA) / _overlap(fwd, fwd, dA),
            _overlap(bwd, sim, dA) / _overlap(bwd, bwd, dA))
