"""Isotropic media: plain dielectrics and pole-expanded dispersive media.

Interface and semantics follow ``src/photonfdtd/materials.py:17-268``:

* :class:`Medium` stamps a constant ``eps_r``.
* :class:`DispersiveMedium` stamps ``eps_inf`` and carries poles
  ``strength / (omega0^2 - omega^2 - 2j gamma omega)`` that the stepper
  advances with the ADE recursion (Lorentz: ``strength = d_eps w0^2``;
  Drude: ``omega0 = 0``; a lossless Sellmeier term is a Lorentz pole with
  ``w0 = 2 pi c / sqrt(C)``).
* The cited library (SiO2, Si, Si3N4, LiNbO3, Au, Ag) uses the same
  published coefficient tables (``materials.py:209-268``); a coefficient
  table is data and has to be identical for the indices to match.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import ClassVar, Sequence, Tuple

import numpy as np

_EV_TO_RAD_S = 1.519_267_447e15     # (e / hbar) * 1 eV, rad/s
_C0 = 299_792_458.0


@dataclass(frozen=True)
class Medium:
    """Non-dispersive dielectric with relative permittivity ``permittivity``."""
    permittivity: float = 1.0
    name: str = ""

    is_dispersive: ClassVar[bool] = False

    @property
    def eps_r(self) -> float:
        return float(self.permittivity)

    @classmethod
    def from_index(cls, n: float, name: str = "") -> "Medium":
        return cls(permittivity=float(n) ** 2, name=name)


VACUUM = Medium(permittivity=1.0, name="vacuum")


@dataclass(frozen=True)
class Pole:
    """One susceptibility pole; all three parameters in rad/s units."""
    strength: float
    omega0: float
    gamma: float = 0.0


@dataclass(frozen=True)
class DispersiveMedium:
    """``eps(w) = eps_inf + sum_p strength_p / (w0_p^2 - w^2 - 2j gamma_p w)``."""
    eps_inf: float
    poles: Tuple[Pole, ...] = ()
    name: str = ""

    is_dispersive: ClassVar[bool] = True

    @property
    def eps_r(self) -> float:
        return float(self.eps_inf)

    def eps_model(self, freq_hz):
        """Complex permittivity, e^{-i w t} convention (loss: Im > 0)."""
        w = 2.0 * np.pi * np.asarray(freq_hz, dtype=float)
        eps = np.full(w.shape, complex(self.eps_inf)) if w.ndim else complex(self.eps_inf)
        for p in self.poles:
            eps = eps + p.strength / (p.omega0 ** 2 - w ** 2 - 2j * p.gamma * w)
        return eps

    def index(self, freq_hz):
        return np.sqrt(self.eps_model(freq_hz))

    def at_wavelength(self, wavelength_m: float, name: str = "") -> Medium:
        """Fixed-index :class:`Medium` with this model's Re(eps) at one wavelength."""
        eps = float(np.real(self.eps_model(_C0 / float(wavelength_m))))
        return Medium(permittivity=eps,
                      name=name or f"{self.name}@{wavelength_m * 1e9:.0f}nm")

    def max_pole_omega(self) -> float:
        return max((p.omega0 for p in self.poles), default=0.0)

    @classmethod
    def lorentz(cls, eps_inf: float, resonances: Sequence[Tuple[float, float, float]],
                name: str = "") -> "DispersiveMedium":
        """Resonances ``(delta_eps, f0_Hz, linewidth_Hz)``; ``gamma = pi * linewidth``."""
        poles = []
        for d_eps, f0, df in resonances:
            w0 = 2.0 * math.pi * f0
            poles.append(Pole(strength=d_eps * w0 ** 2, omega0=w0, gamma=math.pi * df))
        return cls(eps_inf=float(eps_inf), poles=tuple(poles), name=name)

    @classmethod
    def drude(cls, eps_inf: float, f_plasma_Hz: float, f_collision_Hz: float,
              name: str = "") -> "DispersiveMedium":
        wp = 2.0 * math.pi * f_plasma_Hz
        return cls(eps_inf=float(eps_inf),
                   poles=(Pole(strength=wp ** 2, omega0=0.0,
                               gamma=math.pi * f_collision_Hz),),
                   name=name)

    @classmethod
    def sellmeier(cls, terms: Sequence[Tuple[float, float]], name: str = "") -> "DispersiveMedium":
        """Terms ``(B, C[um^2])`` of ``n^2 - 1 = sum B lam^2/(lam^2 - C)``; eps_inf = 1."""
        poles = []
        for B, C_um2 in terms:
            w0 = 2.0 * math.pi * _C0 / (math.sqrt(C_um2) * 1e-6)
            poles.append(Pole(strength=B * w0 ** 2, omega0=w0, gamma=0.0))
        return cls(eps_inf=1.0, poles=tuple(poles), name=name)


def _lorentz_drude(eps_inf: float, wp_eV: float,
                   rows: Sequence[Tuple[float, float, float]], name: str) -> DispersiveMedium:
    """Rakic (1998) rows ``(f_j, Gamma_j[eV], omega_j[eV])``; row 0 is Drude."""
    wp = wp_eV * _EV_TO_RAD_S
    poles = tuple(Pole(strength=f * wp ** 2, omega0=w * _EV_TO_RAD_S,
                       gamma=0.5 * G * _EV_TO_RAD_S) for f, G, w in rows)
    return DispersiveMedium(eps_inf=float(eps_inf), poles=poles, name=name)


def silica() -> DispersiveMedium:
    """Fused silica, Malitson (1965) Sellmeier."""
    return DispersiveMedium.sellmeier(
        [(0.6961663, 0.0684043 ** 2), (0.4079426, 0.1162414 ** 2),
         (0.8974794, 9.896161 ** 2)], name="SiO2 (Malitson 1965)")


def silicon() -> DispersiveMedium:
    """Crystalline silicon, Salzberg & Villa (1957) Sellmeier."""
    return DispersiveMedium.sellmeier(
        [(10.6684293, 0.301516485 ** 2), (0.0030434748, 1.13475115 ** 2),
         (1.54133408, 1104.0 ** 2)], name="Si (Salzberg-Villa 1957)")


def silicon_nitride() -> DispersiveMedium:
    """LPCVD Si3N4, Luke et al. (2015) Sellmeier."""
    return DispersiveMedium.sellmeier(
        [(3.0249, 0.1353406 ** 2), (40314.0, 1239.842 ** 2)],
        name="Si3N4 (Luke 2015)")


def lithium_niobate(axis: str = "o") -> DispersiveMedium:
    """Congruent LiNbO3, Zelmon et al. (1997); ``axis`` 'o' or 'e'."""
    tables = {
        "o": [(2.6734, 0.01764), (1.2290, 0.05914), (12.614, 474.60)],
        "e": [(2.9804, 0.02047), (0.5981, 0.0666), (8.9543, 416.08)],
    }
    if axis not in tables:
        raise ValueError("axis must be 'o' (ordinary) or 'e' (extraordinary)")
    return DispersiveMedium.sellmeier(tables[axis], name=f"LiNbO3-{axis} (Zelmon 1997)")


def gold() -> DispersiveMedium:
    """Gold, Rakic et al. (1998) Lorentz-Drude."""
    return _lorentz_drude(1.0, 9.03, [
        (0.760, 0.053, 0.000), (0.024, 0.241, 0.415), (0.010, 0.345, 0.830),
        (0.071, 0.870, 2.969), (0.601, 2.494, 4.304), (4.384, 2.214, 13.32),
    ], name="Au (Rakic 1998 LD)")


def silver() -> DispersiveMedium:
    """Silver, Rakic et al. (1998) Lorentz-Drude."""
    return _lorentz_drude(1.0, 9.01, [
        (0.845, 0.048, 0.000), (0.065, 3.886, 0.816), (0.124, 0.452, 4.481),
        (0.011, 0.065, 8.185), (0.840, 0.916, 9.083), (5.646, 2.419, 20.29),
    ], name="Ag (Rakic 1998 LD)")
