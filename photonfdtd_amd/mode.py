"""Full-vectorial finite-difference waveguide mode solver
(reference ``src/photonfdtd/mode.py:128-387``: ``ModeSolver``, ``ModeResult``).

The guide is invariant along x, fields vary as ``exp(+i beta x - i w t)``,
and the cross-section (y, z) is a uniform grid with node (iy, iz) at flat
index ``iy * nz + iz`` (z fastest). The unknowns are the transverse electric
fields ``E_u = Ez`` (u = z) and ``E_v = Ey`` (v = y); ``beta^2`` is the
eigenvalue of the coupled operator of Zhu and Brown (Opt. Express 10, 853
(2002)) on the Yee-staggered difference operators:

    U_u, U_v   forward differences along z and y, the neighbour past the last
               node dropped (zero tangential E on the boundary);
    V_u, V_v   backward differences, -U_u^T and -U_v^T;

with ``eps`` the permittivity (diagonal, isotropic), ``ezi = 1/eps`` and
``k2 = k0^2``, the blocks are built here from

    A_u = k2 + U_u ezi V_u        B_u = eps + V_v U_v / k2     C_uv = U_u ezi V_v
    A_v = k2 + U_v ezi V_v        B_v = eps + V_u U_u / k2     C_vu = U_v ezi V_u

    P_uu = A_u B_u - C_uv V_u U_v / k2      P_uv = C_uv B_v - A_u V_v U_u / k2
    P_vv = A_v B_v - C_vu V_v U_u / k2      P_vu = C_vu B_u - A_v V_u U_v / k2

which is the published operator regrouped (the same matrix up to rounding).
The guided modes are the eigenvalues just below ``(k0 n_max)^2``, found by
shift-invert Arnoldi (scipy's ARPACK ``eigs``). Ex follows from Gauss's law
and H from Faraday's law with central differences (display fields; ``n_eff``
and the transverse E are the solved quantities).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .constants import C_0, MU_0


@dataclass
class ModeResult:
    """Solved modes, sorted by descending effective index. Field arrays are
    ``(num_modes, ny, nz)`` complex, each mode scaled to max |E_t| = 1;
    ``psi`` is that |E_t|; ``eps_r`` the cross-section permittivity."""
    wavelength: float
    n_eff: np.ndarray
    psi: np.ndarray
    y: np.ndarray
    z: np.ndarray
    Ey: Optional[np.ndarray] = None
    Ez: Optional[np.ndarray] = None
    Ex: Optional[np.ndarray] = None
    Hx: Optional[np.ndarray] = None
    Hy: Optional[np.ndarray] = None
    Hz: Optional[np.ndarray] = None
    fields: Optional[dict] = None
    eps_r: Optional[np.ndarray] = None


def _axis_window(coords: np.ndarray, center: float, size: float) -> np.ndarray:
    return (coords >= center - size / 2) & (coords <= center + size / 2)


class ModeSolver:
    """Guided modes of the (y, z) cross-section of ``structures`` (boxes; a
    2-tuple centre/size reads as (y, z), a 3-tuple as (x, y, z)).

    ``size = (Ly, Lz)``, ``cell_size`` a scalar or ``(dy, dz)``; at least 4
    cells per axis. ``polarization`` "te" keeps Ey-dominant modes, "tm"
    Ez-dominant ones, None all modes by descending n_eff. Dispersive media
    enter with the real part of their permittivity at ``wavelength``.
    """

    def __init__(self, size: Tuple[float, float], cell_size, structures: Sequence,
                 wavelength: float, background_eps: float = 1.0, num_modes: int = 1,
                 polarization: Optional[str] = None) -> None:
        if isinstance(cell_size, (int, float)):
            dy = dz = float(cell_size)
        else:
            dy, dz = (float(v) for v in cell_size)
        ny, nz = int(round(float(size[0]) / dy)), int(round(float(size[1]) / dz))
        if ny < 4 or nz < 4:
            raise ValueError("Need at least 4 cells per axis")
        if polarization is not None:
            polarization = polarization.lower()
            if polarization not in ("te", "tm"):
                raise ValueError("polarization must be None, 'te' or 'tm'")
        self.dy, self.dz, self.ny, self.nz = dy, dz, ny, nz
        self.y = (np.arange(ny) - (ny - 1) / 2) * dy
        self.z = (np.arange(nz) - (nz - 1) / 2) * dz
        self.wavelength = float(wavelength)
        self.num_modes = int(num_modes)
        self.polarization = polarization
        self.eps_r = np.full((ny, nz), float(background_eps))
        for st in structures:
            self._paint(st)

    def _paint(self, box) -> None:
        """Stamp one box's permittivity on the cross-section (later wins)."""
        c, s = list(box.center), list(box.size)
        (cy, cz), (sy, sz) = (c[:2], s[:2]) if len(c) == 2 else (c[1:3], s[1:3])
        rows = np.flatnonzero(_axis_window(self.y, cy, sy))
        cols = np.flatnonzero(_axis_window(self.z, cz, sz))
        if rows.size == 0 or cols.size == 0:
            return
        med = box.medium
        if getattr(med, "is_dispersive", False):
            value = float(np.real(med.eps_model(C_0 / self.wavelength)))
        else:
            value = med.eps_r
        self.eps_r[np.ix_(rows, cols)] = value

    def operator(self, k0: float) -> sp.csc_matrix:
        """The 2N x 2N matrix P acting on [E_u; E_v] = [Ez; Ey]."""
        ny, nz = self.ny, self.nz
        n = ny * nz
        fwd_z = (sp.eye(nz, k=1) - sp.eye(nz)) / self.dz
        Uu = sp.kron(sp.eye(ny), fwd_z, format="csr")
        Uv = ((sp.eye(n, k=nz) - sp.eye(n)) / self.dy).tocsr()
        Vu, Vv = (-Uu.T).tocsr(), (-Uv.T).tocsr()
        eps = self.eps_r.reshape(n).astype(float)
        E = sp.diags(eps)
        ezi = sp.diags(1.0 / eps)
        k2 = k0 * k0
        I = sp.eye(n, format="csr")
        A_u = k2 * I + Uu @ ezi @ Vu
        A_v = k2 * I + Uv @ ezi @ Vv
        B_u = E + Vv @ Uv / k2
        B_v = E + Vu @ Uu / k2
        C_uv = Uu @ ezi @ Vv
        C_vu = Uv @ ezi @ Vu
        P_uu = A_u @ B_u - C_uv @ Vu @ Uv / k2
        P_vv = A_v @ B_v - C_vu @ Vv @ Uu / k2
        P_uv = C_uv @ B_v - A_u @ Vv @ Uu / k2
        P_vu = C_vu @ B_u - A_v @ Vu @ Uv / k2
        return sp.bmat([[P_uu, P_uv], [P_vu, P_vv]], format="csc")

    def solve(self) -> ModeResult:
        ny, nz = self.ny, self.nz
        n = ny * nz
        k0 = 2 * math.pi / self.wavelength
        P = self.operator(k0)
        shift = (k0 * float(np.sqrt(self.eps_r.max()))) ** 2
        extra = 0 if self.polarization is None else 4     # candidates for the filter
        count = min(self.num_modes + extra, 2 * n - 2)
        try:
            vals, vecs = spla.eigs(P, k=count, sigma=shift, which="LM")
        except spla.ArpackNoConvergence as err:
            vals, vecs = err.eigenvalues, err.eigenvectors
        order = np.argsort(-vals.real)
        beta2, vecs = vals.real[order], vecs[:, order]
        n_eff_all = np.sqrt(np.clip(beta2, 0.0, None)) / k0
        ez_all, ey_all = vecs[:n], vecs[n:]

        picks = list(range(vecs.shape[1]))
        if self.polarization is not None:
            y_dominant = [i for i in picks
                          if np.sum(np.abs(ey_all[:, i]) ** 2) >= np.sum(np.abs(ez_all[:, i]) ** 2)]
            picks = (y_dominant if self.polarization == "te"
                     else [i for i in picks if i not in y_dominant])
        picks = picks[:self.num_modes] or list(range(min(self.num_modes, vecs.shape[1])))
        m = len(picks)
        n_eff = n_eff_all[picks]
        Ez = ez_all[:, picks].T.reshape(m, ny, nz)
        Ey = ey_all[:, picks].T.reshape(m, ny, nz)
        mag = np.sqrt(np.abs(Ey) ** 2 + np.abs(Ez) ** 2)
        psi = np.empty((m, ny, nz))
        for i in range(m):
            peak = mag[i].max()
            if peak > 0:
                Ey[i] /= peak
                Ez[i] /= peak
                psi[i] = mag[i] / peak
            else:
                psi[i] = mag[i]
        Ex, Hx, Hy, Hz = self._other_components(Ey, Ez, n_eff, k0)
        fields = {"Ex": Ex, "Ey": Ey, "Ez": Ez, "Hx": Hx, "Hy": Hy, "Hz": Hz}
        return ModeResult(wavelength=self.wavelength, n_eff=n_eff, psi=psi, y=self.y, z=self.z,
                          Ex=Ex, Ey=Ey, Ez=Ez, Hx=Hx, Hy=Hy, Hz=Hz, fields=fields,
                          eps_r=self.eps_r.copy())

    def _other_components(self, Ey, Ez, n_eff, k0):
        """Ex from div(eps E) = 0 and H = curl E / (i w mu0), d/dx -> i beta,
        central differences on the node grid (y axis 0, z axis 1)."""
        dy, dz, eps = self.dy, self.dz, self.eps_r
        iwmu = 1j * C_0 * k0 * MU_0
        Ex, Hx, Hy, Hz = (np.zeros_like(Ey) for _ in range(4))
        for i in range(Ey.shape[0]):
            beta = n_eff[i] * k0
            if beta == 0:
                continue
            div_t = np.gradient(eps * Ey[i], dy, axis=0) + np.gradient(eps * Ez[i], dz, axis=1)
            Ex[i] = -div_t / (1j * beta * eps)
            Hx[i] = (np.gradient(Ez[i], dy, axis=0) - np.gradient(Ey[i], dz, axis=1)) / iwmu
            Hy[i] = (np.gradient(Ex[i], dz, axis=1) - 1j * beta * Ez[i]) / iwmu
            Hz[i] = (1j * beta * Ey[i] - np.gradient(Ex[i], dy, axis=0)) / iwmu
        return Ex, Hx, Hy, Hz
