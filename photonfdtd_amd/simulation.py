"""``Simulation`` scene assembly and the ``backend="hip"`` dispatch.

Mirrors the reference's setup path (``src/photonfdtd/simulation.py:361-808``):
per-array precision map, dispersion detection, source expansion, permittivity
stamping, Courant time step, CPML tables and ADE pole tables. ``run()``
hands the assembled scene to the HIP stepper (``hip_backend.run_hip``), which
plays the role the reference's ``rust-cuda`` branch plays
(``simulation.py:896-904`` -> ``rustbackend.py:457-461``).

Only the GPU stepper ships here: the NumPy/JAX/Rust steppers of the
reference are not part of this package, so ``backend`` accepts ``"hip"``
(and ``"auto"``, which means ``"hip"``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from .boundaries import CPMLParams, cpml_coeffs_1axis
from .constants import C_0
from .grid import Grid
from .sources import PointDipole

FIELD_KEYS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
_E_KEYS = FIELD_KEYS[:3]
ARRAY_KEYS = FIELD_KEYS + ("eps_r", "psi", "monitors")
COMPUTE_KEYS = FIELD_KEYS + ("eps_r", "psi")
_GROUPS = ("default", "compute", "fields", "e_fields", "h_fields")
_DTYPES = {"float32": np.float32, "float64": np.float64}
BACKENDS = ("auto", "hip")


def resolve_precision(precision) -> Dict[str, type]:
    """``{array_key: numpy dtype}`` from a dtype string or a per-array dict.

    Precedence (``simulation.py:279-324``): individual key, then the
    ``e_fields``/``h_fields`` group, then ``fields``, then ``compute``, then
    ``default`` (itself defaulting to float64). ``monitors`` only listens to
    its own key and ``default``.
    """
    if isinstance(precision, str):
        if precision not in _DTYPES:
            raise ValueError("precision must be 'float32' or 'float64'")
        return {k: _DTYPES[precision] for k in ARRAY_KEYS}
    if not isinstance(precision, dict):
        raise TypeError("precision must be a str ('float32'/'float64') or a dict "
                        "mapping array keys to dtype strings")
    unknown = set(precision) - set(ARRAY_KEYS) - set(_GROUPS)
    if unknown:
        raise ValueError(f"unknown precision key(s) {sorted(unknown)}; allowed keys: "
                         f"{sorted(set(ARRAY_KEYS) | set(_GROUPS))}")
    for k, v in precision.items():
        if v not in _DTYPES:
            raise ValueError(f"precision[{k!r}] must be 'float32' or 'float64', got {v!r}")
    default = _DTYPES[precision.get("default", "float64")]

    def first(*keys):
        for k in keys:
            if k in precision:
                return _DTYPES[precision[k]]
        return default

    out = {}
    for k in FIELD_KEYS:
        out[k] = first(k, "e_fields" if k in _E_KEYS else "h_fields", "fields", "compute")
    out["eps_r"] = first("eps_r", "compute")
    out["psi"] = first("psi", "compute")
    out["monitors"] = first("monitors")
    return out


@dataclass
class Result:
    """Run output, same fields as ``simulation.py:327-336``."""
    times: np.ndarray
    fields: Dict[str, Dict[str, np.ndarray]] = field(default_factory=dict)
    monitor_times: Dict[str, np.ndarray] = field(default_factory=dict)
    flux: Dict[str, float] = field(default_factory=dict)
    dft: Dict[str, Dict[str, np.ndarray]] = field(default_factory=dict)
    dft_freqs: Dict[str, np.ndarray] = field(default_factory=dict)


class Simulation:
    def __init__(self, grid: Grid, structures: Sequence = (), sources: Sequence = (),
                 monitors: Sequence = (), run_time: float = 1e-12, courant: float = 0.99,
                 cpml_params: CPMLParams = CPMLParams(), verbose: bool = False,
                 use_gpu: bool = False, use_numba: bool = False, use_jax: bool = False,
                 precision="float64", subpixel: bool = False, subpixel_factor: int = 3,
                 backend: str = "auto") -> None:
        if backend not in BACKENDS:
            raise ValueError(
                f"backend must be one of {BACKENDS}; this package ships the HIP "
                "stepper only (the NumPy/JAX/Rust steppers are not part of it)")
        if use_gpu or use_numba or use_jax:
            raise ValueError("use_gpu / use_numba / use_jax select steppers that this "
                             "package does not ship; use backend='hip'")
        self.backend = backend
        self.precision = precision
        self.dtypes = resolve_precision(precision)
        self.dtype = self.dtypes["Ex"]
        self.subpixel = bool(subpixel)
        self.subpixel_factor = int(subpixel_factor)
        self.grid = grid
        self.structures = list(structures)
        self._has_dispersion = any(getattr(getattr(s, "medium", None), "is_dispersive", False)
                                   for s in self.structures)
        if self._has_dispersion and self.subpixel:
            raise NotImplementedError("Subpixel smoothing combined with dispersive media "
                                      "is not yet supported. Use one or the other.")
        self.particle_sources: List = []
        self.sources: List[PointDipole] = []
        for s in sources:
            self.add_source(s)
        self.monitors = list(monitors)
        self.run_time = float(run_time)
        self.courant = float(courant)
        self.cpml_params = cpml_params
        self.verbose = verbose
        self.progress_callback: Optional[Callable[[int, int], None]] = None

        self._eps_r: Optional[np.ndarray] = None
        self._eps_r_custom = False
        self._materialize_eps_r()

        inv_d2 = 0.0
        for d in grid.cell_size:
            if d > 0:
                inv_d2 += 1.0 / d ** 2
        self.dt = self.courant / (C_0 * math.sqrt(inv_d2))
        self.n_steps = int(math.ceil(self.run_time / self.dt))
        self._build_cpml()

    # -- scene assembly ------------------------------------------------- #
    def add_source(self, source) -> None:
        if isinstance(source, PointDipole):
            self.sources.append(source)
        elif hasattr(source, "expand"):
            self.sources.extend(source.expand(self.grid))
        elif type(source).__name__ == "ChargedParticle":
            self.particle_sources.append(source)
        else:
            raise TypeError(f"Unsupported source type: {type(source).__name__}. "
                            "Pass a PointDipole or ModeSource.")

    def add_monitor(self, monitor) -> None:
        self.monitors.append(monitor)

    def _materialize_eps_r(self) -> np.ndarray:
        eps = np.ones(self.grid.shape, dtype=self.dtypes["eps_r"])
        for s in self.structures:
            s.stamp(self.grid, eps)
        self._eps_r = eps
        return eps

    @property
    def eps_r(self) -> np.ndarray:
        if self._eps_r is None:
            self._materialize_eps_r()
        return self._eps_r

    @eps_r.setter
    def eps_r(self, value) -> None:
        self._eps_r = value
        self._eps_r_custom = True

    def _smoothed_eps_components(self):
        """``(eps_xx, eps_yy, eps_zz)`` of subpixel smoothing; a custom scalar
        ``eps_r`` has no geometry to supersample and is used for all three
        (``simulation.py:546-559``)."""
        from .smoothing import smooth_permittivity
        if self._eps_r_custom:
            eps = self.eps_r
            return (eps, eps, eps)
        return smooth_permittivity(self.grid, self.structures, background_eps=1.0,
                                   factor=self.subpixel_factor, dtype=self.dtypes["eps_r"])

    def _build_cpml(self) -> None:
        self._b_e, self._c_e, self._b_h, self._c_h = [], [], [], []
        for ax in range(3):
            n = self.grid.shape[ax]
            d = self.grid.cell_size[ax] if n > 1 else 1.0
            npml = self.grid.pml_layers[ax] if n > 1 else 0
            be, ce, bh, ch = cpml_coeffs_1axis(n, d, npml, self.dt, self.cpml_params)
            self._b_e.append(be)
            self._c_e.append(ce)
            self._b_h.append(bh)
            self._c_h.append(ch)

    def ade_media(self):
        """Dispersive regions after overlap resolution, in first-use order
        (the HIP driver's ``ade_tables``, ``simulation.py:572-636``)."""
        from .hip_backend import ade_tables
        return ade_tables(self)

    # -- dispatch ------------------------------------------------------- #
    def run(self, out_of_core: bool = False, tile_cells: Optional[int] = None,
            ooc_workdir: Optional[str] = None) -> Result:
        if out_of_core:
            raise NotImplementedError("out_of_core stepping is not provided; the HIP "
                                      "stepper keeps the whole domain in HBM")
        from .hip_backend import run_hip
        return run_hip(self)
