"""photonfdtd_amd -- MI355X-native FDTD stepping behind photonfdtd's scene API.

``Simulation(..., backend="hip").run()`` time-steps the Yee grid with CPML,
ADE dispersion and on-device DFT accumulation in hand-written gfx950 HIP
kernels reached through the ``libpfhip.so`` C ABI (``include/pfhip.h``).
"""
from .constants import C_0, EPS_0, MU_0, ETA_0, Q_E
from .grid import Grid
from .materials import (Medium, DispersiveMedium, Pole, silica, silicon, silicon_nitride,
                        lithium_niobate, gold, silver)
from . import materials
from .geometry import Box, PolySlab
from .sources import GaussianPulse, PointDipole, ModeSource, UniModeSource
from .monitors import FieldMonitor, DFTMonitor, FluxMonitor
from .storage import CompressedFieldSeries
from .boundaries import CPMLParams, cpml_coeffs_1axis
from .simulation import Simulation, Result
from .mode import ModeResult, ModeSolver
from .smatrix import mode_amplitudes, port_fields, s_parameters
from .smoothing import smooth_permittivity

__all__ = [
    "C_0", "EPS_0", "MU_0", "ETA_0", "Q_E", "Grid", "Medium", "DispersiveMedium", "Pole",
    "materials", "silica", "silicon", "silicon_nitride", "lithium_niobate", "gold", "silver",
    "Box", "PolySlab", "GaussianPulse", "PointDipole", "ModeSource", "UniModeSource",
    "FieldMonitor", "DFTMonitor", "FluxMonitor", "CPMLParams", "cpml_coeffs_1axis",
    "Simulation", "Result", "ModeSolver", "ModeResult", "mode_amplitudes", "port_fields",
    "s_parameters", "smooth_permittivity", "CompressedFieldSeries",
]

__version__ = "0.1.0"
