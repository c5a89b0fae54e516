"""Batched 2-D stepping: many independent xy-plane simulations in one launch.

The reference runs one simulation per ``Simulation.run()``
(``simulation.py:896-904``); a batch of design candidates is a Python loop of
runs. ``SimBatch`` holds n simulations of one precision on the device and
advances them together: one march kernel per time step (launched per step,
replayed from a captured hipGraph, or looped inside one persistent launch)
walks a list of wavefront chunks (a 62-column strip over a run of rows) that
spans every grid, so uneven sizes share the launch without padding. Each simulation keeps the
single-run semantics of ``backend="hip"``: the same tables
(``hip_backend.build_desc``) and the same per-node arithmetic, so a batched
simulation reproduces its single run bit for bit.

Only the polarisation group that the sources excite is stepped (TE: Ex, Ey,
Hz; TM: Ez, Hx, Hy); the other group stays exactly zero in 2-D, as it does in
the reference's six-component update. Supported per simulation: CPML, soft
sources, dispersive (ADE) media and DFT monitors. FieldMonitor and
FluxMonitor need per-simulation host stops and are rejected.

A batch of 3-D grids (every nz >= 2) runs the single-run march kernels each
step over the tiles of all its simulations (each tile names its
simulation), then one ADE, one source and one DFT-capture launch for all of
them; only the per-step stream mode exists for 3-D. The same
bit-for-bit parity with single runs holds. 3-D batches need the march
kernel's conditions: E-side sources only and CPML on slabs.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import _native as N
from .hip_backend import build_desc, ce_of, exact_mode, validate
from .monitors import DFTMonitor, FieldMonitor, FluxMonitor
from .simulation import Result

TE_COMPONENTS = ("Ex", "Ey", "Hz")
TM_COMPONENTS = ("Ez", "Hx", "Hy")
MODES = {"stream": 0, "persistent": 1, "graph": 2}


def polarisation(sim) -> int:
    """0 for TE (sources on Ex/Ey/Hz), 1 for TM (Ez/Hx/Hy)."""
    comps = {s.component for s in sim.sources}
    if not comps:
        raise ValueError("a batched simulation needs at least one source")
    if comps <= set(TE_COMPONENTS):
        return 0
    if comps <= set(TM_COMPONENTS):
        return 1
    raise ValueError(f"sources mix the TE and TM groups: {sorted(comps)}")


def _check(sim) -> np.dtype:
    npdt = validate(sim)
    nx, ny, nz = sim.grid.shape
    if nx < 2 or ny < 2:
        raise ValueError(f"batched stepping needs nx, ny >= 2, got shape {sim.grid.shape}")
    for m in sim.monitors:
        if isinstance(m, (FieldMonitor, FluxMonitor)):
            raise ValueError(f"monitor {m.name!r}: a batch supports DFTMonitor only")
    if nz == 1:
        polarisation(sim)
    elif any(s.component.startswith("H") for s in sim.sources):
        raise ValueError("a 3-D batch needs E-side sources only (the march kernel's condition)")
    return npdt


class SimBatch:
    """n simulations of one precision advanced together on the device: all
    2-D (xy-plane) or all 3-D. For 3-D, ``waves`` is the wavefronts per march
    workgroup (10, 12 or 16) and ``rows`` the x planes per tile (1-128);
    ``None`` picks the defaults of the grid kind."""

    def __init__(self, sims: Sequence, mode: Optional[str] = None, waves: Optional[int] = None,
                 rows: Optional[int] = None):
        if not sims:
            raise ValueError("empty batch")
        dts = {_check(s) for s in sims}
        if len(dts) != 1:
            raise ValueError("a batch needs one compute precision")
        kinds = {s.grid.shape[2] > 1 for s in sims}
        if len(kinds) != 1:
            raise ValueError("a batch holds 2-D grids only or 3-D grids only")
        self.is3d = kinds.pop()
        d_mode, d_waves, d_rows = ("stream", 10, 128) if self.is3d else ("graph", 4, 64)
        mode = d_mode if mode is None else mode
        waves = d_waves if waves is None else waves
        rows = d_rows if rows is None else rows
        self.npdt = dts.pop()
        self.sims = list(sims)
        self.lib = N.load_library()
        n = len(sims)
        descs = (N.PfDesc * n)()
        pols = (C.c_int32 * n)()
        keep: List = []
        self.slots: List[Dict[str, tuple]] = []
        built: Dict[int, tuple] = {}        # the same Simulation object shares its tables
        for i, sim in enumerate(sims):
            if id(sim) not in built:
                d, k, t, slots, _ = build_desc(sim, self.npdt, exact_mode(self.npdt),
                                               0 if self.is3d else N.PF_FLAG_GENERIC)
                built[id(sim)] = (d, slots, 0 if self.is3d else polarisation(sim))
                keep += [d, k, t]
            d, slots, pol = built[id(sim)]
            descs[i] = d
            pols[i] = pol
            self.slots.append(slots)
        h = C.c_void_p()
        N.check(self.lib, self.lib.pf_batch_create(descs, pols, n, C.byref(h)), "pf_batch_create")
        self.h = h
        self.step = 0
        self.n_steps = max(s.n_steps for s in sims)
        del keep
        self.configure(mode, waves, rows)

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.pf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, mode: str, waves: int = 4, rows: int = 64) -> None:
        """Launch mode, wavefronts per workgroup (1, 2, 4, 8) and owned rows
        per wavefront chunk; for 3-D: "stream", march wavefronts per
        workgroup (10, 12, 16) and x planes per tile; 3-D runs in "stream" or
        "graph" mode (pairs of steps replayed from a captured hipGraph)."""
        if mode not in MODES:
            raise ValueError(f"mode must be one of {sorted(MODES)}")
        N.check(self.lib, self.lib.pf_batch_configure(self.h, MODES[mode], int(waves), int(rows)),
                "pf_batch_configure")
        self.mode, self.waves, self.rows = mode, int(waves), int(rows)

    def reset(self) -> None:
        """Zero every field, psi, ADE and DFT state; back to step 0."""
        N.check(self.lib, self.lib.pf_batch_reset(self.h), "pf_batch_reset")
        self.step = 0

    def persistent_grid(self, waves: int = 4) -> Dict[str, int]:
        """Occupancy and workgroup count of the persistent mode (no launch)."""
        per_cu, n_cu, grid = C.c_int32(), C.c_int32(), C.c_int32()
        N.check(self.lib, self.lib.pf_batch_persistent_grid(self.h, int(waves), C.byref(per_cu),
                                                            C.byref(n_cu), C.byref(grid)),
                "pf_batch_persistent_grid")
        return dict(per_cu=per_cu.value, n_cu=n_cu.value, grid=grid.value)

    def tune(self, trial_steps: int = 100,
             modes: Optional[Sequence[str]] = None,
             waves: Optional[Sequence[int]] = None, rows: Optional[Sequence[int]] = None,
             log: Optional[Callable[[str], None]] = None) -> Dict:
        """Autotuner: time every (mode, wavefronts per workgroup, rows per
        chunk) candidate over ``trial_steps`` steps from a reset state with
        HIP events, keep the fastest, and reset the batch to step 0."""
        if self.is3d:
            d_modes, d_waves, d_rows = ("stream", "graph"), (10, 12, 16), (32, 64, 128)
        else:
            d_modes, d_waves, d_rows = ("graph", "stream", "persistent"), (2, 4, 8), (16, 32, 64, 128)
        modes = d_modes if modes is None else modes
        waves = d_waves if waves is None else waves
        rows = d_rows if rows is None else rows
        trial_steps = max(1, min(int(trial_steps), self.n_steps))
        trials = []
        for m in modes:
            for w in waves:
                for r in rows:
                    self.configure(m, w, r)
                    self.reset()
                    self.advance(trial_steps)
                    ms = self.last_ms() / trial_steps
                    trials.append(dict(mode=m, waves=w, rows=r, ms_per_step=ms))
                    if log is not None:
                        log(f"[tune] {m} waves {w} rows {r}: {ms:.4f} ms/step")
        best = min(trials, key=lambda t: t["ms_per_step"])
        self.configure(best["mode"], best["waves"], best["rows"])
        self.reset()
        return dict(best, trials=trials)

    def advance(self, n: Optional[int] = None) -> None:
        n = self.n_steps - self.step if n is None else int(n)
        N.check(self.lib, self.lib.pf_batch_advance(self.h, n), "pf_batch_advance")
        self.step += n

    def set_eps(self, i: int, eps: np.ndarray) -> None:
        """New permittivity grid for simulation i (takes effect from the next
        step; call ``reset`` to start a fresh run)."""
        sim = self.sims[i]
        ce = ce_of(sim, np.asarray(eps).reshape(sim.grid.shape), self.npdt)
        N.check(self.lib, self.lib.pf_batch_write_ce(self.h, int(i), ce.ctypes.data),
                "pf_batch_write_ce")

    def last_ms(self) -> float:
        """Device time of the last ``advance`` (HIP events on the batch stream)."""
        ms = C.c_double()
        N.check(self.lib, self.lib.pf_batch_last_ms(self.h, C.byref(ms)), "pf_batch_last_ms")
        return ms.value

    def field(self, i: int, comp: str) -> np.ndarray:
        out = np.empty(self.sims[i].grid.shape, dtype=self.npdt)
        N.check(self.lib, self.lib.pf_batch_read_field(self.h, int(i), N.COMP_CODE[comp],
                                                       out.ctypes.data), "pf_batch_read_field")
        return out

    def dft(self, i: int, name: str) -> Dict[str, np.ndarray]:
        mon, comps, count, n_freq = self.slots[i][name]
        out = {}
        for slot, c in enumerate(comps):
            buf = np.empty((n_freq,) + count, dtype=np.complex128)
            N.check(self.lib, self.lib.pf_batch_dft_read(
                self.h, int(i), mon, slot, buf.view(np.float64).ctypes.data_as(
                    C.POINTER(C.c_double))), "pf_batch_dft_read")
            out[c] = buf
        return out

    def results(self) -> List[Result]:
        out = []
        for i, sim in enumerate(self.sims):
            r = Result(times=np.arange(sim.n_steps) * sim.dt)
            for m in sim.monitors:
                if isinstance(m, DFTMonitor):
                    r.dft_freqs[m.name] = np.asarray(m.freqs, dtype=np.float64)
                    r.dft[m.name] = self.dft(i, m.name)
            out.append(r)
        return out


def run_batch(sims: Sequence, mode: Optional[str] = None, tune: bool = False,
              trial_steps: int = 100) -> List[Result]:
    """Run every simulation to its own step count; one ``Result`` each."""
    b = SimBatch(sims, mode=mode)
    try:
        if tune:
            b.tune(trial_steps)
        b.advance()
        return b.results()
    finally:
        b.close()


def dft_flux(fields: Dict[str, np.ndarray], axis: str, dA: float) -> np.ndarray:
    """Spectral flux Re(E x H*) . n summed over a DFT plane, per frequency.

    ``fields`` maps component names to (n_freq, ...) complex arrays of one
    DFTMonitor; missing components count as zero. The reference has no
    spectral flux (its FluxMonitor is time-domain, simulation.py:1440-1459);
    the Poynting components follow the same S . n of that function.
    """
    z = None
    pairs = {"x": (("Ey", "Hz", 1), ("Ez", "Hy", -1)),
             "y": (("Ez", "Hx", 1), ("Ex", "Hz", -1)),
             "z": (("Ex", "Hy", 1), ("Ey", "Hx", -1))}[axis]
    for e, h, sign in pairs:
        if e in fields and h in fields:
            term = sign * np.real(fields[e] * np.conj(fields[h]))
            z = term if z is None else z + term
    if z is None:
        raise ValueError(f"no Poynting pair for axis {axis!r} in {sorted(fields)}")
    return z.reshape(z.shape[0], -1).sum(axis=1) * dA
