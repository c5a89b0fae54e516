"""``backend="hip"``: the Python driver of the gfx950 stepper.

Same shape as the reference's Rust/CUDA driver (``src/photonfdtd/rustbackend.py``):

* ``validate``  -- feature gate with ``NotImplementedError`` (``:85-120``);
* ``build_tables`` -- ce, CPML tables, soft-source tables (``:135-166``), plus
  the ADE, DFT-phase and flux tables the reference keeps on the host;
* ``HipEngine`` -- ``advance(step0, n)`` / ``get(comp)`` with a cache cleared
  on every advance (``:222-235,297-304``);
* ``run_loop`` -- chunked stepping with progress ticks and FieldMonitor
  recording (``:310-447``). Unlike the reference, DFT and flux accumulate on
  the device every step, so the loop stops only at FieldMonitor steps.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from dataclasses import dataclass
from typing import Dict, List

import numpy as np

from . import _native as N
from .constants import EPS_0, MU_0
from .monitors import DFTMonitor, FieldMonitor, FluxMonitor
from .simulation import Result
from .storage import CompressedFieldSeries, get_codec

# Every name this module takes from its package (constants, monitors,
# Simulation.Result) exists under the same name in the reference package, and
# the only Simulation / Grid attributes it reads are the reference's own
# (grid, structures, sources, monitors, dt, n_steps, dtypes, eps_r,
# _b_e/_c_e/_b_h/_c_h, subpixel, particle_sources, progress_callback,
# verbose; grid.shape/cell_size/coords/index_at), so the file can sit in the
# reference package as its HIP driver beside _native.py.

_AXIS = {"x": 0, "y": 1, "z": 2}
# The stepping arrays whose precision must agree (simulation.py:274).
_COMPUTE_KEYS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "eps_r", "psi")


def nearest_node(grid, axis: int, position: float) -> int:
    """Index of the node nearest ``position`` along ``axis``; the first wins
    a tie (the argmin of ``rustbackend.py:350-352``, ``simulation.py:1013``)."""
    return int(np.argmin(np.abs(np.asarray(grid.coords[axis]) - position)))


@dataclass
class AdeMedium:
    """One dispersive material's cells and pole recursion coefficients.

    ``P^{n+1} = a P^n + b P^{n-1} + c E^n`` with ``a = (2 - w0^2 dt^2)/(1+g dt)``,
    ``b = (g dt - 1)/(1+g dt)``, ``c = eps0 S dt^2/(1+g dt)``; E is corrected by
    ``-inv_eps * sum_p (P^{n+1} - P^n)``, ``inv_eps = 1/(eps0 eps_inf)``
    (``simulation.py:572-669``).
    """
    cells: np.ndarray          # flat C-order indices, ascending
    abc: np.ndarray            # (n_pole, 3) float64
    inv_eps: float


def ade_tables(sim) -> List[AdeMedium]:
    """Dispersive regions after overlap resolution, in first-use order, from
    ``sim.structures`` (the table half of ``_build_ade``, ``simulation.py:572-636``).

    Structures are stamped in order; a dispersive medium is keyed by object
    identity and a later non-dispersive structure releases the cells it
    covers. Raises the reference's ``ValueError`` when a pole has
    ``omega0 dt >= 2`` (``simulation.py:593-609``).
    """
    grid, dt = sim.grid, sim.dt
    owner = np.full(grid.shape, -1, dtype=np.int32)
    media: List = []
    ids: Dict[int, int] = {}
    for st in sim.structures:
        med = getattr(st, "medium", None)
        mask = st.region_mask(grid)
        if getattr(med, "is_dispersive", False):
            if id(med) not in ids:
                ids[id(med)] = len(media)
                media.append(med)
            owner[mask] = ids[id(med)]
        else:
            owner[mask] = -1
    if not media:
        return []
    worst = max(m.max_pole_omega() * dt for m in media)
    if worst >= 2.0:
        raise ValueError(
            f"A dispersive medium has a pole with omega0*dt = {worst:.2f} >= 2, which makes "
            "the explicit ADE update unstable at this resolution. Use a finer grid, drop "
            "poles outside the band, or DispersiveMedium.at_wavelength() for a fixed-index "
            "Medium.")
    flat = owner.ravel()
    out = []
    for mid, med in enumerate(media):
        cells = np.flatnonzero(flat == mid)
        if cells.size == 0:
            continue
        rows = []
        for p in med.poles:
            den = 1.0 + p.gamma * dt
            rows.append(((2.0 - p.omega0 ** 2 * dt ** 2) / den, (p.gamma * dt - 1.0) / den,
                         (EPS_0 * p.strength * dt ** 2) / den))
        out.append(AdeMedium(cells=cells.astype(np.int64),
                             abc=np.asarray(rows, dtype=np.float64).reshape(-1, 3),
                             inv_eps=1.0 / (EPS_0 * med.eps_inf)))
    return out


def validate(sim) -> np.dtype:
    """Feature gate; returns the single compute dtype."""
    dts = {np.dtype(sim.dtypes[k]) for k in _COMPUTE_KEYS}
    if dts not in ({np.dtype(np.float64)}, {np.dtype(np.float32)}):
        raise NotImplementedError("backend='hip' needs one uniform compute precision "
                                  "('float32' or 'float64'; monitor storage may differ).")
    if sim.particle_sources:
        raise NotImplementedError("ChargedParticle sources are not supported on backend='hip'.")
    for m in sim.monitors:
        if isinstance(m, FieldMonitor) and m.compression is not None:
            get_codec(m.compression)       # ImportError / ValueError before any device work
    for s in sim.sources:
        if s.component not in N.COMP_CODE:
            raise ValueError(f"unknown source component {s.component!r}")
    return dts.pop()


def exact_mode(dtype) -> bool:
    """Reference operation order: on for float64, off for float32 unless
    ``PHOTONFDTD_HIP_EXACT`` says otherwise."""
    env = os.environ.get("PHOTONFDTD_HIP_EXACT")
    if env is not None:
        return env == "1"
    return np.dtype(dtype) == np.float64


def _snap_box(shape, ds, plane=None, zplane=None):
    """Box of ``field[::ds, ::ds, ::ds]`` with an optional single plane."""
    start = [0, 0, 0]
    count = [(n + ds - 1) // ds for n in shape]
    stride = [ds, ds, ds]
    if zplane is not None:                 # FieldMonitor plane_z keeps x/y stride
        start[2], count[2], stride[2] = zplane, 1, 1
    if plane is not None:
        ax, ci = plane
        start[ax], count[ax], stride[ax] = ci, 1, 1
    return start, count, stride


def ce_of(sim, eps, npdt) -> np.ndarray:
    """ce = dt / (eps_r eps0) in the field dtype (``simulation.py:953-954``)."""
    eps = np.asarray(eps, dtype=sim.dtypes["eps_r"])
    return np.ascontiguousarray((sim.dt / (eps * EPS_0)).astype(npdt, copy=False))


def build_tables(sim, npdt) -> Dict:
    grid = sim.grid
    dt = sim.dt
    n_steps = sim.n_steps
    nx, ny, nz = grid.shape
    t: Dict = {}
    if sim.subpixel:
        # one coefficient per E component from the smoothed diagonal tensor
        # (simulation.py:956-969); the generic kernels run such a handle
        exx, eyy, ezz = sim._smoothed_eps_components()
        t["ce"] = ce_of(sim, exx, npdt)
        t["ce_y"] = ce_of(sim, eyy, npdt)
        t["ce_z"] = ce_of(sim, ezz, npdt)
    else:
        t["ce"] = ce_of(sim, sim.eps_r, npdt)
    t["cell"] = tuple(d if d > 0 else 1.0 for d in grid.cell_size)
    t["cpml"] = [np.ascontiguousarray(a, dtype=np.float64)
                 for arrs in (sim._b_e, sim._c_e, sim._b_h, sim._c_h) for a in arrs]

    # Soft sources: waveform rows keyed by (waveform object, E/H Yee time).
    steps = np.arange(n_steps)
    rows, row_of = [], {}
    entries = []
    for order, src in enumerate(sim.sources):
        i, j, k = grid.index_at(src.position)
        comp = N.COMP_CODE[src.component]
        is_h = comp >= 3
        key = (id(src.waveform), is_h)
        if key not in row_of:
            row_of[key] = len(rows)
            tt = (steps + (0.5 if is_h else 1.0)) * dt
            rows.append(np.asarray(src.waveform(tt), dtype=np.float64))
        entries.append((comp, (i * ny + j) * nz + k, order, float(src.amplitude), row_of[key]))
    entries.sort(key=lambda e: (e[0], e[1], e[2]))
    t["src_comp"] = np.array([e[0] for e in entries], dtype=np.int32)
    t["src_cell"] = np.array([e[1] for e in entries], dtype=np.int64)
    t["src_amp"] = np.array([e[3] for e in entries], dtype=np.float64)
    t["src_wave"] = np.array([e[4] for e in entries], dtype=np.int32)
    t["wave"] = (np.ascontiguousarray(np.stack(rows)) if rows
                 else np.zeros((1, max(n_steps, 1)), dtype=np.float64))

    t["ade"] = ade_tables(sim)

    dfts = []
    for m in sim.monitors:
        if not isinstance(m, DFTMonitor):
            continue
        pl = m.plane()
        plane = None if pl is None else (pl[0], nearest_node(grid, pl[0], pl[1]))
        omega = 2.0 * np.pi * np.asarray(m.freqs, dtype=np.float64)
        sample_steps = range(0, n_steps, m.interval)
        ph_e = np.empty((len(sample_steps), omega.size), dtype=np.complex128)
        ph_h = np.empty_like(ph_e)
        for q, step in enumerate(sample_steps):
            ph_e[q] = np.exp(1j * omega * ((step + 1.0) * dt))
            ph_h[q] = np.exp(1j * omega * ((step + 0.5) * dt))
        dfts.append(dict(mon=m, box=_snap_box(grid.shape, m.downsample, plane=plane),
                         n_samples=len(sample_steps), phase_e=ph_e, phase_h=ph_h))
    t["dft"] = dfts

    fluxes = []
    for m in sim.monitors:
        if not isinstance(m, FluxMonitor):
            continue
        ax = _AXIS[m.plane_axis]
        idx = 0 if grid.shape[ax] == 1 else nearest_node(grid, ax, m.plane_position)
        ext = [grid.cell_size[a] if grid.shape[a] > 1 else 1.0 for a in range(3)]
        dA = float(np.prod([ext[a] for a in range(3) if a != ax]))
        fluxes.append(dict(mon=m, axis=ax, index=idx, dA=dA))
    t["flux"] = fluxes
    return t


def build_desc(sim, npdt, exact: bool, flags: int = 0):
    """Fill a ``pf_desc`` for ``sim``. Returns (desc, host arrays that must
    stay alive until pf_create returns, tables, DFT slots, flux names)."""
    npdt = np.dtype(npdt).type
    t = build_tables(sim, npdt)
    keep: List = []                       # host arrays alive until pf_create returns
    d = N.PfDesc()
    d.dtype = N.PF_F64 if npdt == np.float64 else N.PF_F32
    d.flags = int(flags)
    d.exact = 1 if exact else 0
    d.shape[:] = sim.grid.shape
    d.cell[:] = t["cell"]
    d.ch = sim.dt / MU_0
    d.ce = t["ce"].ctypes.data
    if "ce_y" in t:
        d.ce_y = t["ce_y"].ctypes.data
        d.ce_z = t["ce_z"].ctypes.data
    cp = t["cpml"]
    for a in range(3):
        d.b_e[a] = N.as_ptr(cp[a], C.c_double)
        d.c_e[a] = N.as_ptr(cp[3 + a], C.c_double)
        d.b_h[a] = N.as_ptr(cp[6 + a], C.c_double)
        d.c_h[a] = N.as_ptr(cp[9 + a], C.c_double)
    d.n_steps = t["wave"].shape[1]
    d.n_src = t["src_comp"].size
    d.src_comp = N.as_ptr(t["src_comp"], C.c_int32)
    d.src_cell = N.as_ptr(t["src_cell"], C.c_int64)
    d.src_amp = N.as_ptr(t["src_amp"], C.c_double)
    d.src_wave = N.as_ptr(t["src_wave"], C.c_int32)
    d.n_wave = t["wave"].shape[0]
    d.wave = N.as_ptr(t["wave"], C.c_double)

    ade = (N.PfAdeDesc * max(len(t["ade"]), 1))()
    for n_, med in enumerate(t["ade"]):
        cells = np.ascontiguousarray(med.cells, dtype=np.int64)
        abc = np.ascontiguousarray(med.abc, dtype=np.float64)
        keep += [cells, abc]
        ade[n_].n_cells = cells.size
        ade[n_].cells = N.as_ptr(cells, C.c_int64)
        ade[n_].n_poles = abc.shape[0]
        ade[n_].abc = N.as_ptr(abc, C.c_double)
        ade[n_].inv_eps = med.inv_eps
    d.n_ade = len(t["ade"])
    d.ade = ade

    dft = (N.PfDftDesc * max(len(t["dft"]), 1))()
    dft_slots: Dict[str, tuple] = {}
    for n_, spec in enumerate(t["dft"]):
        m = spec["mon"]
        comps = [N.COMP_CODE[c] for c in m.components]
        if len(set(comps)) != len(comps) or not 1 <= len(comps) <= 6:
            raise ValueError(f"DFTMonitor {m.name!r}: components must be distinct field names")
        ph_e = np.ascontiguousarray(spec["phase_e"]).view(np.float64)
        ph_h = np.ascontiguousarray(spec["phase_h"]).view(np.float64)
        keep += [ph_e, ph_h]
        dd = dft[n_]
        dd.n_comp = len(comps)
        for s, c in enumerate(comps):
            dd.comps[s] = c
        start, count, stride = spec["box"]
        dd.box.start[:], dd.box.count[:], dd.box.stride[:] = start, count, stride
        dd.n_freq = len(m.freqs)
        dd.interval = m.interval
        dd.n_samples = spec["n_samples"]
        dd.phase_e = N.as_ptr(ph_e, C.c_double)
        dd.phase_h = N.as_ptr(ph_h, C.c_double)
        dd.dt = sim.dt
        dft_slots[m.name] = (n_, list(m.components), tuple(count), len(m.freqs))
    d.n_dft = len(t["dft"])
    d.dft = dft

    flux = (N.PfFluxDesc * max(len(t["flux"]), 1))()
    flux_names: List[str] = []
    for n_, spec in enumerate(t["flux"]):
        flux[n_].axis = spec["axis"]
        flux[n_].index = spec["index"]
        flux[n_].dA = spec["dA"]
        flux[n_].dt = sim.dt
        flux_names.append(spec["mon"].name)
    d.n_flux = len(t["flux"])
    d.flux = flux
    return d, keep, t, dft_slots, flux_names


def streaming_mode(sim, npdt) -> bool:
    """Whether ``run_hip`` keeps the domain in host memory and streams x-slab
    tiles through the device (``rustbackend.py:264-269``): forced by
    ``PHOTONFDTD_HIP_STREAM=1``, refused by ``=0``, otherwise chosen when the
    resident stepper (two field levels, ce and its march copy, psi) would
    take more than 85 % of the free device memory."""
    env = os.environ.get("PHOTONFDTD_HIP_STREAM")
    if env in ("0", "1"):
        return env == "1"
    nx, ny, nz = sim.grid.shape
    need_mb = 15.0 * nx * ny * nz * np.dtype(npdt).itemsize / 1e6
    _, free_mb, _ = device_info()
    return need_mb > 0.85 * free_mb


class HipEngine:
    """One simulation behind the C ABI: device-resident, or with
    ``streaming`` host-resident and streamed through the device in x-slab
    tiles (``PHOTONFDTD_HIP_TB`` steps per visit, ``PHOTONFDTD_HIP_TILE``
    planes per tile)."""

    def __init__(self, sim, npdt, exact: bool, flags: int = 0, streaming: bool = False):
        self.lib = N.load_library()
        self.npdt = np.dtype(npdt).type
        self.shape = sim.grid.shape
        self.streaming = bool(streaming)
        if self.streaming:
            flags |= N.PF_FLAG_STREAM
        d, keep, self.t, self.dft_slots, self.flux_names = build_desc(sim, self.npdt, exact, flags)
        if self.streaming:
            what = ("dispersive media" if d.n_ade else "flux monitors" if d.n_flux else
                    "subpixel smoothing" if d.ce_y else None)
            if what:
                raise NotImplementedError(
                    f"{what} need the device-resident HIP stepper, and this run streams its "
                    "domain through the device (set PHOTONFDTD_HIP_STREAM=0 to keep it resident)")
            d.t_block = int(os.environ.get("PHOTONFDTD_HIP_TB", "32"))
            d.tile_rows = int(os.environ.get("PHOTONFDTD_HIP_TILE", "0"))

        h = C.c_void_p()
        N.check(self.lib, self.lib.pf_create(C.byref(d), C.byref(h)), "pf_create")
        self.h = h
        self._cache: Dict[str, np.ndarray] = {}
        del keep

    def close(self):
        if getattr(self, "h", None):
            self.lib.pf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def advance(self, step0: int, n: int) -> None:
        self._cache.clear()
        N.check(self.lib, self.lib.pf_advance(self.h, int(step0), int(n)), "pf_advance")

    def get(self, comp: str) -> np.ndarray:
        if comp not in self._cache:
            out = np.empty(self.shape, dtype=self.npdt)
            N.check(self.lib, self.lib.pf_read_field(self.h, N.COMP_CODE[comp],
                                                     out.ctypes.data), "pf_read_field")
            self._cache[comp] = out
        return self._cache[comp]

    def read_box(self, comp: str, start, count, stride) -> np.ndarray:
        box = N.PfBox()
        box.start[:], box.count[:], box.stride[:] = start, count, stride
        out = np.empty(tuple(count), dtype=self.npdt)
        N.check(self.lib, self.lib.pf_read_box(self.h, N.COMP_CODE[comp], C.byref(box),
                                               out.ctypes.data), "pf_read_box")
        return out

    def read_box_q(self, comp: str, start, count, stride, bits: int):
        """A box quantised on the device for compressed storage: (int8 or
        float16 array, scale) as ``CompressedFieldSeries`` stores a frame."""
        box = N.PfBox()
        box.start[:], box.count[:], box.stride[:] = start, count, stride
        out = np.empty(tuple(count), dtype=np.int8 if bits == 8 else np.float16)
        scale = C.c_double()
        N.check(self.lib, self.lib.pf_read_box_q(self.h, N.COMP_CODE[comp], C.byref(box), int(bits),
                                                 out.ctypes.data, C.byref(scale)), "pf_read_box_q")
        return out, float(scale.value)

    def set(self, comp: str, values: np.ndarray) -> None:
        arr = np.ascontiguousarray(values, dtype=self.npdt).reshape(self.shape)
        self._cache.pop(comp, None)
        N.check(self.lib, self.lib.pf_write_field(self.h, N.COMP_CODE[comp], arr.ctypes.data),
                "pf_write_field")

    def info(self) -> Dict[str, int]:
        """Stepping path, z-row padding and march tile counts (``pf_stepper_info``)."""
        out = N.PfInfo()
        N.check(self.lib, self.lib.pf_stepper_info(self.h, C.byref(out)), "pf_stepper_info")
        return {name: int(getattr(out, name)) for name, _ in out._fields_}

    def dft_flush(self) -> None:
        """Fold every buffered DFT sample into the accumulators (synchronous)."""
        N.check(self.lib, self.lib.pf_dft_flush(self.h), "pf_dft_flush")

    def dft(self, name: str) -> Dict[str, np.ndarray]:
        mon, comps, count, n_freq = self.dft_slots[name]
        out = {}
        for slot, c in enumerate(comps):
            buf = np.empty((n_freq,) + count, dtype=np.complex128)
            N.check(self.lib, self.lib.pf_dft_read(self.h, mon, slot,
                                                   buf.view(np.float64).ctypes.data_as(
                                                       C.POINTER(C.c_double))), "pf_dft_read")
            out[c] = buf
        return out

    def field_tensor(self, comp: str):
        """The current level of field ``comp`` as a torch tensor on the device
        that aliases the stepper's memory (``pf_field_ptr``; no copy). It is
        valid until the next advance, ``set`` or ``close``; writes through it
        change the simulation state."""
        import torch
        v = N.PfView()
        N.check(self.lib, self.lib.pf_field_ptr(self.h, N.COMP_CODE[comp], C.byref(v)),
                "pf_field_ptr")
        return N.to_torch(v, self, torch.cuda.current_device())

    def dft_tensors(self, name: str):
        """``{comp: complex128 tensor (n_freq, *box)}`` on the device, aliasing
        the accumulators of DFT monitor ``name`` after every buffered sample
        is folded in (``pf_dft_ptr``; no copy). Valid until the next advance
        or ``close``."""
        import torch
        mon, comps, count, n_freq = self.dft_slots[name]
        v = N.PfView()
        N.check(self.lib, self.lib.pf_dft_ptr(self.h, mon, C.byref(v)), "pf_dft_ptr")
        acc = torch.view_as_complex(N.to_torch(v, self, torch.cuda.current_device()))
        return {c: acc[slot].reshape((n_freq,) + tuple(count)) for slot, c in enumerate(comps)}

    def enable_checkpoints(self, interval: int) -> None:
        N.check(self.lib, self.lib.pf_checkpoint_enable(self.h, int(interval)),
                "pf_checkpoint_enable")

    def adjoint(self, cotangents: Dict[str, Dict[str, np.ndarray]]) -> np.ndarray:
        """dJ/dce per node given dJ/d(DFT) as complex arrays (Re/Im parts of
        the derivative w.r.t. the real and imaginary accumulator parts)."""
        keep, items = [], []
        for name, comps in cotangents.items():
            mon, names, count, n_freq = self.dft_slots[name]
            for c, g in comps.items():
                arr = np.ascontiguousarray(np.asarray(g, dtype=np.complex128)
                                           .reshape((n_freq,) + count))
                keep.append(arr)
                items.append((mon, names.index(c), arr))
        cot = (N.PfCotangent * max(len(items), 1))()
        for n_, (mon, slot, arr) in enumerate(items):
            cot[n_].mon, cot[n_].slot = mon, slot
            cot[n_].g = arr.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
        grad = np.zeros(self.shape, dtype=np.float64)
        N.check(self.lib, self.lib.pf_adjoint(self.h, len(items), cot, N.as_ptr(grad, C.c_double)),
                "pf_adjoint")
        return grad

    def adjoint_device(self, cotangents):
        """``adjoint`` with the cotangents given as torch tensors on the
        device; returns dJ/dce as a float64 device tensor (no host copies)."""
        import torch
        keep, items = [], []
        for name, comps in cotangents.items():
            mon, names, count, n_freq = self.dft_slots[name]
            for c, g in comps.items():
                t = torch.view_as_real(g.to(torch.complex128).reshape((n_freq,) + tuple(count))
                                       .contiguous())
                keep.append(t)
                items.append((mon, names.index(c), t))
        cot = (N.PfCotangent * max(len(items), 1))()
        for n_, (mon, slot, t) in enumerate(items):
            cot[n_].mon, cot[n_].slot = mon, slot
            cot[n_].g = C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double))
        grad = torch.zeros(self.shape, dtype=torch.float64, device=keep[0].device if keep else "cuda")
        torch.cuda.synchronize()
        N.check(self.lib, self.lib.pf_adjoint(self.h, len(items), cot,
                                              C.cast(C.c_void_p(grad.data_ptr()),
                                                     C.POINTER(C.c_double))), "pf_adjoint")
        return grad

    def flux(self) -> Dict[str, float]:
        if not self.flux_names:
            return {}
        buf = np.zeros(len(self.flux_names), dtype=np.float64)
        N.check(self.lib, self.lib.pf_flux_read(self.h, N.as_ptr(buf, C.c_double)),
                "pf_flux_read")
        return {n: float(v) for n, v in zip(self.flux_names, buf)}


def device_info():
    lib = N.load_library()
    name = C.create_string_buffer(256)
    free_mb, total_mb = C.c_int64(), C.c_int64()
    N.check(lib, lib.pf_device_info(name, 256, C.byref(free_mb), C.byref(total_mb)),
            "pf_device_info")
    return name.value.decode(), int(free_mb.value), int(total_mb.value)


def run_loop(sim, engine: HipEngine) -> Result:
    """Chunked stepping, FieldMonitor recording and Result assembly."""
    grid = sim.grid
    dt = sim.dt
    n_steps = sim.n_steps
    mon_dtype = sim.dtypes["monitors"]
    result = Result(times=np.arange(n_steps) * dt)

    rec: Dict[str, dict] = {}
    for m in sim.monitors:
        if isinstance(m, FieldMonitor):
            if m.times is not None:
                steps = {int(round(tm / dt)) for tm in m.times}
            else:
                steps = set(range(0, n_steps, m.interval))
            steps = {s for s in steps if 0 <= s < n_steps}
            zplane = (None if m.plane_z is None else
                      int(np.argmin(np.abs(np.asarray(grid.coords[2]) - m.plane_z))))
            rec[m.name] = dict(mon=m, steps=steps, n=len(steps), rows={c: None for c in m.components},
                               times=[], box=_snap_box(grid.shape, m.downsample, zplane=zplane),
                               codec=None)
        elif isinstance(m, DFTMonitor):
            result.dft_freqs[m.name] = np.asarray(m.freqs, dtype=np.float64)

    def record(r, step):
        row = len(r["times"])
        start, count, stride = r["box"]
        m = r["mon"]
        if m.compression is not None:
            # quantised on the device; the series runs the byte codec
            # (storage.py:128-141 on the reference's host)
            if r["codec"] is None:
                r["codec"] = get_codec(m.compression)
            for c in m.components:
                q, scale = engine.read_box_q(c, start, count, stride, m.compression_bits)
                if r["rows"][c] is None:
                    r["rows"][c] = CompressedFieldSeries(q.shape, mon_dtype, r["codec"],
                                                         bits=m.compression_bits)
                r["rows"][c].append_quantized(q, scale)
            r["times"].append(step * dt)
            return
        full = count == list(grid.shape) and stride == [1, 1, 1]
        for c in m.components:
            snap = engine.get(c) if full else engine.read_box(c, start, count, stride)
            if r["rows"][c] is None:
                r["rows"][c] = np.empty((r["n"],) + snap.shape, dtype=mon_dtype)
            r["rows"][c][row] = snap
        r["times"].append(step * dt)

    stops = sorted(set().union(*(r["steps"] for r in rec.values()))) if rec else []
    tick = max(n_steps // 100, 1)
    pos = 0

    def advance_to(end):
        nonlocal pos
        while pos < end:
            n = min(end - pos, tick if sim.progress_callback else end - pos)
            engine.advance(pos, n)
            pos += n
            if sim.progress_callback is not None:
                sim.progress_callback(min(pos, n_steps), n_steps)

    for s in stops:
        advance_to(s + 1)
        for r in rec.values():
            if s in r["steps"]:
                record(r, s)
    advance_to(n_steps)
    if sim.progress_callback is not None:
        sim.progress_callback(n_steps, n_steps)

    for m in sim.monitors:
        if isinstance(m, FieldMonitor):
            r = rec[m.name]
            if r["times"]:
                if m.compression is not None:
                    for series in r["rows"].values():
                        series.finalize()
                result.fields[m.name] = dict(r["rows"])
                result.monitor_times[m.name] = np.array(r["times"])
            else:
                result.fields[m.name] = {c: np.zeros((0,) + grid.shape) for c in m.components}
                result.monitor_times[m.name] = np.array([])
        elif isinstance(m, DFTMonitor):
            result.dft[m.name] = engine.dft(m.name)
        elif isinstance(m, FluxMonitor):
            pass
    result.flux.update(engine.flux())
    return result


def run_hip(sim) -> Result:
    """Run ``sim`` on the gfx950 stepper and return a ``Result``."""
    npdt = validate(sim)
    streaming = streaming_mode(sim, npdt)
    engine = HipEngine(sim, npdt, exact_mode(npdt), streaming=streaming)
    if sim.verbose:
        name, free_mb, total_mb = device_info()
        mode = "streaming (host-resident)" if streaming else "device-resident"
        print(f"[photonfdtd_amd] HIP backend ({mode}), device: {name} "
              f"({free_mb}/{total_mb} MB free)", file=sys.stderr, flush=True)
    try:
        return run_loop(sim, engine)
    finally:
        engine.close()
