"""NumPy restatement of the reference's in-core FDTD time loop (the oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Follows ``src/photonfdtd/simulation.py`` operation by operation so a float64
run here is the reference's float64 trajectory:

* bulk update per component: ``cbuf = dA - dB`` (forward differences
  ``(a[1:] - a[:-1]) / d``), ``cbuf *= K`` (``K = ce`` for E, ``-dt/mu0``
  for H), ``F += cbuf`` over the cropped region (``:1252-1289``);
* PML correction afterwards on the slabs with ``c != 0``:
  ``psi = b psi + c d`` then ``F += sign K psi`` (``:1303-1309``);
* per step: H pass, H sources at (n+1/2)dt, capture E^n on dispersive
  cells, E pass, ADE correction, E sources at (n+1)dt, flux, monitors
  (``:1319-1406``);
* DFT: ``acc += exp(1j w t) * f * dt`` in complex128 (``:1116-1133``).

The setup tables (CPML, ce, ADE) are recomputed here from the scene rather
than taken from the package, so a package-side table bug shows up as a
parity failure.
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional

import numpy as np

C0 = 299_792_458.0
EPS0 = 8.854_187_8128e-12
MU0 = 1.256_637_062_12e-6
ETA0 = math.sqrt(MU0 / EPS0)

_FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def cpml_axis(n: int, d: float, npml: int, dt: float, order=3, sigma_factor=0.8,
              alpha_max=0.24, R=1e-7):
    """``(b_e, c_e, b_h, c_h)`` for one axis -- ``boundaries.py:31-97``."""
    if npml == 0:
        z = np.zeros(n)
        return z.copy(), z.copy(), z.copy(), z.copy()
    sigma_max = sigma_factor * (-(order + 1) * np.log(R) / (2.0 * ETA0 * npml * d))

    def bc(u):
        sigma = sigma_max * u ** order
        alpha = alpha_max * (1.0 - u)
        b = np.exp(-(sigma + alpha) * dt / EPS0)
        c = np.zeros_like(sigma)
        m = (sigma + alpha) > 0
        c[m] = sigma[m] / (sigma + alpha)[m] * (b[m] - 1.0)
        return b, c

    span = max(npml - 1, 1)
    out = []
    for shift in (0.0, 0.5):
        idx = np.arange(n) + shift
        b = np.zeros(n)
        c = np.zeros(n)
        b[:npml], c[:npml] = bc(np.clip((npml - 1 - idx[:npml]) / span, 0.0, 1.0))
        b[-npml:], c[-npml:] = bc(np.clip((idx[-npml:] - (n - npml)) / span, 0.0, 1.0))
        out += [b, c]
    return out[0], out[1], out[2], out[3]


def _runs(mask1d):
    """Maximal runs of True as (start, stop)."""
    runs, start = [], None
    for i, v in enumerate(mask1d):
        if v and start is None:
            start = i
        elif not v and start is not None:
            runs.append((start, i))
            start = None
    if start is not None:
        runs.append((start, len(mask1d)))
    return runs


def _ade_tables(sim, dt):
    """Dispersive regions and pole coefficients -- ``simulation.py:572-636``."""
    structures = sim.structures
    if not any(getattr(getattr(s, "medium", None), "is_dispersive", False) for s in structures):
        return []
    owner = np.full(sim.grid.shape, -1, dtype=np.int32)
    media, key = [], {}
    for s in structures:
        med = getattr(s, "medium", None)
        mask = s.region_mask(sim.grid)
        if getattr(med, "is_dispersive", False):
            if id(med) not in key:
                key[id(med)] = len(media)
                media.append(med)
            owner[mask] = key[id(med)]
        else:
            owner[mask] = -1
    w0dt = max((max((p.omega0 for p in m.poles), default=0.0) * dt for m in media), default=0.0)
    if w0dt >= 2.0:
        raise ValueError(f"pole with omega0*dt = {w0dt:.2f} >= 2: ADE update unstable")
    out = []
    for mid, med in enumerate(media):
        idx = np.nonzero(owner == mid)
        if idx[0].size == 0:
            continue
        coeffs = []
        for p in med.poles:
            den = 1.0 + p.gamma * dt
            coeffs.append(((2.0 - p.omega0 ** 2 * dt ** 2) / den, (p.gamma * dt - 1.0) / den,
                           (EPS0 * p.strength * dt ** 2) / den))
        out.append(dict(idx=idx, abc=coeffs, inv=1.0 / (EPS0 * med.eps_inf)))
    return out


def _smoothed_eps(grid, structures, factor):
    """Subpixel-smoothed ``(eps_xx, eps_yy, eps_zz)`` (``smoothing.py:119-172``),
    restated over the whole refined domain at once: each native cell's
    arithmetic / harmonic mean of the permittivity stamped on its ``factor``
    sub-nodes per active axis, and ``eps_ii = eps_a + (eps_h - eps_a) n_i^2``
    with n the gradient of eps_a times its reciprocal norm (0 where the norm
    is below 1e-9 of its maximum)."""
    from types import SimpleNamespace
    shape = grid.shape
    f = [factor if n > 1 else 1 for n in shape]
    coords = []
    for a in range(3):
        n = shape[a] * f[a]
        d = grid.cell_size[a] / factor if shape[a] > 1 else 0.0
        coords.append((np.arange(n) - (n - 1) / 2) * d if n > 1 else np.zeros(1))
    view = SimpleNamespace(shape=tuple(c.size for c in coords), coords=tuple(coords))
    eps = np.ones(view.shape)
    for st in structures:
        eps[st.region_mask(view)] = st.medium.eps_r
    blocks = eps.reshape(shape[0], f[0], shape[1], f[1], shape[2], f[2])
    arith = blocks.mean(axis=(1, 3, 5))
    harm = 1.0 / (1.0 / blocks).mean(axis=(1, 3, 5))
    g = [np.gradient(arith, grid.cell_size[a], axis=a) if shape[a] > 1 else np.zeros(shape)
         for a in range(3)]
    mag = np.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
    ok = mag > max(mag.max() * 1e-9, np.finfo(np.float64).tiny)
    inv = np.where(ok, 1.0 / np.where(ok, mag, 1.0), 0.0)
    unit = [gi * inv for gi in g]
    return tuple(arith + (harm - arith) * u ** 2 for u in unit)


def run_oracle(sim, return_state: bool = False):
    """Time-step ``sim`` (a ``photonfdtd_amd.Simulation``) on the CPU.

    Returns ``(result_dict, state)`` where ``result_dict`` has the same keys as
    the reference ``Result`` (``times, fields, monitor_times, flux, dft,
    dft_freqs``) and ``state`` (when requested) holds the final six fields.
    """
    from photonfdtd_amd.monitors import DFTMonitor, FieldMonitor, FluxMonitor

    grid = sim.grid
    nx, ny, nz = grid.shape
    shape = grid.shape
    cell = tuple(d if d > 0 else 1.0 for d in grid.cell_size)
    dt = sim.dt
    n_steps = sim.n_steps
    dts = sim.dtypes
    psi_dt = dts["psi"]
    mon_dt = dts["monitors"]

    F = {c: np.zeros(shape, dtype=dts[c]) for c in _FIELDS}
    eps = np.ones(shape, dtype=dts["eps_r"])
    for s in sim.structures:
        s.stamp(grid, eps)
    ce = (dt / (eps * EPS0)).astype(dts["eps_r"], copy=False)
    ce_of = {"Ex": ce, "Ey": ce, "Ez": ce}
    if getattr(sim, "subpixel", False):
        # per-component coefficients from the smoothed diagonal tensor
        # (simulation.py:956-969)
        comps = _smoothed_eps(grid, sim.structures, sim.subpixel_factor)
        for c, e in zip(("Ex", "Ey", "Ez"), comps):
            ce_of[c] = (dt / (e.astype(dts["eps_r"]) * EPS0)).astype(dts["eps_r"], copy=False)
    ch = dts["Hx"](dt / MU0)

    be, cE, bh, cH = [], [], [], []
    for ax in range(3):
        n = shape[ax]
        b1, c1, b2, c2 = cpml_axis(n, cell[ax] if n > 1 else 1.0,
                                   grid.pml_layers[ax] if n > 1 else 0, dt,
                                   sim.cpml_params.poly_order, sim.cpml_params.sigma_factor,
                                   sim.cpml_params.alpha_max, sim.cpml_params.R_target)
        be.append(b1); cE.append(c1); bh.append(b2); cH.append(c2)

    def crop(ax, e_side):
        if shape[ax] <= 1:
            return slice(None)
        return slice(1, None) if e_side else slice(0, -1)

    # (target, (axis, source) of the + term, (axis, source) of the - term, is_E)
    table = [("Hx", (1, "Ez"), (2, "Ey"), False), ("Hy", (2, "Ex"), (0, "Ez"), False),
             ("Hz", (0, "Ey"), (1, "Ex"), False), ("Ex", (1, "Hz"), (2, "Hy"), True),
             ("Ey", (2, "Hx"), (0, "Hz"), True), ("Ez", (0, "Hy"), (1, "Hx"), True)]
    plans = {False: [], True: []}
    for tgt, (axA, srcA), (axB, srcB), is_e in table:
        reg = [slice(None)] * 3
        reg[axA] = crop(axA, is_e)
        reg[axB] = crop(axB, is_e)
        reg = tuple(reg)
        rshape = F[tgt][reg].shape
        terms = []
        for ax, src, other, sign in ((axA, srcA, axB, 1.0), (axB, srcB, axA, -1.0)):
            if shape[ax] <= 1:
                continue
            csl = crop(ax, is_e)
            bcoef = (be if is_e else bh)[ax][csl]
            ccoef = (cE if is_e else cH)[ax][csl]
            slabs = []
            for s0, s1 in _runs(ccoef != 0.0):
                sl = [slice(None)] * 3
                sl[ax] = slice(s0, s1)
                rs = [1, 1, 1]
                rs[ax] = s1 - s0
                pshape = list(rshape)
                pshape[ax] = s1 - s0
                slabs.append((tuple(sl), bcoef[s0:s1].astype(psi_dt).reshape(rs),
                              ccoef[s0:s1].astype(psi_dt).reshape(rs),
                              np.zeros(pshape, dtype=psi_dt)))
            terms.append((ax, src, other, sign, slabs))
        plans[is_e].append((tgt, reg, ce_of[tgt][reg] if is_e else None, terms))

    scratch = np.empty(shape, dtype=dts["Ex"])

    def deriv(a, ax):
        hi = [slice(None)] * 3
        lo = [slice(None)] * 3
        hi[ax] = slice(1, None)
        lo[ax] = slice(0, -1)
        return (a[tuple(hi)] - a[tuple(lo)]) / cell[ax]

    def sweep(is_e):
        for tgt, reg, Kreg, terms in plans[is_e]:
            ds = []
            for ax, src, other, sign, slabs in terms:
                d = deriv(F[src], ax)
                if shape[other] > 1:
                    o = [slice(None)] * 3
                    o[other] = crop(other, is_e)
                    d = d[tuple(o)]
                ds.append(d)
            if not ds:
                continue
            buf = scratch[reg]
            for n_, ((ax, src, other, sign, slabs), d) in enumerate(zip(terms, ds)):
                if n_ == 0:
                    if sign > 0:
                        np.copyto(buf, d)
                    else:
                        np.negative(d, out=buf)
                elif sign > 0:
                    np.add(buf, d, out=buf)
                else:
                    np.subtract(buf, d, out=buf)
            Freg = F[tgt][reg]
            np.multiply(buf, Kreg if is_e else -ch, out=buf)
            Freg += buf
            for (ax, src, other, sign, slabs), d in zip(terms, ds):
                for sl, b, c, psi in slabs:
                    psi[...] = b * psi + c * d[sl]
                    k = Kreg[sl] if is_e else -ch
                    Freg[sl] += sign * k * psi

    ade = _ade_tables(sim, dt)
    for st in ade:
        n_c = st["idx"][0].size
        npole = len(st["abc"])
        st["cur"] = [[np.zeros(n_c, dtype=dts["eps_r"]) for _ in range(3)] for _ in range(npole)]
        st["prev"] = [[np.zeros(n_c, dtype=dts["eps_r"]) for _ in range(3)] for _ in range(npole)]

    src_cells = [(s, grid.index_at(s.position)) for s in sim.sources]

    # monitors
    res = dict(times=np.arange(n_steps) * dt, fields={}, monitor_times={}, flux={}, dft={},
               dft_freqs={})
    fm_steps, fm_rows, fm_z, fm_times, fm_n = {}, {}, {}, {}, {}
    dft_acc, dft_w, dft_plane = {}, {}, {}
    for m in sim.monitors:
        if isinstance(m, FieldMonitor):
            # A compressed monitor records the snapshots the reference hands to
            # CompressedFieldSeries.append (simulation.py:1088-1097): the field
            # values, uncompressed; the quantisation is checked on its own.
            steps = ({int(round(t / dt)) for t in m.times} if m.times is not None
                     else set(range(0, n_steps, m.interval)))
            fm_steps[m.name] = steps
            fm_n[m.name] = sum(1 for s in steps if 0 <= s < n_steps)
            fm_rows[m.name] = {c: None for c in m.components}
            fm_times[m.name] = []
            if m.plane_z is not None:
                zi = int(np.argmin(np.abs(np.asarray(grid.coords[2]) - m.plane_z)))
                fm_z[m.name] = slice(zi, zi + 1)
            else:
                fm_z[m.name] = None
        elif isinstance(m, DFTMonitor):
            dft_acc[m.name] = {c: None for c in m.components}
            dft_w[m.name] = 2.0 * np.pi * np.asarray(m.freqs, dtype=np.float64)
            res["dft_freqs"][m.name] = np.asarray(m.freqs, dtype=np.float64)
            pl = m.plane()
            dft_plane[m.name] = None if pl is None else (
                pl[0], int(np.argmin(np.abs(np.asarray(grid.coords[pl[0]]) - pl[1]))))
        elif isinstance(m, FluxMonitor):
            res["flux"][m.name] = 0.0

    def field_view(f, ds, zsl):
        if zsl is not None:
            return f[::ds, ::ds, zsl]
        return f[::ds, ::ds, ::ds] if ds > 1 else f

    def plane_view(f, ds, plane):
        sl = [slice(None, None, ds) if ds > 1 else slice(None) for _ in range(3)]
        if plane is not None:
            sl[plane[0]] = slice(plane[1], plane[1] + 1)
        return f[tuple(sl)]

    def flux(m):
        ax = {"x": 0, "y": 1, "z": 2}[m.plane_axis]
        co = grid.coords[ax]
        i = 0 if co.size == 1 else int(np.argmin(np.abs(co - m.plane_position)))
        ext = [grid.cell_size[a] if shape[a] > 1 else 1.0 for a in range(3)]
        if ax == 0:
            S = F["Ey"][i] * F["Hz"][i] - F["Ez"][i] * F["Hy"][i]
            dA = ext[1] * ext[2]
        elif ax == 1:
            S = F["Ez"][:, i] * F["Hx"][:, i] - F["Ex"][:, i] * F["Hz"][:, i]
            dA = ext[0] * ext[2]
        else:
            S = F["Ex"][:, :, i] * F["Hy"][:, :, i] - F["Ey"][:, :, i] * F["Hx"][:, :, i]
            dA = ext[0] * ext[1]
        return float(np.sum(S) * dA)

    E = ("Ex", "Ey", "Ez")
    for step in range(n_steps):
        t_h = (step + 0.5) * dt
        t_e = (step + 1.0) * dt
        sweep(False)
        for src, (i, j, k) in src_cells:
            if src.component[0] == "H":
                F[src.component][i, j, k] += src.amplitude * src.waveform(np.array([t_h]))[0]
        for st in ade:
            st["En"] = [F[c][st["idx"]].copy() for c in E]
        sweep(True)
        for st in ade:
            for c in range(3):
                delta = None
                for p, (a, b, cc) in enumerate(st["abc"]):
                    pc, pp = st["cur"][p][c], st["prev"][p][c]
                    pn = a * pc + b * pp + cc * st["En"][c]
                    inc = pn - pc
                    delta = inc if delta is None else delta + inc
                    st["prev"][p][c] = pc
                    st["cur"][p][c] = pn
                if delta is not None:
                    F[E[c]][st["idx"]] -= st["inv"] * delta
        for src, (i, j, k) in src_cells:
            if src.component[0] == "E":
                F[src.component][i, j, k] += src.amplitude * src.waveform(np.array([t_e]))[0]
        for m in sim.monitors:
            if isinstance(m, FieldMonitor):
                if step in fm_steps[m.name]:
                    row = len(fm_times[m.name])
                    for c in m.components:
                        snap = field_view(F[c], m.downsample, fm_z[m.name])
                        if fm_rows[m.name][c] is None:
                            fm_rows[m.name][c] = np.empty(
                                (fm_n[m.name],) + snap.shape,
                                dtype=snap.dtype if m.compression is not None else mon_dt)
                        fm_rows[m.name][c][row] = snap
                    fm_times[m.name].append(step * dt)
            elif isinstance(m, DFTMonitor):
                if step % m.interval == 0:
                    for c in m.components:
                        snap = plane_view(F[c], m.downsample, dft_plane[m.name])
                        t = t_e if c[0] == "E" else t_h
                        ph = np.exp(1j * dft_w[m.name] * t).reshape((-1,) + (1,) * snap.ndim)
                        if dft_acc[m.name][c] is None:
                            dft_acc[m.name][c] = np.zeros((ph.shape[0],) + snap.shape,
                                                          dtype=np.complex128)
                        dft_acc[m.name][c] += ph * snap[None, ...] * dt
            elif isinstance(m, FluxMonitor):
                res["flux"][m.name] += flux(m) * dt

    for m in sim.monitors:
        if isinstance(m, FieldMonitor):
            if fm_times[m.name]:
                res["fields"][m.name] = dict(fm_rows[m.name])
                res["monitor_times"][m.name] = np.array(fm_times[m.name])
            else:
                res["fields"][m.name] = {c: np.zeros((0,) + shape) for c in m.components}
                res["monitor_times"][m.name] = np.array([])
        elif isinstance(m, DFTMonitor):
            res["dft"][m.name] = {
                c: (dft_acc[m.name][c] if dft_acc[m.name][c] is not None else
                    np.zeros((len(m.freqs),) + shape, dtype=np.complex128))
                for c in m.components}
    if return_state:
        return res, F
    return res
