"""2-D inverse design with the surrogate: scene, FDTD labels, and the two loops.

The problem: a 512 x 512-cell TM domain (dx = 20 nm, 12-cell CPML, float32)
in silica (eps 2.085); an input silicon waveguide (eps 12.1, 0.4 um) on the
axis, an output waveguide offset by 0.8 um, and between them a 128 x 128-cell
design region whose permittivity is 2.085 + rho (12.1 - 2.085). An Ez dipole
in the input guide launches a 1.55 um, 10 fs pulse; a DFT plane across the
output guide records (Ez, Hy) at 16 frequencies in [c/1.7 um, c/1.4 um]. The
label of a design is its transmission spectrum: the x-directed spectral flux
-Re(Ez conj(Hy)) through a 1.6 um window on the output guide, divided by the
same flux for a straight guide through an empty design region. The objective
is the mean transmission over the band.

Labels come from the batched HIP stepper (``SimBatch``: one persistent set of
device simulations whose permittivity is replaced per design); the plain
method is gradient ascent with the hand-written adjoint (``value_and_grad_eps``);
the surrogate method proposes candidate batches from the network's input
gradient, verifies every candidate with FDTD, and falls back to an adjoint step
when the network's error on the chosen design exceeds 2 %.
"""
from __future__ import annotations

import math
import time
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

DX = 20e-9
N_GRID = 512
N_DESIGN = 128
EPS_LO, EPS_HI = 2.085, 12.1
WG_WIDTH = 0.4e-6
OFFSET = 0.8e-6
WINDOW = 0.8e-6          # half width of the output flux window
N_FREQ = 16


def random_design(seed: int, n: int = N_DESIGN, sigma: float = 4.0) -> np.ndarray:
    """Binary map: white noise (default_rng(seed)) -> Gaussian filter (sigma
    cells) -> threshold at 0."""
    from scipy.ndimage import gaussian_filter
    noise = np.random.default_rng(seed).standard_normal((n, n))
    return (gaussian_filter(noise, sigma) > 0).astype(np.float32)


def gray_design(seed: int, n: int = N_DESIGN, sigma: float = 4.0) -> np.ndarray:
    """Continuous map: the filtered noise of `random_design` through a
    logistic of random sharpness (1 = soft gray, 8 = almost binary)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.standard_normal((n, n)), sigma)
    beta = rng.uniform(1.0, 8.0)
    return (1.0 / (1.0 + np.exp(-beta * f / f.std()))).astype(np.float32)


def training_designs(n: int, seed0: int = 0) -> np.ndarray:
    """Half binary, half gray maps (seeds seed0 .. seed0 + n - 1)."""
    return np.stack([random_design(seed0 + k) if k % 2 == 0 else gray_design(seed0 + k)
                     for k in range(n)])


def freqs(pf) -> np.ndarray:
    return np.linspace(pf.C_0 / 1.7e-6, pf.C_0 / 1.4e-6, N_FREQ)


def scene(pf, steps: int = 4000, precision: str = "float32", straight: bool = False,
          n: int = N_GRID, backend: str = "hip"):
    """The design scene with an empty (eps 2.085) design region; `straight`
    gives the normalisation scene (one straight guide, no offset)."""
    L = n * DX
    hd = N_DESIGN * DX / 2
    far = L / 2 + 1e-6
    out_y = 0.0 if straight else OFFSET
    structures = [pf.Box(center=(0, 0), size=(3 * L, 3 * L), medium=pf.Medium(EPS_LO))]
    wg = pf.Medium(EPS_HI)
    if straight:
        structures.append(pf.Box(center=(0, 0), size=(2 * far, WG_WIDTH), medium=wg))
    else:
        structures.append(pf.Box(center=(-(far + hd) / 2, 0), size=(far - hd, WG_WIDTH), medium=wg))
        structures.append(pf.Box(center=((far + hd) / 2, out_y), size=(far - hd, WG_WIDTH),
                                 medium=wg))
    src = pf.PointDipole(position=(-L / 2 + 1.5e-6, 0.0), component="Ez",
                         waveform=pf.GaussianPulse(freq0=pf.C_0 / 1.55e-6, fwhm=10e-15))
    mon = pf.DFTMonitor(name="out", components=("Ez", "Hy"), freqs=freqs(pf), plane_axis="x",
                        plane_position=L / 2 - 1.5e-6)
    grid = pf.Grid(size=(L, L), cell_size=DX, pml_layers=(12, 12, 0))
    probe = pf.Simulation(grid, structures=structures, sources=[src], monitors=[mon],
                          run_time=1e-15, precision=precision, backend=backend)
    return pf.Simulation(grid, structures=structures, sources=[src], monitors=[mon],
                         run_time=(steps - 0.5) * probe.dt, precision=precision, backend=backend)


def design_window(sim) -> tuple:
    """(i0, j0): first grid index of the centred design window."""
    nx, ny, _ = sim.grid.shape
    return (nx - N_DESIGN) // 2, (ny - N_DESIGN) // 2


def eps_with(sim, rho: np.ndarray) -> np.ndarray:
    """Full permittivity grid of `sim` with the design window set from rho."""
    eps = np.array(sim.eps_r, dtype=np.float64, copy=True)
    i0, j0 = design_window(sim)
    r = np.clip(np.asarray(rho, dtype=np.float64), 0, 1)
    eps[i0:i0 + N_DESIGN, j0:j0 + N_DESIGN, 0] = EPS_LO * (1 - r) + EPS_HI * r   # exact at 0, 1
    return eps


def flux_window(sim, out_y: float = OFFSET) -> np.ndarray:
    """Boolean mask over the y nodes of the output flux window."""
    y = np.asarray(sim.grid.coords[1])
    return np.abs(y - out_y) <= WINDOW


def plane_flux(ez: np.ndarray, hy: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """-Re(Ez conj(Hy)) summed over the window, per frequency; ez/hy are
    (n_freq, 1, ny, 1) DFT planes."""
    e = ez.reshape(ez.shape[0], -1)[:, mask]
    h = hy.reshape(hy.shape[0], -1)[:, mask]
    return -np.real(e * np.conj(h)).sum(axis=1)


def reference_flux(pf, steps: int = 4000, precision: str = "float32") -> np.ndarray:
    """Output-window flux of the straight guide (the normalisation)."""
    from .batch import run_batch
    sim = scene(pf, steps, precision, straight=True)
    r = run_batch([sim])[0]
    return plane_flux(r.dft["out"]["Ez"], r.dft["out"]["Hy"], flux_window(sim, 0.0))


class Labeler:
    """Transmission spectra of designs from the batched HIP stepper."""

    def __init__(self, pf, ref_flux: np.ndarray, batch: int = 256, steps: int = 4000,
                 precision: str = "float32", mode: str = "graph", waves: int = 8, rows: int = 16):
        from .batch import SimBatch
        self.sim = scene(pf, steps, precision)
        self.batch = batch
        self.ref = np.asarray(ref_flux)
        self.mask = flux_window(self.sim)
        self.b = SimBatch([self.sim] * batch, mode=mode, waves=waves, rows=rows)
        self.one = SimBatch([self.sim], mode=mode, waves=waves, rows=rows) if batch > 1 else self.b
        self.solves = 0

    def close(self) -> None:
        self.b.close()
        self.one.close()

    def __call__(self, designs: np.ndarray) -> np.ndarray:
        designs = np.asarray(designs, dtype=np.float32).reshape(-1, N_DESIGN, N_DESIGN)
        out = np.empty((designs.shape[0], N_FREQ))
        for c0 in range(0, designs.shape[0], self.batch):
            chunk = designs[c0:c0 + self.batch]
            b = self.one if chunk.shape[0] == 1 else self.b
            for i, rho in enumerate(chunk):
                b.set_eps(i, eps_with(self.sim, rho))
            b.reset()
            b.advance()
            for i in range(chunk.shape[0]):
                d = b.dft(i, "out")
                out[c0 + i] = plane_flux(d["Ez"], d["Hy"], self.mask) / self.ref
            self.solves += chunk.shape[0]
        return out


def objective(spectra: np.ndarray) -> np.ndarray:
    """Mean transmission over the band, per design."""
    return np.asarray(spectra).reshape(-1, N_FREQ).mean(axis=1)


# ---- plain adjoint gradient ascent -------------------------------------------------------

def adjoint_design(pf, ref_flux: np.ndarray, rho0: np.ndarray, iters: int, lr: float = 0.05,
                   steps: int = 4000, precision: str = "float32",
                   target: Optional[float] = None, log: Optional[Callable] = None) -> Dict:
    """Adam ascent on the continuous density with the HIP adjoint gradient;
    each iteration is one forward + one adjoint solve (the objective of the
    design before the step is the verified value of that iteration)."""
    import torch
    from .adjoint import value_and_grad_eps
    sim0 = scene(pf, steps, precision)
    mask = torch.tensor(flux_window(sim0))
    ref = torch.tensor(np.asarray(ref_flux))
    i0, j0 = design_window(sim0)

    def loss(out):
        ez, hy = out["dft"]["out"]["Ez"], out["dft"]["out"]["Hy"]
        e = ez.reshape(ez.shape[0], -1)[:, mask]
        h = hy.reshape(hy.shape[0], -1)[:, mask]
        flux = -(e * torch.conj(h)).real.sum(dim=1)
        return (flux / ref).mean()

    rho = np.array(rho0, dtype=np.float64)
    m = np.zeros_like(rho)
    v = np.zeros_like(rho)
    hist = []
    t0 = time.perf_counter()
    for it in range(iters):
        sim = scene(pf, steps, precision)
        sim.eps_r = eps_with(sim, rho).astype(sim.dtypes["eps_r"])
        val, g = value_and_grad_eps(sim, loss)
        grad = np.asarray(g)[i0:i0 + N_DESIGN, j0:j0 + N_DESIGN, 0] * (EPS_HI - EPS_LO)
        hist.append(dict(iter=it, objective=float(val), wall_s=time.perf_counter() - t0))
        if log:
            log(f"[adjoint] iter {it}: J = {val:.4f} ({hist[-1]['wall_s']:.1f} s)")
        if target is not None and val >= target:
            break
        m = 0.9 * m + 0.1 * grad
        v = 0.999 * v + 0.001 * grad * grad
        mh, vh = m / (1 - 0.9 ** (it + 1)), v / (1 - 0.999 ** (it + 1))
        rho = np.clip(rho + lr * mh / (np.sqrt(vh) + 1e-12), 0.0, 1.0)
    return dict(history=hist, rho=rho)


# ---- surrogate-guided search -----------------------------------------------------------

def train(net, X: np.ndarray, Y: np.ndarray, epochs: int = 10, batch: int = 64,
          lr: float = 1e-3, seed: int = 0, log: Optional[Callable] = None) -> List[float]:
    """Adam on standardised targets (per-frequency mean and deviation of the
    first training set, kept on the network for later fine-tuning)."""
    import torch
    if getattr(net, "y_mu", None) is None:
        net.y_mu = np.asarray(Y, dtype=np.float64).mean(axis=0)
        net.y_sd = np.asarray(Y, dtype=np.float64).std(axis=0) + 1e-12
    rng = np.random.default_rng(seed)
    Z = (np.asarray(Y, dtype=np.float64) - net.y_mu) / net.y_sd
    Xd = torch.tensor(np.asarray(X, dtype=np.float32), device="cuda")
    Zd = torch.tensor(Z.astype(np.float32), device="cuda")
    bs = min(batch, len(X))
    losses = []
    for ep in range(epochs):
        perm = rng.permutation(len(X))
        tot = 0.0
        nb = 0
        for b0 in range(0, len(X) - bs + 1, bs):
            idx = torch.tensor(perm[b0:b0 + bs], device="cuda")
            net.zero_grad()
            tot += float(net.mse_step(Xd[idx], Zd[idx]))
            net.adam(lr)
            nb += 1
        losses.append(tot / max(nb, 1))
        if log:
            log(f"[train] epoch {ep}: mse {losses[-1]:.3e} (standardised)")
    return losses


def predict(net, X: np.ndarray, batch: int = 256) -> np.ndarray:
    import torch
    out = []
    for b0 in range(0, len(X), batch):
        x = torch.tensor(np.asarray(X[b0:b0 + batch], dtype=np.float32), device="cuda")
        out.append(net.forward(x).cpu().numpy())
    return np.concatenate(out) * net.y_sd + net.y_mu


def propose_from(net, starts: np.ndarray, iters: int = 30, step: float = 0.02,
                 snapshots: int = 3) -> np.ndarray:
    """Candidates from normalised gradient ascent on the network's predicted
    objective, one path per starting design (one batched forward/backward
    per iteration); the iterates at `snapshots` evenly spaced iterations of
    every path are the candidates, in order (snapshot, start)."""
    import torch
    x = np.asarray(starts, dtype=np.float32).reshape(-1, N_DESIGN, N_DESIGN).copy()
    dy = torch.tensor(np.tile(net.y_sd / N_FREQ, (len(x), 1)).astype(np.float32), device="cuda")
    keep = {max(1, round(iters * (q + 1) / snapshots)) for q in range(snapshots)}
    out = []
    for it in range(1, iters + 1):
        net.forward(torch.tensor(x, device="cuda"))
        net.zero_grad()
        g = net.backward(dy, input_grad=True).cpu().numpy()
        scale = np.abs(g).reshape(len(x), -1).max(axis=1)[:, None, None] + 1e-20
        x = np.clip(x + step * g / scale, 0, 1).astype(np.float32)
        if it in keep:
            out.append(x.copy())
    net.zero_grad()
    return np.concatenate(out)


def surrogate_design(pf, net, labeler: Labeler, ref_flux: np.ndarray, rho0: np.ndarray,
                     rounds: int, target: Optional[float] = None, candidates: int = 48,
                     pool: Optional[tuple] = None, iters: int = 30, step: float = 0.02,
                     finetune_epochs: int = 3, tol: float = 0.02, seed: int = 0,
                     steps: int = 4000, precision: str = "float32",
                     log: Optional[Callable] = None) -> Dict:
    """Rounds of: start the network's gradient ascent from the best verified
    designs known so far (the labelled training set, if given as `pool` =
    (designs, spectra), plus every design verified in earlier rounds), verify
    the candidates with one batched FDTD run, keep the best verified design
    and fine-tune the network on the verified batch. If the network's
    relative error on the kept design exceeds `tol`, also take one plain
    adjoint step from it. A round whose verified designs reach `target`
    ends the loop right after the verification, and one whose adjoint step
    reaches it ends before the fine-tuning (both only serve later rounds)."""
    best = np.array(rho0, dtype=np.float32)
    best_j = float(objective(labeler(best[None]))[0])
    if pool is not None:
        P = np.asarray(pool[0], dtype=np.float32).reshape(-1, N_DESIGN, N_DESIGN)
        PJ = objective(pool[1]).astype(np.float64)
    else:
        P = np.empty((0, N_DESIGN, N_DESIGN), np.float32)
        PJ = np.empty(0)
    P = np.concatenate([P, best[None]])
    PJ = np.concatenate([PJ, [best_j]])
    n_start = max(1, candidates // 3)
    hist = [dict(round=-1, objective=best_j, wall_s=0.0, fallback=False)]
    t0 = time.perf_counter()
    for r in range(rounds):
        top = np.argsort(PJ)[::-1][:n_start]
        cand = propose_from(net, P[top], iters=iters, step=step)[:candidates]
        pred = objective(predict(net, cand))
        spec = labeler(cand)
        ver = objective(spec)
        P = np.concatenate([P, cand])
        PJ = np.concatenate([PJ, ver])
        k = int(np.argmax(ver))
        err = abs(pred[k] - ver[k]) / max(abs(ver[k]), 1e-12)
        fallback = err > tol
        if ver[k] > best_j:
            best, best_j = cand[k], float(ver[k])
        if target is not None and best_j >= target:
            hist.append(dict(round=r, objective=best_j, predicted=float(pred[k]), error=float(err),
                             fallback=False, wall_s=time.perf_counter() - t0))
            if log:
                log(f"[surrogate] round {r}: best J = {best_j:.4f} (candidate pred {pred[k]:.4f} vs "
                    f"FDTD {ver[k]:.4f}, err {err:.1%}); target reached; {hist[-1]['wall_s']:.1f} s")
            break
        if fallback:
            step_a = adjoint_design(pf, ref_flux, best, iters=1, steps=steps, precision=precision)
            rho_a = step_a["rho"].astype(np.float32)
            ja = float(objective(labeler(rho_a[None]))[0])
            P = np.concatenate([P, rho_a[None]])
            PJ = np.concatenate([PJ, [ja]])
            if ja > best_j:
                best, best_j = rho_a, ja
        if target is None or best_j < target:
            train(net, cand, spec, epochs=finetune_epochs, batch=min(64, len(cand)), lr=5e-4,
                  seed=seed + r)
        hist.append(dict(round=r, objective=best_j, predicted=float(pred[k]), error=float(err),
                         fallback=bool(fallback), wall_s=time.perf_counter() - t0))
        if log:
            log(f"[surrogate] round {r}: best J = {best_j:.4f} (candidate pred {pred[k]:.4f} vs "
                f"FDTD {ver[k]:.4f}, err {err:.1%}{', adjoint fallback' if fallback else ''}; "
                f"{hist[-1]['wall_s']:.1f} s)")
        if target is not None and best_j >= target:
            break
    return dict(history=hist, rho=best, objective=best_j)
