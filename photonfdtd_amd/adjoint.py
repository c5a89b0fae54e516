"""Gradients of monitor losses w.r.t. permittivity through the HIP stepper.

Mirrors the reference's differentiable entry point
``value_and_grad_eps(sim, loss)`` (``src/photonfdtd/jaxbackend.py:666-722``):
``loss`` maps the monitor outputs ``{'dft': {name: {comp: complex}}}`` to a
scalar, and the gradient w.r.t. the per-node ``eps_r`` flows through
``ce = dt / (eps_r * EPS_0)`` and the whole time evolution. Dispersive media and
subpixel smoothing are rejected, as in the reference.

The forward pass is the HIP stepper with field checkpoints every
``checkpoint_interval`` steps kept on the device; the backward pass is the
hand-written discrete adjoint (``pf_adjoint``): time-reversed transposed Yee +
CPML steps with the DFT cotangents as sources, each checkpoint segment
recomputed once. ``FDTDFunction`` wraps both as a ``torch.autograd.Function``
so the loss can be any torch expression of the DFT tensors. With ``eps`` on
the GPU the DFT outputs, their cotangents and the gradient stay on the device
(``pf_dft_ptr``, device cotangents in ``pf_adjoint``); with ``eps`` on the CPU
they are host arrays, as in the reference.
"""
from __future__ import annotations

import copy
from typing import Callable, Dict, List, Tuple

import numpy as np

from . import _native as N
from .constants import EPS_0
from .hip_backend import HipEngine, exact_mode, run_loop, validate
from .monitors import DFTMonitor


def _require_torch():
    try:
        import torch
    except ImportError as e:                                   # pragma: no cover
        raise RuntimeError("the adjoint entry points need PyTorch (autograd plumbing)") from e
    return torch


def _check_sim(sim) -> None:
    if sim.subpixel or getattr(sim, "_has_dispersion", False):
        raise NotImplementedError(
            "value_and_grad_eps differentiates the scalar per-node eps_r; subpixel-smoothed "
            "tensors and dispersive pole models are not parameterised by it.")


def _dft_keys(sim) -> List[Tuple[str, str]]:
    return [(m.name, c) for m in sim.monitors if isinstance(m, DFTMonitor) for c in m.components]


def _with_eps(sim, eps: np.ndarray):
    """A copy of ``sim`` whose permittivity grid is ``eps`` (same shape)."""
    s = copy.copy(sim)
    s.monitors = list(sim.monitors)
    s.eps_r = np.asarray(eps, dtype=sim.dtypes["eps_r"]).reshape(sim.grid.shape)
    return s


def _checkpointed_engine(run, npdt, interval):
    """A stepper that stores checkpoints for the adjoint: the 64-lane march on
    dense z rows when nz is a multiple of 64 (its fields are then in the
    generic layout the reverse sweep reads), else the generic kernels."""
    if run.grid.shape[2] % 64 == 0:
        eng = HipEngine(run, npdt, exact_mode(npdt), flags=N.PF_FLAG_DENSE64)
        try:
            eng.enable_checkpoints(interval)
            return eng
        except (RuntimeError, ValueError):
            eng.close()    # the stepper did not end up in that layout
    eng = HipEngine(run, npdt, exact_mode(npdt), flags=N.PF_FLAG_GENERIC)
    eng.enable_checkpoints(interval)
    return eng


def _make_function():
    torch = _require_torch()

    class FDTDFunction(torch.autograd.Function):
        """eps (nx, ny, nz) float64 -> the DFT accumulators of ``sim``, in
        ``_dft_keys(sim)`` order, as complex128 tensors."""

        @staticmethod
        def forward(ctx, eps, sim, checkpoint_interval):
            _check_sim(sim)
            run = _with_eps(sim, eps.detach().cpu().numpy())
            npdt = validate(run)
            eng = _checkpointed_engine(run, npdt, checkpoint_interval)
            res = run_loop(run, eng)
            ctx.engine = eng
            ctx.keys = _dft_keys(sim)
            ctx.save_for_backward(eps)
            ctx.dt = run.dt
            if eps.is_cuda:
                # device handoff: a device copy of the accumulators (the
                # engine's own memory stays with the engine until backward)
                dev = {n: eng.dft_tensors(n) for n in {k for k, _ in ctx.keys}}
                return tuple(dev[n][c].clone() for n, c in ctx.keys)
            return tuple(torch.from_numpy(np.ascontiguousarray(res.dft[n][c]))
                         for n, c in ctx.keys)

        @staticmethod
        def backward(ctx, *grads):
            (eps,) = ctx.saved_tensors
            eng = ctx.engine
            if eps.is_cuda:
                cot_d: Dict[str, Dict] = {}
                for (n, c), g in zip(ctx.keys, grads):
                    if g is not None:
                        cot_d.setdefault(n, {})[c] = g
                grad_ce = eng.adjoint_device(cot_d)
                eng.close()
                e = eps.detach().to(torch.float64)
                grad_eps = grad_ce * (-ctx.dt / (e * e * EPS_0))
                return grad_eps.to(eps.dtype), None, None
            cot: Dict[str, Dict[str, np.ndarray]] = {}
            for (n, c), g in zip(ctx.keys, grads):
                if g is None:
                    continue
                cot.setdefault(n, {})[c] = g.detach().cpu().numpy()
            grad_ce = eng.adjoint(cot)
            eng.close()
            e = eps.detach().cpu().numpy()
            # ce = dt / (eps * EPS_0)  ->  dce/deps = -dt / (eps^2 EPS_0)
            grad_eps = grad_ce * (-ctx.dt / (e * e * EPS_0))
            return torch.from_numpy(grad_eps).to(eps.dtype), None, None

    return FDTDFunction


_FUNC = None


def FDTDFunction():
    """The ``torch.autograd.Function`` class (built on first use)."""
    global _FUNC
    if _FUNC is None:
        _FUNC = _make_function()
    return _FUNC


def dft_outputs(sim, eps, checkpoint_interval: int = 100):
    """Differentiable DFT outputs ``{name: {comp: complex tensor}}`` for a torch
    ``eps`` tensor of the grid's shape, on ``eps``'s device."""
    outs = FDTDFunction().apply(eps, sim, int(checkpoint_interval))
    d: Dict[str, Dict] = {}
    for (n, c), t in zip(_dft_keys(sim), outs):
        d.setdefault(n, {})[c] = t
    return d


def value_and_grad_eps(sim, loss: Callable, checkpoint_interval: int = 100,
                       device: str = "cpu"):
    """``(value, d value / d eps_r)`` for a scalar ``loss(out)``, ``out`` =
    ``{'dft': {name: {comp: complex tensor}}}`` (``jaxbackend.py:666-722``).
    ``device="cuda"`` hands the DFT tensors to ``loss`` on the GPU (no host
    copies of the outputs, cotangents or gradient until the returned array)."""
    torch = _require_torch()
    _check_sim(sim)
    eps = torch.tensor(np.asarray(sim.eps_r, dtype=np.float64), device=device,
                       requires_grad=True)
    val = loss({"dft": dft_outputs(sim, eps, checkpoint_interval)})
    val.backward()
    return float(val.detach()), eps.grad.detach().cpu().numpy()
