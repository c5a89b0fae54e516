"""Learned surrogate for 2-D inverse design: design map -> transmission spectrum.

The reference has no learned component; this module is the surrogate named
in the tier brief, built on the HIP stepper: the training labels come from
the batched gfx950 FDTD stepper (``photonfdtd_amd.batch``), the network runs
on hand-written MFMA kernels (``csrc/nn_mfma.hip``: BF16 operands, FP32
accumulation, FP32 master weights under Adam), and every design it proposes
is verified by a full FDTD solve, with the plain adjoint step as fallback.

Network: a strided 3x3 convolution encoder over the 128 x 128 design map
(channels 32-64-128-256-128, resolution 128 -> 4) whose 4 x 4 x 128 output is
flattened (it keeps where the material is, which decides the coupling) into a
two-layer spectral head (2048 -> 256 -> n_freq), about 2.0 M parameters. In
NHWC order the flatten is a reshape, so the head layers are 1 x 1
convolutions on a 1 x 1 map.
PyTorch provides device memory and the stream only; every layer's forward
and backward pass and the optimizer run in this package's kernels.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N

# (cin, cout, ksize, stride, relu) of the encoder; its output is flattened
# into the 1x1 head layers.
ENCODER: Tuple[Tuple[int, int, int, int, int], ...] = (
    (1, 32, 3, 1, 1), (32, 32, 3, 2, 1),
    (32, 64, 3, 1, 1), (64, 64, 3, 2, 1),
    (64, 128, 3, 1, 1), (128, 128, 3, 2, 1),
    (128, 256, 3, 1, 1), (256, 256, 3, 2, 1),
    (256, 128, 3, 2, 1),
)
FLAT = 4 * 4 * 128       # encoder output at a 128 x 128 input
HEAD_HIDDEN = 256


def _torch():
    try:
        import torch
    except ImportError as e:                                   # pragma: no cover
        raise RuntimeError("the surrogate needs PyTorch for device memory") from e
    return torch


def layer_table(n_out: int = 16) -> List[Tuple[int, int, int, int, int]]:
    return list(ENCODER) + [(FLAT, HEAD_HIDDEN, 1, 1, 1), (HEAD_HIDDEN, n_out, 1, 1, 0)]


def n_params(n_out: int = 16) -> int:
    return sum(co * k * k * ci + co for ci, co, k, _, _ in layer_table(n_out))


def _nn_check(lib, status: int, what: str) -> None:
    if status != 0:
        raise RuntimeError(f"{what}: {lib.pf_nn_last_error().decode(errors='replace')}")


class SpectralSurrogate:
    """The network, its parameters (FP32 master + BF16 copy), Adam state and
    gradient buffers on one GPU."""

    def __init__(self, n_out: int = 16, in_size: int = 128, seed: int = 0, device: str = "cuda"):
        torch = _torch()
        self.torch = torch
        self.lib = N.load_library()
        self.dev = torch.device(device)
        if self.dev.type != "cuda":
            raise RuntimeError("the surrogate kernels run on the GPU only")
        self.n_out, self.in_size = n_out, in_size
        self.layers = layer_table(n_out)
        gen = torch.Generator().manual_seed(seed)
        self.p: List[Dict] = []
        for ci, co, k, _, _ in self.layers:
            std = math.sqrt(2.0 / (k * k * ci))
            w = (torch.randn((co, k, k, ci), generator=gen) * std).to(self.dev)
            b = torch.zeros(co, device=self.dev)
            self.p.append(dict(w=w, b=b, wb=w.to(torch.bfloat16), gw=torch.zeros_like(w),
                               gb=torch.zeros_like(b), mw=torch.zeros_like(w),
                               vw=torch.zeros_like(w), mb=torch.zeros_like(b),
                               vb=torch.zeros_like(b)))
        self.step_count = 0
        self._acts: List = []
        self._shapes: List[Tuple[int, int, int]] = []

    # ---- plumbing ----
    def _stream(self):
        return C.c_void_p(self.torch.cuda.current_stream(self.dev).cuda_stream)

    @staticmethod
    def _ptr(t):
        return C.c_void_p(t.data_ptr()) if t is not None else None

    def _desc(self, n, h, w, li):
        ci, co, k, s, relu = self.layers[li]
        d = N.PfNnConv()
        d.n, d.h, d.w, d.cin, d.cout = n, h, w, ci, co
        d.ksize, d.stride, d.relu = k, s, relu
        d.out_f32 = 1 if li == len(self.layers) - 1 else 0
        return d

    def n_params(self) -> int:
        return sum(q["w"].numel() + q["b"].numel() for q in self.p)

    # ---- passes ----
    def forward(self, x):
        """x: (N, H, W) float32 design maps on the device -> (N, n_out) float32."""
        torch = self.torch
        n, h, w = x.shape
        s = self._stream()
        cur = torch.empty((n, h, w, 1), dtype=torch.bfloat16, device=self.dev)
        _nn_check(self.lib, self.lib.pf_nn_to_bf16(x.numel(), self._ptr(x.contiguous()),
                                                   self._ptr(cur), s), "pf_nn_to_bf16")
        self._acts, self._shapes = [cur], [(h, w, 1)]
        n_enc = len(ENCODER)
        for li, (ci, co, k, st, relu) in enumerate(self.layers):
            if li == n_enc:                                    # flatten (NHWC: a reshape)
                if cur[0].numel() != ci:
                    raise ValueError(f"encoder output {tuple(cur.shape[1:])} does not flatten "
                                     f"to {ci} features; the input must be {self.in_size}^2")
                self._enc_shape = tuple(cur.shape)
                cur, h, w = cur.reshape(n, 1, 1, ci), 1, 1
                self._acts.append(cur)
                self._shapes.append((1, 1, ci))
            d = self._desc(n, h, w, li)
            ho, wo = (h - 1) // st + 1, (w - 1) // st + 1
            last = li == len(self.layers) - 1
            out = torch.empty((n, ho, wo, co), dtype=torch.float32 if last else torch.bfloat16,
                              device=self.dev)
            q = self.p[li]
            _nn_check(self.lib, self.lib.pf_nn_conv_forward(C.byref(d), self._ptr(cur),
                                                            self._ptr(q["wb"]), self._ptr(q["b"]),
                                                            self._ptr(out), s),
                      "pf_nn_conv_forward")
            cur, h, w = out, ho, wo
            self._acts.append(cur)
            self._shapes.append((h, w, co))
        return cur.view(n, -1)

    def backward(self, dy, input_grad: bool = False):
        """Accumulate parameter gradients for dL/dy = dy (N, n_out) float32;
        returns dL/dx (N, H, W) when input_grad is set."""
        torch = self.torch
        s = self._stream()
        n = dy.shape[0]
        n_enc = len(ENCODER)
        g = dy.contiguous().view(*self._acts[-1].shape)
        # activation index of the input of layer li: acts[li] before the
        # flatten, acts[li + 1] after it (the flattened view is inserted)
        for li in range(len(self.layers) - 1, -1, -1):
            ai = li if li < n_enc else li + 1
            x_in, y_out = self._acts[ai], self._acts[ai + 1]
            h, w, _ = self._shapes[ai]
            d = self._desc(n, h, w, li)
            q = self.p[li]
            need_dx = li > 0 or input_grad
            dx = torch.empty(x_in.shape, dtype=torch.float32, device=self.dev) if need_dx else None
            _nn_check(self.lib, self.lib.pf_nn_conv_backward(
                C.byref(d), self._ptr(x_in), self._ptr(q["wb"]), self._ptr(y_out), self._ptr(g),
                self._ptr(dx), self._ptr(q["gw"]), self._ptr(q["gb"]), s), "pf_nn_conv_backward")
            if li == n_enc:                                    # through the flatten
                dx = dx.reshape(self._enc_shape)
            g = dx
        return g.view(n, self._shapes[0][0], self._shapes[0][1]) if input_grad else None

    def mse_step(self, x, target):
        """Forward, MSE loss and backward (gradients accumulate)."""
        torch = self.torch
        pred = self.forward(x)
        dpred = torch.empty_like(pred)
        loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        _nn_check(self.lib, self.lib.pf_nn_mse(pred.numel(), self._ptr(pred),
                                               self._ptr(target.contiguous()), self._ptr(dpred),
                                               self._ptr(loss), self._stream()), "pf_nn_mse")
        self.backward(dpred)
        return loss

    def adam(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
             eps: float = 1e-8) -> None:
        self.step_count += 1
        s = self._stream()
        for q in self.p:
            for pw, gw, m, v, pb in ((q["w"], q["gw"], q["mw"], q["vw"], q["wb"]),
                                     (q["b"], q["gb"], q["mb"], q["vb"], None)):
                _nn_check(self.lib, self.lib.pf_nn_adam(pw.numel(), self._ptr(pw), self._ptr(gw),
                                                        self._ptr(m), self._ptr(v), self._ptr(pb),
                                                        lr, beta1, beta2, eps, self.step_count, s),
                          "pf_nn_adam")

    def zero_grad(self) -> None:
        for q in self.p:
            q["gw"].zero_()
            q["gb"].zero_()

    def state(self) -> Dict[str, np.ndarray]:
        """FP32 master weights on the host (w as [cout, k, k, cin])."""
        out = {}
        for i, q in enumerate(self.p):
            out[f"w{i}"] = q["w"].detach().cpu().numpy()
            out[f"b{i}"] = q["b"].detach().cpu().numpy()
        return out
