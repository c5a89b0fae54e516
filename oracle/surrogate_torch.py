"""CPU restatement of the spectral surrogate network (test infrastructure).

The checker for the hand-written MFMA kernels of
``photonfdtd_amd/csrc/nn_mfma.hip``: the same layers in PyTorch float32 on the
CPU, with the device path's BF16 roundings applied at the same places
(inputs, weights and every stored activation; the data-gradient GEMM's dZ
operand), so the only remaining difference is the FP32 summation order.
The reference has no learned component, so nothing here cites it; the layer
table is the one in ``photonfdtd_amd/surrogate.py``.
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import torch


def bf16(t: torch.Tensor) -> torch.Tensor:
    """Round to bfloat16 (nearest even) and back to float32."""
    return t.to(torch.bfloat16).to(torch.float32)


def conv_layer(x_nhwc, w_okkc, b, stride: int, relu: bool, out_f32: bool):
    """One layer as the device computes it: BF16 operands, FP32 accumulate,
    bias, ReLU, output rounded to BF16 unless kept in FP32."""
    k = w_okkc.shape[1]
    x = x_nhwc.permute(0, 3, 1, 2)
    w = w_okkc.permute(0, 3, 1, 2)
    y = torch.nn.functional.conv2d(x, w, b, stride=stride, padding=k // 2)
    if relu:
        y = torch.relu(y)
    y = y.permute(0, 2, 3, 1)
    return y if out_f32 else bf16(y)


def network(x_nhw, params: Sequence[Tuple[torch.Tensor, torch.Tensor]],
            layers: Sequence[Tuple[int, int, int, int, int]], n_encoder: int):
    """Forward pass; params are (w [cout, k, k, cin] BF16-valued float32, b)."""
    cur = bf16(x_nhw.unsqueeze(-1))
    for li, ((ci, co, k, s, relu), (w, b)) in enumerate(zip(layers, params)):
        if li == n_encoder:                      # flatten NHWC into the 1x1 head
            cur = cur.reshape(cur.shape[0], 1, 1, -1)
        cur = conv_layer(cur, w, b, s, bool(relu), li == len(layers) - 1)
    return cur.reshape(cur.shape[0], -1)
