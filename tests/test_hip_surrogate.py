"""GPU: the surrogate's MFMA kernels against the CPU restatement.

Forward outputs and the data / weight / bias gradients of single layers and
of the whole network are compared with PyTorch float32 on the CPU
(``oracle/surrogate_torch.py``) with the same BF16 roundings, so the bound
covers FP32 summation order and the BF16 rounding of the gradient operand.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    return float((a - b).norm() / b.norm().clamp_min(1e-30))


def run_layer(n, h, w, ci, co, k, s, relu, out_f32, seed=0):
    import ctypes as C
    from oracle.surrogate_torch import bf16, conv_layer
    from photonfdtd_amd import _native as N
    lib = N.load_library()
    g = torch.Generator().manual_seed(seed)
    x = bf16(torch.randn(n, h, w, ci, generator=g))
    wt = bf16(torch.randn(co, k, k, ci, generator=g) / np.sqrt(k * k * ci))
    b = torch.randn(co, generator=g) * 0.1
    ho, wo = (h - 1) // s + 1, (w - 1) // s + 1
    dy = torch.randn(n, ho, wo, co, generator=g)
    dev = torch.device("cuda")
    xd, wd = x.to(torch.bfloat16).to(dev), wt.to(torch.bfloat16).to(dev)
    bd, dyd = b.to(dev), dy.to(dev)
    yd = torch.empty(n, ho, wo, co, dtype=torch.float32 if out_f32 else torch.bfloat16, device=dev)
    dxd = torch.empty(n, h, w, ci, device=dev)
    dwd = torch.zeros(co, k, k, ci, device=dev)
    dbd = torch.zeros(co, device=dev)
    d = N.PfNnConv()
    d.n, d.h, d.w, d.cin, d.cout, d.ksize, d.stride, d.relu, d.out_f32 = (
        n, h, w, ci, co, k, s, int(relu), int(out_f32))
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: C.c_void_p(t.data_ptr())
    assert lib.pf_nn_conv_forward(C.byref(d), p(xd), p(wd), p(bd), p(yd), st) == 0
    assert lib.pf_nn_conv_backward(C.byref(d), p(xd), p(wd), p(yd), p(dyd), p(dxd), p(dwd),
                                   p(dbd), st) == 0
    torch.cuda.synchronize()
    # CPU reference
    y_ref = conv_layer(x, wt, b, s, relu, out_f32)
    xr = x.clone().requires_grad_(True)
    wr = wt.clone().requires_grad_(True)
    br = b.clone().requires_grad_(True)
    pre = torch.nn.functional.conv2d(xr.permute(0, 3, 1, 2), wr.permute(0, 3, 1, 2), br,
                                     stride=s, padding=k // 2).permute(0, 2, 3, 1)
    mask = (yd.float().cpu() > 0).float() if relu else torch.ones_like(dy)
    dz = bf16(dy * mask)
    (pre * dz).sum().backward()
    return dict(y=rel(yd.float(), y_ref), dx=rel(dxd, xr.grad), dw=rel(dwd, wr.grad),
                db=rel(dbd, (dy * mask).sum(dim=(0, 1, 2))))


@pytest.mark.parametrize("cfg", [
    (2, 20, 18, 1, 32, 3, 1, True, False),
    (2, 17, 16, 32, 64, 3, 2, True, False),
    (3, 9, 9, 64, 40, 3, 1, False, True),
    (5, 1, 1, 256, 512, 1, 1, True, False),
    (5, 1, 1, 512, 16, 1, 1, False, True),
    (3, 15, 13, 1, 64, 3, 2, True, False),    # one input channel, stride 2
    (2, 9, 8, 16, 24, 1, 2, True, False),     # 1x1 stride 2: odd pixels get no gradient
    (4, 12, 12, 128, 256, 3, 2, True, False), # 128-wide tiles
])
def test_layer_matches_cpu(cfg):
    print(f"\n[case] layer {cfg}", flush=True)
    e = run_layer(*cfg)
    print(f"[layer] {e}", flush=True)
    assert e["y"] < 8e-3 and e["dx"] < 1e-5 and e["dw"] < 1e-5 and e["db"] < 1e-5


def test_network_matches_cpu():
    from oracle.surrogate_torch import bf16, network
    from photonfdtd_amd.surrogate import ENCODER, SpectralSurrogate
    print("\n[case] full network fwd/bwd", flush=True)
    net = SpectralSurrogate(seed=3)
    g = torch.Generator().manual_seed(1)
    x = (torch.rand(4, 128, 128, generator=g) > 0.5).float()
    y = net.forward(x.cuda())
    dy = torch.randn(4, 16, generator=g)
    net.zero_grad()
    dx = net.backward(dy.cuda(), input_grad=True)
    torch.cuda.synchronize()
    params = [(bf16(q["w"].cpu()).requires_grad_(True), q["b"].cpu().clone().requires_grad_(True))
              for q in net.p]
    xr = x.clone().requires_grad_(True)
    y_ref = network(xr, params, net.layers, len(ENCODER))
    ey = rel(y, y_ref)
    (y_ref * dy).sum().backward()
    ew = max(rel(q["gw"], pw.grad) for q, (pw, _) in zip(net.p, params))
    eb = max(rel(q["gb"], pb.grad) for q, (_, pb) in zip(net.p, params))
    ex = rel(dx, xr.grad)
    print(f"[network] y {ey:.2e} dW {ew:.2e} db {eb:.2e} dx {ex:.2e}", flush=True)
    assert ey < 2e-2 and ew < 5e-2 and eb < 5e-2 and ex < 5e-2


def test_training_reduces_loss():
    from photonfdtd_amd.surrogate import SpectralSurrogate
    print("\n[case] training on a synthetic map -> spectrum task", flush=True)
    g = torch.Generator().manual_seed(0)
    x = (torch.rand(64, 128, 128, generator=g) > 0.5).float()
    # targets: a smooth function of quadrant fill fractions
    q = torch.stack([x[:, :64, :64].mean((1, 2)), x[:, 64:, :64].mean((1, 2)),
                     x[:, :64, 64:].mean((1, 2)), x[:, 64:, 64:].mean((1, 2))], 1)
    f = torch.linspace(0, 1, 16)
    t = (q[:, :1] * torch.cos(3 * f) + q[:, 1:2] * f + q[:, 2:3] * (1 - f) ** 2
         + q[:, 3:4] * torch.sin(2 * f))
    net = SpectralSurrogate(seed=0)
    xd, td = x.cuda(), t.cuda()
    losses = []
    for it in range(150):
        net.zero_grad()
        loss = net.mse_step(xd, td)
        net.adam(lr=1e-3)
        losses.append(float(loss))
    print(f"[train] loss {losses[0]:.4e} -> {losses[-1]:.4e}", flush=True)
    assert np.isfinite(losses).all() and losses[-1] < 0.1 * losses[0]
