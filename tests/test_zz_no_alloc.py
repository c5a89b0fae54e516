"""GPU: pf_nn_conv_backward_noalloc (weight-gradient partials in the
library's static scratch array, no allocation) gives the same bits as
pf_nn_conv_backward (stream-ordered workspace) through the C ABI: two calls
in a row, the second on a smaller layer, then the first shape again (the
scratch keeps the previous call's contents; every partial the reduction
reads is rewritten by the call), and a layer too large for the scratch
array is rejected with PF_E_LIMIT before anything is launched."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

PF_E_LIMIT = -4


def _layer(n, h, w, ci, co, k, s, relu, out_f32, seed):
    from photonfdtd_amd import _native as N
    dev = torch.device("cuda")
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(n, h, w, ci, generator=g).to(torch.bfloat16).to(dev)
    wt = (torch.randn(co, k, k, ci, generator=g) / np.sqrt(k * k * ci)).to(torch.bfloat16).to(dev)
    b = (torch.randn(co, generator=g) * 0.1).to(dev)
    ho, wo = (h - 1) // s + 1, (w - 1) // s + 1
    dy = torch.randn(n, ho, wo, co, generator=g).to(dev)
    y = torch.empty(n, ho, wo, co, dtype=torch.float32 if out_f32 else torch.bfloat16, device=dev)
    d = N.PfNnConv()
    d.n, d.h, d.w, d.cin, d.cout, d.ksize, d.stride, d.relu, d.out_f32 = (
        n, h, w, ci, co, k, s, int(relu), int(out_f32))
    return d, x, wt, b, y, dy


def _backward(lib, fn, d, x, wt, y, dy):
    dev = x.device
    dx = torch.full((d.n, d.h, d.w, d.cin), float("nan"), device=dev)
    dw = torch.zeros(d.cout, d.ksize, d.ksize, d.cin, device=dev)
    db = torch.zeros(d.cout, device=dev)
    p = lambda t: C.c_void_p(t.data_ptr())
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = getattr(lib, fn)(C.byref(d), p(x), p(wt), p(y), p(dy), p(dx), p(dw), p(db), st)
    assert rc == 0, (fn, rc, lib.pf_nn_last_error())
    torch.cuda.synchronize()
    return dx.cpu(), dw.cpu(), db.cpu()


def test_noalloc_matches_allocating_backward():
    from photonfdtd_amd import _native as N
    lib = N.load_library()
    shapes = [(32, 64, 64, 16, 32, 3, 1, True, False, 1),  # 32 pixel splits of the dW GEMM
              (2, 16, 16, 16, 32, 3, 1, True, True, 2),    # smaller layer (one split), FP32 output
              (32, 64, 64, 16, 32, 3, 1, True, False, 1)]  # the first shape again
    for shape in shapes:
        print(f"\n[case] noalloc conv backward {shape}", flush=True)
        d, x, wt, b, y, dy = _layer(*shape)
        p = lambda t: C.c_void_p(t.data_ptr())
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        assert lib.pf_nn_conv_forward(C.byref(d), p(x), p(wt), p(b), p(y), st) == 0
        want = _backward(lib, "pf_nn_conv_backward", d, x, wt, y, dy)
        got = _backward(lib, "pf_nn_conv_backward_noalloc", d, x, wt, y, dy)
        for name, a, bb in zip(("dx", "dw", "db"), got, want):
            assert torch.equal(a, bb), name
            assert torch.isfinite(a).all() and a.abs().max() > 0, name


def test_noalloc_rejects_layer_beyond_scratch():
    from photonfdtd_amd import _native as N
    lib = N.load_library()
    print("\n[case] noalloc conv backward over the scratch limit", flush=True)
    # cout * 9 * cin = 1024 * 9 * 1024 floats > 4 Mi: rejected on the host
    d, x, wt, b, y, dy = _layer(1, 4, 4, 1024, 1024, 3, 1, False, True, 3)
    p = lambda t: C.c_void_p(t.data_ptr())
    dw = torch.zeros(1024, 3, 3, 1024, device=x.device)
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.pf_nn_conv_backward_noalloc(C.byref(d), p(x), p(wt), p(y), p(dy), None, p(dw), None, st)
    assert rc == PF_E_LIMIT
    assert b"scratch" in lib.pf_nn_last_error()
    torch.cuda.synchronize()
