"""Mode amplitudes on an x-normal port plane (reference ``smatrix.py:41-118``),
pinned by what the inner product fixes: a field a*m+ + b*m- decomposes into
(a, b) for any complex a, b (this needs <m+, m-> = 0 and the self-overlap
normalisation), the result does not depend on the mode's scale or on dA, and
the closed form of <m, m> for a real mode is 2 dA sum (Ey Hz - Ez Hy).
Mode profiles are real up to one global phase, as a lossless guide's are."""
import numpy as np
import pytest

from photonfdtd_amd.smatrix import (PORT_COMPONENTS, PortField, inner, mode_amplitudes,
                                    port_fields, s_parameters)


def _mode(rng, shape=(1, 7, 5)):
    """A guided-mode-like profile: real transverse fields times one global
    phase (E* x H is then real, which makes the two directions orthogonal)."""
    phase = np.exp(1j * rng.uniform(0, 2 * np.pi))
    return {c: phase * rng.standard_normal(shape) for c in PORT_COMPONENTS}


def _combo(mode, a, b):
    """Fields (n_freq, plane) equal to a_f m+ + b_f m- per frequency."""
    out = {}
    for c in PORT_COMPONENTS:
        sign = -1.0 if c.startswith("H") else 1.0
        out[c] = a[:, None, None, None] * mode[c][None] + sign * b[:, None, None, None] * mode[c][None]
    return out


def test_decomposition_recovers_amplitudes():
    rng = np.random.default_rng(3)
    mode = _mode(rng)
    a = rng.standard_normal(4) + 1j * rng.standard_normal(4)
    b = rng.standard_normal(4) + 1j * rng.standard_normal(4)
    ap, am = mode_amplitudes(_combo(mode, a, b), mode, dA=0.37)
    np.testing.assert_allclose(ap, a, rtol=1e-12)
    np.testing.assert_allclose(am, b, rtol=1e-12)


def test_directions_orthogonal_and_self_overlap_closed_form():
    rng = np.random.default_rng(4)
    real = {c: rng.standard_normal((1, 6, 4)) for c in PORT_COMPONENTS}
    m = PortField.from_dict(real, keep_leading=False)
    assert abs(inner(m, m.reversed(), 2.0)) < 1e-12
    want = 2 * 2.0 * np.sum(real["Ey"] * real["Hz"] - real["Ez"] * real["Hy"])
    assert inner(m, m, 2.0) == pytest.approx(want, rel=1e-12)


def test_scale_and_area_invariance():
    rng = np.random.default_rng(5)
    mode = _mode(rng)
    f = {c: rng.standard_normal((3, 1, 7, 5)) + 1j * rng.standard_normal((3, 1, 7, 5))
         for c in PORT_COMPONENTS}
    a1, b1 = mode_amplitudes(f, mode, 1.0)
    scaled = {c: (2.5 - 1.5j) * v for c, v in mode.items()}
    a2, b2 = mode_amplitudes(f, scaled, 7.0)
    # <s m, f> / <s m, s m> = conj(s) <m, f> / |s|^2 <m, m> = alpha / s
    np.testing.assert_allclose(a2 * (2.5 - 1.5j), a1, rtol=1e-12)
    np.testing.assert_allclose(b2 * (2.5 - 1.5j), b1, rtol=1e-12)


def test_torch_tensors_and_gradient():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(6)
    mode = _mode(rng)
    f = {c: torch.tensor(rng.standard_normal((2, 1, 7, 5)) + 1j * rng.standard_normal((2, 1, 7, 5)),
                         requires_grad=True) for c in PORT_COMPONENTS}
    ap, _ = mode_amplitudes(f, {c: torch.tensor(v) for c, v in mode.items()}, 1.0)
    ref, _ = mode_amplitudes({c: v.detach().numpy() for c, v in f.items()}, mode, 1.0)
    np.testing.assert_allclose(ap.detach().numpy(), ref, rtol=1e-12)
    (ap.abs() ** 2).sum().backward()
    assert all(v.grad is not None for v in f.values())


def test_port_fields_and_s_parameters():
    rng = np.random.default_rng(7)
    mode = _mode(rng)

    class R:
        dft = {"out": _combo(mode, np.array([0.5 + 0.1j]), np.array([0.2j]))}

    class M:
        pass

    mr = M()
    for c in PORT_COMPONENTS:
        setattr(mr, c, mode[c][None])
    ap, am = s_parameters(R(), "out", mr, 0, dA=1.0)
    np.testing.assert_allclose(ap, [0.5 + 0.1j], rtol=1e-12)
    np.testing.assert_allclose(am, [0.2j], rtol=1e-12, atol=1e-14)

    class Bad:
        dft = {"out": {"Ey": 0}}

    with pytest.raises(ValueError):
        port_fields(Bad(), "out")
