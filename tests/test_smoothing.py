"""Subpixel smoothing (reference ``smoothing.py:119-172``): the package's
tensors equal the ones the reference's own path wrote into the subpixel
fixtures, bit for bit; a cell inside one material keeps that material's
permittivity; an axis-aligned face between
nodes gives the harmonic mean on the normal component and the arithmetic
mean on the tangential ones (closed form for a half-filled cell)."""
import numpy as np
import pytest

import golden_io as G
import scenes


@pytest.mark.parametrize("key", ["subpixel2d", "subpixel3d"])
def test_tensors_equal_reference_fixture(key):
    import photonfdtd_amd as pf
    sim = scenes.SCENES[key](pf, "hip")
    gold = G.load(key)
    for c, e in zip(("xx", "yy", "zz"), sim._smoothed_eps_components()):
        assert np.array_equal(e, gold[f"eps_{c}"]), c


def test_half_filled_cells_closed_form():
    """A box face on the midpoint between two x nodes, factor 2: the cells
    on the face are half eps1, half eps2 along x only."""
    import photonfdtd_amd as pf
    from photonfdtd_amd.smoothing import smooth_permittivity
    d = 50e-9
    grid = pf.Grid(size=(20 * d, 12 * d), cell_size=d, pml_layers=(0, 0, 0))
    x_face = grid.coords[0][10]            # the box ends on node 10's centre line
    box = pf.Box(center=(x_face - 5 * d, 0.0), size=(10 * d, 40 * d), medium=pf.Medium(9.0))
    exx, eyy, ezz = smooth_permittivity(grid, [box], factor=2)
    e_in, e_out = 9.0, 1.0
    inside = box.region_mask(grid)[:, 0, 0]
    assert exx[2, 5, 0] == e_in and exx[15, 5, 0] == e_out
    # node 10 holds the face: its sub-nodes at +-d/4 are half in, half out
    a = (e_in + e_out) / 2
    h = 2 / (1 / e_in + 1 / e_out)
    assert inside[10] and not inside[11]
    assert exx[10, 5, 0] == pytest.approx(h, rel=1e-12)     # normal along x: harmonic
    assert eyy[10, 5, 0] == pytest.approx(a, rel=1e-12)     # tangential: arithmetic
    assert ezz[10, 5, 0] == pytest.approx(a, rel=1e-12)
