"""The HIP driver depends only on the reference package's own surface.

``photonfdtd_amd/hip_backend.py`` is the module a maintainer drops into the
reference package as its ``backend="hip"`` driver (beside ``_native.py``).
Two checks keep that true without importing the reference:

* every name the driver imports from a sibling module exists under that name
  in the reference module of the same name (the lists below are read off the
  reference source, file:line given per entry);
* ``build_desc`` runs on stand-in ``Simulation`` / ``Grid`` objects that
  expose only the reference's attribute names (anything else raises
  ``AttributeError``) and produces the same tables as on this package's own
  ``Simulation``.
"""
import ast
from pathlib import Path

import numpy as np
import pytest

import scenes

ROOT = Path(__file__).resolve().parent.parent

# Names each reference module defines (src/photonfdtd/...).
REFERENCE_NAMES = {
    "constants": {"C_0", "EPS_0", "MU_0"},                        # constants.py:3-8
    "monitors": {"FieldMonitor", "DFTMonitor", "FluxMonitor"},   # monitors.py:18,95,177
    "simulation": {"Simulation", "Result", "_COMPUTE_KEYS"},     # simulation.py:274,327,361
    "storage": {"CompressedFieldSeries", "get_codec"},           # storage.py:80,104
}
# Modules shipped with the driver itself (copied beside it).
SHIPPED = {"_native"}

# Attributes of the reference Simulation the driver may read
# (simulation.py:361-528,794-808; rustbackend.py:85-166 reads the same set).
SIM_ATTRS = ("grid", "structures", "sources", "monitors", "dt", "n_steps", "dtypes", "eps_r",
             "_b_e", "_c_e", "_b_h", "_c_h", "subpixel", "particle_sources",
             "progress_callback", "verbose", "_smoothed_eps_components")
# Attributes of the reference Grid (grid.py:17-96).
GRID_ATTRS = ("size", "cell_size", "pml_layers", "shape", "coords", "ndim", "index_at")


class _Only:
    """Exposes exactly the listed attributes of ``obj``."""

    def __init__(self, obj, names, **override):
        d = {n: getattr(obj, n) for n in names}
        d.update(override)
        object.__setattr__(self, "_d", d)

    def __getattr__(self, name):
        d = object.__getattribute__(self, "_d")
        if name not in d:
            raise AttributeError(f"the reference object has no attribute {name!r}")
        return d[name]


def _reference_view(sim):
    grid = _Only(sim.grid, GRID_ATTRS)
    return _Only(sim, SIM_ATTRS, grid=grid)


def test_driver_imports_only_reference_names():
    tree = ast.parse((ROOT / "photonfdtd_amd" / "hip_backend.py").read_text())
    seen = set()
    for node in ast.walk(tree):
        if isinstance(node, ast.ImportFrom) and node.level == 1:
            if node.module is None:
                for a in node.names:
                    assert a.name in SHIPPED, f"from . import {a.name}: not shipped with the driver"
                    seen.add(a.name)
                continue
            assert node.module in REFERENCE_NAMES, f"from .{node.module}: no such reference module"
            for a in node.names:
                assert a.name in REFERENCE_NAMES[node.module], \
                    f"from .{node.module} import {a.name}: the reference module lacks it"
                seen.add(a.name)
    assert {"_native", "Result", "DFTMonitor"} <= seen


def _tables_equal(a, b, path="t"):
    if isinstance(a, dict):
        assert set(a) == set(b), path
        for k in a:
            if k == "mon":
                assert a[k] is b[k], path
            else:
                _tables_equal(a[k], b[k], f"{path}.{k}")
    elif isinstance(a, (list, tuple)):
        assert len(a) == len(b), path
        for i, (x, y) in enumerate(zip(a, b)):
            _tables_equal(x, y, f"{path}[{i}]")
    elif isinstance(a, np.ndarray):
        assert np.array_equal(a, b), path
    elif hasattr(a, "__dataclass_fields__"):
        for f in a.__dataclass_fields__:
            _tables_equal(getattr(a, f), getattr(b, f), f"{path}.{f}")
    else:
        assert a == b, path


@pytest.mark.parametrize("key", ["metals_xz", "mode3d", "flux2d", "te2d_poly", "box3d"])
def test_build_tables_on_reference_surface(key):
    import photonfdtd_amd as pf
    from photonfdtd_amd import hip_backend as hb
    sim = scenes.SCENES[key](pf, "hip")
    for npdt in (np.float64, np.float32):
        own = hb.build_tables(sim, npdt)
        ref = hb.build_tables(_reference_view(sim), npdt)
        _tables_equal(own, ref)
        hb.validate(_reference_view(sim))
        # the descriptor itself fills from the stand-in too
        hb.build_desc(_reference_view(sim), npdt, exact=True)


def test_stand_in_rejects_unknown_attributes():
    import photonfdtd_amd as pf
    v = _reference_view(scenes.SCENES["box3d"](pf, "hip"))
    with pytest.raises(AttributeError):
        v.ade_media
    with pytest.raises(AttributeError):
        v.grid.n_cells


def test_ade_tables_match_per_cell_owner():
    """ade_tables against a brute-force owner map: the last structure that
    covers a cell decides it; dispersive media keyed by identity."""
    import photonfdtd_amd as pf
    from photonfdtd_amd.hip_backend import ade_tables
    sim = scenes.SCENES["metals_xz"](pf, "hip")
    owner = {}
    for st in sim.structures:
        for c in np.flatnonzero(st.region_mask(sim.grid).ravel()):
            owner[int(c)] = st.medium if st.medium.is_dispersive else None
    media = ade_tables(sim)
    got = {}
    for m in media:
        for c in m.cells:
            assert int(c) not in got
            got[int(c)] = m
    want = {c for c, med in owner.items() if med is not None}
    assert set(got) == want
    # cells of one reference medium object share one table entry
    for c in want:
        same = [d for d in want if owner[d] is owner[c]]
        assert all(got[d] is got[c] for d in same[:50])
