"""The oracle against the golden fixtures written by the reference's NumPy path.

The oracle must reproduce every reference trajectory bit for bit (float64 and
float32); these are the checks that pin it before it is trusted as the
checker of the HIP stepper.
"""
import numpy as np
import pytest

import golden_io as G
from scenes import F32_SCENES, SCENES


@pytest.mark.parametrize("key", list(SCENES))
def test_oracle_reproduces_reference_fp64(key):
    import photonfdtd_amd as pf
    from oracle.fdtd_numpy import run_oracle
    gold = G.load(key)
    sim = SCENES[key](pf, "hip")
    assert sim.n_steps == int(gold["n_steps"])
    assert sim.dt == float(gold["dt"])
    assert G.compare(run_oracle(sim), gold, 0.0) == 0.0


@pytest.mark.parametrize("key", list(F32_SCENES))
def test_oracle_reproduces_reference_fp32(key):
    import photonfdtd_amd as pf
    from oracle.fdtd_numpy import run_oracle
    gold = G.load(key + "_f32")
    assert G.compare(run_oracle(SCENES[key](pf, "hip", precision="float32")), gold, 0.0) == 0.0


@pytest.mark.parametrize("key", list(SCENES))
def test_scene_tables_match_reference(key):
    """dt, n_steps, CPML tables, stamped eps and expanded sources (package
    setup path, independent of the oracle's own tables)."""
    import photonfdtd_amd as pf
    gold = G.load(key)
    sim = SCENES[key](pf, "hip")
    assert np.array_equal(np.asarray(sim.eps_r), gold["eps_r"])
    for a in range(3):
        assert np.array_equal(sim._b_e[a], gold[f"cpml_b_e{a}"])
        assert np.array_equal(sim._c_e[a], gold[f"cpml_c_e{a}"])
        assert np.array_equal(sim._b_h[a], gold[f"cpml_b_h{a}"])
        assert np.array_equal(sim._c_h[a], gold[f"cpml_c_h{a}"])
    pos = np.array([list(s.position) + [0.0] * (3 - len(s.position)) for s in sim.sources],
                   dtype=np.float64).reshape(-1, 3)
    assert np.array_equal(pos, gold["src_pos"])
    assert np.array_equal(np.array([s.amplitude for s in sim.sources]), gold["src_amp"])
    assert list(gold["src_comp"]) == [s.component for s in sim.sources]


def test_oracle_cpml_tables_match_package():
    """The oracle's own CPML restatement equals the package's table builder."""
    from oracle.fdtd_numpy import cpml_axis
    from photonfdtd_amd.boundaries import cpml_coeffs_1axis
    for n, d, npml, dt in ((80, 50e-9, 12, 1e-16), (37, 20e-9, 5, 3e-17), (9, 1e-7, 1, 2e-16)):
        a = cpml_axis(n, d, npml, dt)
        b = cpml_coeffs_1axis(n, d, npml, dt)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
