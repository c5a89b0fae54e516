"""GPU: batched 3-D stepping reproduces single runs bit for bit.

Per step, the march launches (psi-free and CPML tiles) cover the tiles of
every 3-D simulation in the batch, then one launch each runs the ADE media,
the E-side sources and the DFT captures of all of them. Each batched
simulation must equal its own ``backend="hip"`` run
(fields and DFT planes) for uneven grid shapes, uneven step counts, Drude
and Lorentz boxes, several march shapes (wavefronts per workgroup, planes
per tile) and both operation orders; float64 must also match the NumPy
oracle to 1e-12 relative L2.
"""
import numpy as np
import pytest

import golden_io as G

pytestmark = pytest.mark.gpu

COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def _scene(precision, shape, run_time, seed, interval=1):
    """3-D grid of the given node counts with CPML of different widths per
    axis, a dielectric box, a Drude box, a two-pole Lorentz box, two E
    dipoles and DFT planes on x and y; the structure placement follows the
    seed. (The media keep omega0 * dt far below the ADE stability limit: a
    silicon Sellmeier box at this dt grows without bound in the oracle too.)"""
    import photonfdtd_amd as pf
    rng = np.random.default_rng(seed)
    dx = 40e-9
    nx, ny, nz = shape
    grid = pf.Grid(size=(nx * dx, ny * dx, nz * dx), cell_size=dx, pml_layers=(5, 4, 6))
    L = np.array([nx, ny, nz]) * dx

    def box(frac, medium):
        c = (rng.uniform(-0.15, 0.15, 3) * L).tolist()
        return pf.Box(center=tuple(c), size=tuple((np.asarray(frac) * L).tolist()), medium=medium)

    drude = pf.DispersiveMedium.drude(1.0, f_plasma_Hz=80e12, f_collision_Hz=5e12)
    st = [box((0.5, 0.4, 0.5), pf.Medium(float(rng.uniform(2.0, 6.0)))),
          box((0.2, 0.2, 0.15), drude),
          box((0.15, 0.2, 0.2), pf.DispersiveMedium.lorentz(
              2.0, [(1.5, 150e12, 5e12), (0.5, 300e12, 10e12)]))]
    wave = pf.GaussianPulse(freq0=pf.C_0 / 1.2e-6, fwhm=4e-15)
    p = (rng.uniform(-0.2, 0.2, 3) * L).tolist()
    src = [pf.PointDipole(position=tuple(p), component="Ez", waveform=wave),
           pf.PointDipole(position=(-p[0], p[1], -p[2]), component="Ex", waveform=wave,
                          amplitude=0.5)]
    mon = [pf.DFTMonitor(name="d", components=("Ex", "Ez", "Hy"),
                         freqs=[pf.C_0 / 1.2e-6, pf.C_0 / 1.0e-6], plane_axis="x",
                         plane_position=0.1 * L[0], interval=interval),
           pf.DFTMonitor(name="y", components=("Hz", "Ey"), freqs=[pf.C_0 / 1.1e-6],
                         plane_axis="y", plane_position=-0.1 * L[1], interval=interval)]
    return pf.Simulation(grid, structures=st, sources=src, monitors=mon, run_time=run_time,
                         precision=precision, backend="hip", courant=0.9)


def _scenes(precision, interval=1):
    return [_scene(precision, (40, 30, 70), 8e-15, 1, interval),
            _scene(precision, (25, 44, 130), 5e-15, 2, interval),     # stops early
            _scene(precision, (70, 19, 62), 8e-15, 3, interval)]


def _single(sim):
    from photonfdtd_amd.hip_backend import HipEngine, exact_mode, validate
    npdt = validate(sim)
    eng = HipEngine(sim, npdt, exact_mode(npdt))
    try:
        eng.advance(0, sim.n_steps)
        fields = {c: eng.get(c).copy() for c in COMPS}
        dft = {m.name: eng.dft(m.name) for m in sim.monitors}
        return fields, dft
    finally:
        eng.close()


def _assert_same(batch, i, ref, label):
    fields, dft = ref
    for c in COMPS:
        got = batch.field(i, c)
        assert np.array_equal(got, fields[c]), f"{label}: sim {i} {c} differs"
    for name, comps in dft.items():
        got = batch.dft(i, name)
        for c, v in comps.items():
            assert np.array_equal(got[c], v), f"{label}: sim {i} DFT {name}/{c} differs"
    assert np.abs(fields["Ez"]).max() > 0 and np.isfinite(fields["Ez"]).all()


@pytest.mark.parametrize("mode", ["stream", "graph"])
@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("waves,rows", [(10, 128), (12, 7), (16, 5)])
def test_batch3_matches_single_runs(precision, waves, rows, mode, monkeypatch):
    from photonfdtd_amd.batch import SimBatch
    monkeypatch.setenv("PHOTONFDTD_HIP_EXACT", "1")
    sims = _scenes(precision)
    # the first simulation twice: the same tables in two slots of the batch
    sims = sims + [sims[0]]
    print(f"\n[case] 3-D batch {precision} {mode} waves={waves} rows={rows} "
          f"shapes={[s.grid.shape for s in sims]} steps={[s.n_steps for s in sims]}", flush=True)
    assert len({s.n_steps for s in sims}) > 1
    refs = [_single(s) for s in sims[:3]] + [None]
    refs[3] = refs[0]
    b = SimBatch(sims, mode=mode, waves=waves, rows=rows)
    try:
        b.advance(31)
        b.advance()                                 # in two pieces
        for i in range(len(sims)):
            _assert_same(b, i, refs[i], f"{precision} {waves}/{rows}")
    finally:
        b.close()


@pytest.mark.parametrize("mode", ["stream", "graph"])
def test_batch3_fast_order_reset_and_interval(mode, monkeypatch):
    """Fast operation order (float32), DFT every 3rd step into rings of 5
    samples (folds at both step parities, so graph pairs stop short of them);
    a reset and a second run give the same bits again."""
    from photonfdtd_amd.batch import SimBatch
    monkeypatch.setenv("PHOTONFDTD_HIP_EXACT", "0")
    monkeypatch.setenv("PHOTONFDTD_HIP_DFT_RING", "5")
    sims = _scenes("float32", interval=3)
    print(f"\n[case] 3-D batch fast order, interval 3, ring 5, reset, {mode}", flush=True)
    refs = [_single(s) for s in sims]
    b = SimBatch(sims, mode=mode)
    try:
        b.advance(40)
        b.reset()
        b.advance()
        for i in range(len(sims)):
            _assert_same(b, i, refs[i], "fast/reset")
    finally:
        b.close()


def test_batch3_tune_and_set_eps(monkeypatch):
    """The autotuner's trials leave no state behind, and a new permittivity
    written into one slot equals a single run of that permittivity."""
    from photonfdtd_amd.batch import SimBatch
    monkeypatch.setenv("PHOTONFDTD_HIP_EXACT", "1")
    sims = _scenes("float32")
    other = _scene("float32", sims[1].grid.shape, 5e-15, 9)   # same grid, other structures
    print("\n[case] 3-D batch tune + set_eps", flush=True)
    b = SimBatch(sims)
    try:
        best = b.tune(trial_steps=20, rows=(32, 128), log=lambda m: print(m, flush=True))
        assert best["mode"] in ("stream", "graph") and best["ms_per_step"] > 0
        assert len(best["trials"]) == 12
        b.set_eps(1, other.eps_r)
        b.advance()
        _assert_same(b, 0, _single(sims[0]), "tuned")
        # sim 1 keeps its own dispersive boxes (the ADE tables): compare with
        # a single run of its scene whose permittivity grid is replaced
        ref = _scene("float32", sims[1].grid.shape, 5e-15, 2)
        ref.eps_r = other.eps_r
        _assert_same(b, 1, _single(ref), "set_eps")
    finally:
        b.close()


def test_batch3_fp64_matches_oracle():
    from oracle.fdtd_numpy import run_oracle
    from photonfdtd_amd.batch import run_batch
    sims = [_scene("float64", (30, 24, 66), 4e-15, 4), _scene("float64", (22, 34, 40), 3e-15, 5)]
    print("\n[case] 3-D fp64 batch vs oracle", flush=True)
    res = run_batch(sims)
    for s, r in zip(sims, res):
        ref = run_oracle(s)
        for name, comps in ref["dft"].items():
            for c, v in comps.items():
                e = G.rel_l2(r.dft[name][c], v)
                print(f"[fp64] {name}/{c} rel L2 = {e:.3e}", flush=True)
                assert e <= 1e-12
