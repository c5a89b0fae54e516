"""GPU parity of the two-step march (one launch advances two Yee steps).

The pair kernel recomputes a two-node apron around every tile and keeps the
intermediate step in registers and LDS; it must reproduce the single-step
march bit for bit: fields, both psi sides, the ADE polarisation, E-side
sources of both steps and the DFT samples it writes for the first step.
Runs are stepped in uneven segments so pairs and single steps interleave.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene(pf, precision, interval=1):
    """Awkward 3-D grid (extents off the tile sizes), CPML of different
    widths per axis, a Drude box, a 3-pole Sellmeier box, three E dipoles
    (two on one cell), DFT planes on x and y."""
    grid = pf.Grid(size=(70 * 40e-9, 37 * 40e-9, 130 * 40e-9), cell_size=40e-9,
                   pml_layers=(6, 5, 7))
    drude = pf.DispersiveMedium.drude(1.0, f_plasma_Hz=80e12, f_collision_Hz=5e12)
    st = [pf.Box(center=(0, 0, 0), size=(1e-6, 0.6e-6, 2e-6), medium=pf.Medium(4.0)),
          pf.Box(center=(0.3e-6, 0.1e-6, -0.5e-6), size=(0.4e-6, 0.3e-6, 0.6e-6), medium=drude),
          pf.Box(center=(-0.6e-6, -0.1e-6, 1.2e-6), size=(0.3e-6, 0.2e-6, 0.5e-6),
                 medium=pf.silicon())]
    wave = pf.GaussianPulse(freq0=pf.C_0 / 1.2e-6, fwhm=4e-15)
    src = [pf.PointDipole(position=(-0.4e-6, 0.05e-6, 0.3e-6), component="Ez", waveform=wave),
           pf.PointDipole(position=(-0.4e-6, 0.05e-6, 0.3e-6), component="Ez", waveform=wave,
                          amplitude=-0.5),
           pf.PointDipole(position=(0.2e-6, -0.2e-6, -0.9e-6), component="Ex",
                          waveform=pf.GaussianPulse(freq0=pf.C_0 / 1.0e-6, fwhm=3e-15))]
    dft = pf.DFTMonitor(name="d", components=("Ex", "Ez", "Hy"),
                        freqs=[pf.C_0 / 1.2e-6, pf.C_0 / 1.0e-6], plane_axis="x",
                        plane_position=0.5e-6, interval=interval)
    dfy = pf.DFTMonitor(name="y", components=("Hz", "Ey", "Hx"), freqs=[pf.C_0 / 1.1e-6],
                        plane_axis="y", plane_position=-0.1e-6, interval=interval)
    mon = pf.FieldMonitor(name="s", components=("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"),
                          times=[2.1e-15, 4.4e-15, 6e-15])
    return pf.Simulation(grid, structures=st, sources=src, monitors=[dft, dfy, mon],
                         run_time=7e-15, precision=precision, backend="hip", courant=0.9)


def _pair_launches(pf, precision, monkeypatch, pair):
    """Step the scene directly through the engine; return (fields, dft, pair launches)."""
    from photonfdtd_amd import _native as N
    from photonfdtd_amd.hip_backend import HipEngine
    monkeypatch.setenv("PHOTONFDTD_HIP_PAIR", "1" if pair else "0")
    sim = _scene(pf, precision)
    npdt = np.float32 if precision == "float32" else np.float64
    eng = HipEngine(sim, npdt, True)
    lib = eng.lib
    N.check(lib, lib.pf_timing_enable(eng.h, 1), "pf_timing_enable")
    eng.advance(0, 7)
    eng.advance(7, 12)
    ms = np.zeros(len(N.KERNEL_CLASSES), dtype=np.float64)
    cnt = np.zeros(len(N.KERNEL_CLASSES), dtype=np.int64)
    N.check(lib, lib.pf_timing_read(eng.h, N.as_ptr(ms, N.C.c_double),
                                    N.as_ptr(cnt, N.C.c_int64)), "pf_timing_read")
    fields = {c: eng.get(c) for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")}
    dft = eng.dft("d")
    eng.close()
    return fields, dft, int(cnt[N.KERNEL_CLASSES.index("yee_pair")])


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_pair_engine_launches(precision, monkeypatch):
    """7 + 12 steps: 3 + 6 pair launches and one single step when enabled,
    none when switched off; identical state either way."""
    import photonfdtd_amd as pf
    print(f"\n[case] pair engine {precision}", flush=True)
    monkeypatch.setenv("PHOTONFDTD_HIP_EXACT", "1")
    fa, da, na = _pair_launches(pf, precision, monkeypatch, True)
    fb, db, nb = _pair_launches(pf, precision, monkeypatch, False)
    assert na == 9 and nb == 0, (na, nb)
    for c in fa:
        assert np.array_equal(fa[c], fb[c]), c
    for c in da:
        assert np.array_equal(da[c], db[c]), c
    assert np.abs(fa["Ez"]).max() > 0


@pytest.mark.parametrize("precision", ["float64", "float32"])
@pytest.mark.parametrize("xchunk,nw", [("64", "10"), ("7", "10"), ("64", "12"), ("5", "12")])
def test_pair_bitwise_equals_single_step(precision, xchunk, nw, monkeypatch):
    import photonfdtd_amd as pf
    print(f"\n[case] pair vs single {precision} xchunk={xchunk} nw={nw}", flush=True)
    monkeypatch.setenv("PHOTONFDTD_HIP_EXACT", "1")
    monkeypatch.setenv("PHOTONFDTD_HIP_XCHUNK", xchunk)
    monkeypatch.setenv("PHOTONFDTD_HIP_PAIR_NW", nw)
    monkeypatch.setenv("PHOTONFDTD_HIP_PAIR", "1")
    a = _scene(pf, precision).run()
    monkeypatch.setenv("PHOTONFDTD_HIP_PAIR", "0")
    b = _scene(pf, precision).run()
    for c in a.fields["s"]:
        assert np.array_equal(a.fields["s"][c], b.fields["s"][c]), c
    for m in ("d", "y"):
        for c in a.dft[m]:
            assert np.array_equal(a.dft[m][c], b.dft[m][c]), (m, c)
    assert np.abs(a.fields["s"]["Ez"]).max() > 0


@pytest.mark.parametrize("interval", [2, 3])
def test_pair_dft_interval(interval, monkeypatch):
    """DFT sampling every 2nd / 3rd step: the kernel writes the samples of
    the pair's first step only when that step is on the sampling grid."""
    import photonfdtd_amd as pf
    print(f"\n[case] pair DFT interval {interval}", flush=True)
    monkeypatch.setenv("PHOTONFDTD_HIP_EXACT", "1")
    monkeypatch.setenv("PHOTONFDTD_HIP_PAIR", "1")
    a = _scene(pf, "float64", interval).run()
    monkeypatch.setenv("PHOTONFDTD_HIP_PAIR", "0")
    b = _scene(pf, "float64", interval).run()
    for m in ("d", "y"):
        for c in a.dft[m]:
            assert np.array_equal(a.dft[m][c], b.dft[m][c]), (m, c)
            assert np.abs(a.dft[m][c]).max() > 0


def test_pair_fast_order_bitwise(monkeypatch):
    """Fast operation order (multiply by 1/d) in float32: same bits as the
    single-step march in the same order."""
    import photonfdtd_amd as pf
    print("\n[case] pair fast order float32", flush=True)
    monkeypatch.setenv("PHOTONFDTD_HIP_EXACT", "0")
    monkeypatch.setenv("PHOTONFDTD_HIP_PAIR", "1")
    a = _scene(pf, "float32").run()
    monkeypatch.setenv("PHOTONFDTD_HIP_PAIR", "0")
    b = _scene(pf, "float32").run()
    for c in a.fields["s"]:
        assert np.array_equal(a.fields["s"][c], b.fields["s"][c]), c


@pytest.mark.parametrize("precision", ["float64", "float32"])
def test_pair_single_steps_at_both_polarisation_levels(precision, monkeypatch):
    """Uneven advances (1, 2, 1, 3, 2, 1, 4, 1 steps) put single steps after
    both an even and an odd number of pairs, so the all-media ADE launch of a
    single step runs at both polarisation levels of the two-step march (two
    dispersive media: the Drude and the silicon box). Same bits as single
    steps only."""
    import photonfdtd_amd as pf
    from photonfdtd_amd.hip_backend import HipEngine
    print(f"\n[case] pair single steps at both P levels {precision}", flush=True)
    monkeypatch.setenv("PHOTONFDTD_HIP_EXACT", "1")
    npdt = np.float32 if precision == "float32" else np.float64
    out = []
    for pair in ("1", "0"):
        monkeypatch.setenv("PHOTONFDTD_HIP_PAIR", pair)
        eng = HipEngine(_scene(pf, precision), npdt, True)
        pos = 0
        for n in (1, 2, 1, 3, 2, 1, 4, 1):
            eng.advance(pos, n)
            pos += n
        out.append(({c: eng.get(c).copy() for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")},
                    eng.dft("d")))
        eng.close()
    (fa, da), (fb, db) = out
    for c in fa:
        assert np.array_equal(fa[c], fb[c]), c
    for c in da:
        assert np.array_equal(da[c], db[c]), c
    assert np.abs(fa["Ez"]).max() > 0
