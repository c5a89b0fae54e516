"""GPU parity of the 3-D dispersive slab scene (silicon Sellmeier box, gold
Drude cube, silica cladding, Ey dipole, two 32-frequency x-plane DFT
monitors, 12-cell CPML on six faces) against the float64 oracle.

The grid (48 x 72 x 140) is sized so the 64-lane march runs every kernel
variant: the z tile of nodes [64, 128) lies inside the CPML-free range
[13, 128), and so do x segments and y blocks, so the psi-free tile list is
non-empty and the silicon box crosses those tiles; the tiles that reach into
the z slabs only, the y slabs only, the x slabs only and several slabs each
run their own variant. Bars: float64 in the reference operation order
within 1e-12 relative L2 of the oracle (fields and both DFT planes), float32
in the fast order within 1e-4; the other layouts (dense rows with 62-node
windows, every tile on the all-axis kernel; padded rows of 62 nodes per 64
slots) give the same bits as the 64-lane split launch.
"""
import numpy as np
import pytest

from scenes import dispersive_slab

pytestmark = pytest.mark.gpu

FIELDS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")
# the 64-lane march (the default from 2^25 nodes; this grid is smaller)
M64 = {"PHOTONFDTD_HIP_ZPAD": "64"}


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


@pytest.fixture(scope="module")
def oracle_run():
    import photonfdtd_amd as pf
    from oracle.fdtd_numpy import run_oracle
    print("\n[case] dispersive slab 48x72x140 oracle (float64, 300 steps)", flush=True)
    return run_oracle(dispersive_slab(pf, "hip"), return_state=True)


def _gpu(precision, exact, env, monkeypatch, **scene):
    import photonfdtd_amd as pf
    from photonfdtd_amd.hip_backend import HipEngine
    for k in ("PHOTONFDTD_HIP_ZPAD", "PHOTONFDTD_HIP_SPLIT", "PHOTONFDTD_HIP_PAIR",
              "PHOTONFDTD_HIP_MIX", "PHOTONFDTD_HIP_EVEN_X", "PHOTONFDTD_HIP_XCHUNK",
              "PHOTONFDTD_HIP_POST", "PHOTONFDTD_HIP_SIDE"):
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    sim = dispersive_slab(pf, "hip", precision=precision, **scene)
    eng = HipEngine(sim, np.float64 if precision == "float64" else np.float32, exact)
    info = eng.info()
    eng.advance(0, sim.n_steps)
    out = {c: eng.get(c).astype(np.float64) for c in FIELDS}
    dft = {nm: eng.dft(nm) for nm in ("in", "out")}
    eng.close()
    return info, out, dft


def _worst(out, dft, ref, state):
    w = 0.0
    for c in FIELDS:
        w = max(w, _rel(out[c], state[c]))
    for nm in ("in", "out"):
        for c, v in dft[nm].items():
            w = max(w, _rel(v, ref["dft"][nm][c]))
    return w


def test_fp64_exact_split_march_matches_oracle(oracle_run, monkeypatch):
    ref, state = oracle_run
    print("[case] dispersive slab float64 exact, 64-lane split march", flush=True)
    info, out, dft = _gpu("float64", True, M64, monkeypatch)
    assert info["path"] == 2 and info["zpad"] == 64, info   # the 64-lane march
    assert info["tiles_bulk"] > 0 and info["cells_bulk"] > 0, info
    assert info["cpml_lists"] == 0b1111, info   # several-axis, z-, y- and x-only tiles
    worst = _worst(out, dft, ref, state)
    print(f"[fp64] dispersive slab worst rel L2 = {worst:.3e}; psi-free tiles "
          f"{info['tiles_bulk']} ({info['cells_bulk']} cells), CPML tiles {info['tiles_cpml']}",
          flush=True)
    assert worst <= 1e-12
    assert np.abs(state["Ey"]).max() > 0


def test_fp32_fast_within_1e4_of_oracle(oracle_run, monkeypatch):
    ref, state = oracle_run
    print("[case] dispersive slab float32 fast, 64-lane split march", flush=True)
    info, out, dft = _gpu("float32", False, M64, monkeypatch)
    assert info["tiles_bulk"] > 0 and info["cpml_lists"] == 0b1111, info
    worst = _worst(out, dft, ref, state)
    print(f"[fp32] dispersive slab worst rel L2 vs float64 oracle = {worst:.3e}", flush=True)
    assert worst <= 1e-4


@pytest.mark.parametrize("zpad", ["0", "62"])
@pytest.mark.parametrize("precision,exact", [("float64", True), ("float32", False)])
def test_other_layouts_same_bits_as_split_march(precision, exact, zpad, monkeypatch):
    print(f"[case] dispersive slab 64-lane split vs zpad {zpad} {precision}", flush=True)
    ia, fa, da = _gpu(precision, exact, M64, monkeypatch)
    ib, fb, db = _gpu(precision, exact, {"PHOTONFDTD_HIP_ZPAD": zpad}, monkeypatch)
    assert ia["tiles_bulk"] > 0 and ib["zpad"] == int(zpad), (ia, ib)
    assert (ib["tiles_bulk"] > 0) == (zpad != "0"), ib
    for c in FIELDS:
        assert np.array_equal(fa[c], fb[c]), c
    for nm in ("in", "out"):
        for c in da[nm]:
            assert np.array_equal(da[nm][c], db[nm][c]), (nm, c)


@pytest.mark.parametrize("mix", ["4", "12", "28"])
@pytest.mark.parametrize("precision,exact", [("float64", True), ("float32", False)])
def test_mixed_launch_same_bits_as_per_list_launches(precision, exact, mix, monkeypatch):
    """The z-only (mix 4), z- and y-only (mix 12) or z-, y- and x-only (mix
    28) CPML tiles launched with the psi-free tiles in one spatially ordered
    launch, over x chunks of nearly equal length, give the bits of one launch
    per list over 48-plane chunks."""
    print(f"[case] dispersive slab mixed launch {mix} {precision}", flush=True)
    # one launch per list, x chunks of at most 48 planes with a short last one
    env = dict(M64, PHOTONFDTD_HIP_MIX="0", PHOTONFDTD_HIP_EVEN_X="0", PHOTONFDTD_HIP_XCHUNK="48")
    ia, fa, da = _gpu(precision, exact, env, monkeypatch)
    env = dict(M64, PHOTONFDTD_HIP_MIX=mix)   # x chunks of nearly equal length
    ib, fb, db = _gpu(precision, exact, env, monkeypatch)
    assert ia["cells_bulk"] > 0 and ib["cells_bulk"] > 0 and ib["cpml_lists"] == 0b1111, (ia, ib)
    for c in FIELDS:
        assert np.array_equal(fa[c], fb[c]), c
    for nm in ("in", "out"):
        for c in da[nm]:
            assert np.array_equal(da[nm][c], db[nm][c]), (nm, c)


@pytest.mark.parametrize("precision,exact", [("float64", True), ("float32", False)])
def test_post_step_launch_and_main_stream_x_list_same_bits(precision, exact, monkeypatch):
    """The ADE media, the E-side dipole and both DFT planes' samples in one
    launch after the march, with the x-only CPML list on the main stream
    after the mixed launch (the float32 defaults), give the bits of one
    launch per part (PHOTONFDTD_HIP_POST=0) with the x-only list on the side
    stream (PHOTONFDTD_HIP_SIDE=26).
    At 160 cells in x the DFT planes miss the gold cube and the silicon box,
    so the one-launch path is taken; on the 48-cell grid the "out" plane
    crosses the cube and the stepper keeps one launch per part."""
    print(f"[case] dispersive slab 160x72x140 post-step launch {precision}", flush=True)
    scene = dict(shape=(160, 72, 140), steps=200)
    ia, fa, da = _gpu(precision, exact, dict(M64, PHOTONFDTD_HIP_POST="0", PHOTONFDTD_HIP_SIDE="26"),
                      monkeypatch, **scene)
    ib, fb, db = _gpu(precision, exact, dict(M64, PHOTONFDTD_HIP_SIDE="2"), monkeypatch, **scene)
    assert ia["post_fused"] == 0 and ib["post_fused"] == 1, (ia, ib)
    assert ib["cpml_lists"] == 0b1111 and ib["tiles_bulk"] > 0, ib
    assert float(np.abs(fb["Ey"]).max()) > 0
    for c in FIELDS:
        assert np.array_equal(fa[c], fb[c]), c
    for nm in ("in", "out"):
        assert float(np.abs(db[nm]["Ey"]).max()) > 0, nm
        for c in da[nm]:
            assert np.array_equal(da[nm][c], db[nm][c]), (nm, c)
    print(f"[case] dispersive slab 48x72x140 post-step launch refused {precision}", flush=True)
    ic, _, _ = _gpu(precision, exact, M64, monkeypatch)
    assert ic["post_fused"] == 0, ic


def test_sampled_event_timing_counts(monkeypatch):
    """pf_timing_enable(h, k) brackets the march and aux groups of every k-th
    step only (k = 1: every step); a negative interval is refused."""
    import photonfdtd_amd as pf
    from photonfdtd_amd import _native as N
    from photonfdtd_amd.hip_backend import HipEngine
    print("[case] dispersive slab sampled event timing", flush=True)
    for k in ("PHOTONFDTD_HIP_ZPAD", "PHOTONFDTD_HIP_POST", "PHOTONFDTD_HIP_SIDE"):
        monkeypatch.delenv(k, raising=False)
    monkeypatch.setenv("PHOTONFDTD_HIP_ZPAD", "64")
    sim = dispersive_slab(pf, "hip", steps=60, precision="float32")
    eng = HipEngine(sim, np.float32, False)
    lib = eng.lib
    fused, aux = N.KERNEL_CLASSES.index("yee_fused"), N.KERNEL_CLASSES.index("aux")
    try:
        assert lib.pf_timing_enable(eng.h, -1) == -1   # PF_E_ARG
        for every, step0, n, want in ((10, 0, 40, 4), (1, 40, 20, 20)):
            N.check(lib, lib.pf_timing_enable(eng.h, every), "pf_timing_enable")
            eng.advance(step0, n)
            ms = np.zeros(len(N.KERNEL_CLASSES))
            cnt = np.zeros(len(N.KERNEL_CLASSES), dtype=np.int64)
            N.check(lib, lib.pf_timing_read(eng.h, N.as_ptr(ms, N.C.c_double),
                                            N.as_ptr(cnt, N.C.c_int64)), "pf_timing_read")
            assert cnt[fused] == want and cnt[aux] == want, (every, cnt)
            assert ms[fused] > 0 and ms[aux] > 0, (every, ms)
    finally:
        eng.close()


def test_full_size_64_lane_march_same_bits_as_padded_march(monkeypatch):
    """At 512^3 float32 the 64-lane march's scalar offsets (plane, component
    and psi-slot offsets) reach beyond 1 GiB; the padded march addresses the
    same nodes with per-plane base pointers. 60 steps of the benchmark scene
    on both layouts give the same bits (fields and both DFT planes)."""
    import photonfdtd_amd as pf
    from photonfdtd_amd.hip_backend import HipEngine

    def run(zpad):
        for k in ("PHOTONFDTD_HIP_SPLIT", "PHOTONFDTD_HIP_PAIR", "PHOTONFDTD_HIP_MIX"):
            monkeypatch.delenv(k, raising=False)
        monkeypatch.setenv("PHOTONFDTD_HIP_ZPAD", zpad)
        sim = dispersive_slab(pf, "hip", shape=(512, 512, 512), steps=60, precision="float32")
        eng = HipEngine(sim, np.float32, False)
        info = eng.info()
        eng.advance(0, sim.n_steps)
        out = {c: eng.get(c) for c in FIELDS}
        dft = {nm: eng.dft(nm) for nm in ("in", "out")}
        eng.close()
        return info, out, dft

    print("[case] dispersive slab 512^3 float32, 64-lane vs padded march, 60 steps", flush=True)
    ia, fa, da = run("64")
    assert ia["zpad"] == 64 and ia["tiles_bulk"] > 0, ia
    ib, fb, db = run("62")
    assert ib["zpad"] == 62, ib
    assert float(np.abs(fa["Ey"]).max()) > 0
    # the pulse reaches the "in" plane (27 cells from the dipole) within the
    # 60 steps, not the "out" plane
    assert float(np.abs(da["in"]["Ey"]).max()) > 0
    for c in FIELDS:
        assert np.array_equal(fa[c], fb[c]), c
    for nm in ("in", "out"):
        for c in da[nm]:
            assert np.array_equal(da[nm][c], db[nm][c]), (nm, c)
