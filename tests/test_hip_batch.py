"""GPU: batched 2-D stepping reproduces single runs bit for bit.

Each batched simulation must equal its own ``backend="hip"`` run (same tables,
same per-node arithmetic) in every launch mode, for TE and TM, with uneven
grid sizes, uneven step counts, Lorentz ADE media and DFT planes; float64
batches must also match the NumPy oracle to 1e-12 relative L2.
"""
import numpy as np
import pytest

import golden_io as G

pytestmark = pytest.mark.gpu

COMPS = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def _scenes(precision, component, n=4, seed=0, steps=300, lo=40, hi=90):
    import photonfdtd_amd as pf
    from bench import batched_2d_scenes
    return batched_2d_scenes(pf, n_sims=n, lo=lo, hi=hi, steps=steps, precision=precision,
                             seed=seed, component=component)


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


def _assert_same(batch, i, sim, label):
    fields, dft = _single(sim)
    for c in COMPS:
        got = batch.field(i, c)
        assert np.array_equal(got, fields[c]), f"{label}: sim {i} {c} differs"
    for name, comps in dft.items():
        got = batch.dft(i, name)
        for c, v in comps.items():
            assert np.array_equal(got[c], v), f"{label}: sim {i} DFT {name}/{c} differs"


def _has_ade(sim):
    return any(type(s.medium).__name__ == "DispersiveMedium" for s in sim.structures)


@pytest.mark.parametrize("precision", ["float32", "float64"])
@pytest.mark.parametrize("mode", ["stream", "persistent", "graph"])
def test_te_batch_matches_single_runs(precision, mode):
    from photonfdtd_amd.batch import SimBatch
    sims = _scenes(precision, "Ex", n=4, seed=1)
    sims += _scenes(precision, "Ex", n=1, seed=2, steps=170)     # stops early
    assert any(_has_ade(s) for s in sims)
    print(f"\n[case] TE batch {precision} {mode} shapes={[s.grid.shape[:2] for s in sims]}",
          flush=True)
    b = SimBatch(sims, mode=mode, waves=2, rows=16)
    try:
        b.advance()
        for i, s in enumerate(sims):
            _assert_same(b, i, s, f"TE {precision} {mode}")
    finally:
        b.close()


@pytest.mark.parametrize("mode", ["persistent", "graph"])
def test_tm_batch_matches_single_runs(mode):
    from photonfdtd_amd.batch import SimBatch
    sims = _scenes("float32", "Ez", n=4, seed=5)
    assert any(_has_ade(s) for s in sims)
    print(f"\n[case] TM batch {mode} shapes={[s.grid.shape[:2] for s in sims]}", flush=True)
    b = SimBatch(sims, mode=mode, waves=8, rows=7)
    try:
        b.advance(120)
        b.advance()                       # in two pieces
        for i, s in enumerate(sims):
            _assert_same(b, i, s, f"TM {mode}")
    finally:
        b.close()


def test_mixed_polarisations_in_one_batch():
    from photonfdtd_amd.batch import SimBatch
    sims = _scenes("float32", "Ex", n=2, seed=7) + _scenes("float32", "Ez", n=2, seed=8)
    print("\n[case] mixed TE/TM batch", flush=True)
    b = SimBatch(sims, mode="stream", waves=1, rows=128)
    try:
        b.advance()
        for i, s in enumerate(sims):
            _assert_same(b, i, s, "mixed")
    finally:
        b.close()


def test_tune_resets_state():
    from photonfdtd_amd.batch import SimBatch
    sims = _scenes("float32", "Ex", n=3, seed=11)
    print("\n[case] tune then run", flush=True)
    b = SimBatch(sims)
    try:
        best = b.tune(trial_steps=60, rows=(16, 64), log=lambda m: print(m, flush=True))
        assert best["mode"] in ("graph", "stream", "persistent") and best["ms_per_step"] > 0
        assert len(best["trials"]) == 18
        b.advance()
        for i, s in enumerate(sims):
            _assert_same(b, i, s, "tuned")
    finally:
        b.close()


def test_fp64_batch_matches_oracle():
    from oracle.fdtd_numpy import run_oracle
    from photonfdtd_amd.batch import run_batch
    sims = _scenes("float64", "Ex", n=2, seed=1, steps=200)
    print("\n[case] fp64 batch vs oracle", flush=True)
    res = run_batch(sims)
    for s, r in zip(sims, res):
        ref = run_oracle(s)
        for name, comps in ref["dft"].items():
            for c, v in comps.items():
                e = G.rel_l2(r.dft[name][c], v)
                print(f"[fp64] {name}/{c} rel L2 = {e:.3e}", flush=True)
                assert e <= 1e-12
