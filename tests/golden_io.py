"""Helpers to read the golden fixtures and compare a run against them."""
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    with np.load(GOLDEN / f"{name}.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return 0.0 if den == 0 and num == 0 else num / max(den, 1e-300)


def compare(res, gold, tol, report=None):
    """Max relative-L2 error of every monitor output in ``res`` vs ``gold``.

    ``res`` is a Result or the oracle's result dict. Returns the worst error
    and raises AssertionError when any output exceeds ``tol``.
    """
    get = (lambda k: res[k]) if isinstance(res, dict) else (lambda k: getattr(res, k))
    worst = 0.0
    seen = 0
    for key, ref in gold.items():
        parts = key.split("/")
        if parts[0] == "field":
            got = get("fields")[parts[1]][parts[2]]
        elif parts[0] == "dft":
            got = get("dft")[parts[1]][parts[2]]
        elif parts[0] == "flux":
            got = get("flux")[parts[1]]
        elif parts[0] == "mt":
            got = get("monitor_times")[parts[1]]
            assert np.array_equal(np.asarray(got), ref), f"{key}: monitor times differ"
            continue
        elif parts[0] == "dftf":
            assert np.array_equal(np.asarray(get("dft_freqs")[parts[1]]), ref)
            continue
        else:
            continue
        got = np.asarray(got)
        assert got.shape == ref.shape, f"{key}: shape {got.shape} != {ref.shape}"
        e = rel_l2(got, ref)
        if report is not None:
            report[key] = e
        worst = max(worst, e)
        seen += 1
        assert e <= tol, f"{key}: relative L2 error {e:.3e} > {tol:.1e}"
    assert seen > 0
    return worst
