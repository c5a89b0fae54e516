"""ctypes driver of ``libpfcpu.so`` (``cpu_fdtd.cpp``): the OpenMP restatement
of the reference's compiled CPU stepper (rust/src/kernel.rs, blocked.rs).

TEST AND BASELINE INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``. The
scene's tables come from the HIP driver's ``build_desc`` (the same
``pf_desc`` the GPU library takes); nothing here touches a GPU.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Dict

import numpy as np

LIB = Path(__file__).resolve().parent / "libpfcpu.so"
_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is missing; build it with 'python -m oracle.build_cpu'")
        lib = C.CDLL(str(LIB))
        lib.pfcpu_create.restype = C.c_int
        lib.pfcpu_create.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                     C.POINTER(C.c_void_p)]
        lib.pfcpu_advance.argtypes = [C.c_void_p, C.c_int64, C.c_int64]
        lib.pfcpu_read_field.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        lib.pfcpu_dft_read.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
        lib.pfcpu_destroy.argtypes = [C.c_void_p]
        lib.pfcpu_last_error.restype = C.c_char_p
        lib.pfcpu_threads.restype = C.c_int
        _lib = lib
    return _lib


class CpuStepper:
    """``advance(step0, n)`` / ``get(comp)`` / ``dft(name)`` like the HIP engine."""

    def __init__(self, sim, npdt, threads: int = 0, t_block: int = 0, tile_rows: int = 0):
        from photonfdtd_amd.hip_backend import build_desc
        from photonfdtd_amd import _native as N
        self.lib = load()
        self.npdt = np.dtype(npdt).type
        self.shape = sim.grid.shape
        # the tables must outlive pfcpu_create, which copies them
        d, keep, tables, self.dft_slots, _ = build_desc(sim, self.npdt, exact=False)
        h = C.c_void_p()
        self._check(self.lib.pfcpu_create(C.addressof(d), int(threads), int(t_block),
                                          int(tile_rows), C.byref(h)))
        self.h = h
        self.threads = int(self.lib.pfcpu_threads())
        del keep, tables, N

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.pfcpu_last_error().decode())

    def advance(self, step0: int, n: int) -> None:
        self._check(self.lib.pfcpu_advance(self.h, int(step0), int(n)))

    def get(self, comp: str) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.npdt)
        self._check(self.lib.pfcpu_read_field(self.h, "Ex Ey Ez Hx Hy Hz".split().index(comp),
                                              out.ctypes.data))
        return out

    def dft(self, name: str) -> Dict[str, np.ndarray]:
        mon, comps, count, n_freq = self.dft_slots[name]
        out = {}
        for slot, c in enumerate(comps):
            buf = np.empty((n_freq,) + tuple(count), dtype=np.complex128)
            self._check(self.lib.pfcpu_dft_read(self.h, mon, slot, buf.ctypes.data))
            out[c] = buf
        return out

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.pfcpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
