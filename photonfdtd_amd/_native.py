"""ctypes binding of ``libpfhip.so`` (declarations in ``include/pfhip.h``).

The loader fails loudly: there is no CPU fallback. It plays the role of the
reference's ``_rust_module()`` gate (``src/photonfdtd/rustbackend.py:45-58``),
including its ``RuntimeError`` with build instructions.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpfhip.so"
ABI_VERSION = 7
PF_FLAG_GENERIC = 1

PF_F32, PF_F64 = 0, 1
KERNEL_CLASSES = ("yee_h", "yee_e", "yee_fused", "aux", "yee_pair", "dft_fold")
COMP_CODE = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}

_i32, _i64, _f64, _vp = C.c_int32, C.c_int64, C.c_double, C.c_void_p
_P = C.POINTER


class PfBox(C.Structure):
    _fields_ = [("start", _i64 * 3), ("count", _i64 * 3), ("stride", _i64 * 3)]


class PfAdeDesc(C.Structure):
    _fields_ = [("n_cells", _i64), ("cells", _P(_i64)), ("n_poles", _i32), ("abc", _P(_f64)),
                ("inv_eps", _f64)]


class PfDftDesc(C.Structure):
    _fields_ = [("n_comp", _i32), ("comps", _i32 * 6), ("box", PfBox), ("n_freq", _i32),
                ("interval", _i32), ("n_samples", _i64), ("phase_e", _P(_f64)),
                ("phase_h", _P(_f64)), ("dt", _f64)]


class PfFluxDesc(C.Structure):
    _fields_ = [("axis", _i32), ("index", _i64), ("dA", _f64), ("dt", _f64)]


class PfDesc(C.Structure):
    _fields_ = [("dtype", _i32), ("exact", _i32), ("shape", _i64 * 3), ("cell", _f64 * 3),
                ("ch", _f64), ("ce", _vp),
                ("b_e", _P(_f64) * 3), ("c_e", _P(_f64) * 3), ("b_h", _P(_f64) * 3),
                ("c_h", _P(_f64) * 3),
                ("n_steps", _i64), ("n_src", _i32), ("src_comp", _P(_i32)), ("src_cell", _P(_i64)),
                ("src_amp", _P(_f64)), ("src_wave", _P(_i32)), ("n_wave", _i32),
                ("wave", _P(_f64)), ("n_ade", _i32), ("ade", _P(PfAdeDesc)), ("n_dft", _i32),
                ("dft", _P(PfDftDesc)), ("n_flux", _i32), ("flux", _P(PfFluxDesc)),
                ("flags", _i32), ("ce_y", _vp), ("ce_z", _vp)]


class PfNnConv(C.Structure):
    _fields_ = [("n", _i32), ("h", _i32), ("w", _i32), ("cin", _i32), ("cout", _i32),
                ("ksize", _i32), ("stride", _i32), ("relu", _i32), ("out_f32", _i32)]


class PfInfo(C.Structure):
    _fields_ = [("path", _i32), ("zpad", _i32), ("zrow", _i32), ("march_nw", _i32),
                ("pml_nw", _i32), ("cpml_lists", _i32), ("tiles_bulk", _i64), ("tiles_cpml", _i64),
                ("cells_bulk", _i64), ("cells_cpml", _i64), ("pair_tiles_bulk", _i64),
                ("pair_tiles_cpml", _i64), ("dft_ring", _i64)]


class PfCotangent(C.Structure):
    _fields_ = [("mon", _i32), ("slot", _i32), ("g", _P(_f64))]


# Exported symbols and their signatures; tests check every one is present.
SIGNATURES = {
    "pf_create": (_i32, [_P(PfDesc), _P(_vp)]),
    "pf_destroy": (_i32, [_vp]),
    "pf_advance": (_i32, [_vp, _i64, _i64]),
    "pf_read_field": (_i32, [_vp, _i32, _vp]),
    "pf_read_box": (_i32, [_vp, _i32, _P(PfBox), _vp]),
    "pf_read_box_q": (_i32, [_vp, _i32, _P(PfBox), _i32, _vp, _P(_f64)]),
    "pf_write_field": (_i32, [_vp, _i32, _vp]),
    "pf_dft_read": (_i32, [_vp, _i32, _i32, _P(_f64)]),
    "pf_dft_flush": (_i32, [_vp]),
    "pf_stepper_info": (_i32, [_vp, _P(PfInfo)]),
    "pf_flux_read": (_i32, [_vp, _P(_f64)]),
    "pf_checkpoint_enable": (_i32, [_vp, _i32]),
    "pf_adjoint": (_i32, [_vp, _i32, _P(PfCotangent), _P(_f64)]),
    "pf_timing_enable": (_i32, [_vp, _i32]),
    "pf_timing_read": (_i32, [_vp, _P(_f64), _P(_i64)]),
    "pf_device_info": (_i32, [C.c_char_p, _i32, _P(_i64), _P(_i64)]),
    "pf_last_error": (C.c_char_p, []),
    "pf_abi_version": (_i32, []),
    "pf_batch_create": (_i32, [_P(PfDesc), _P(_i32), _i32, _P(_vp)]),
    "pf_batch_destroy": (_i32, [_vp]),
    "pf_batch_configure": (_i32, [_vp, _i32, _i32, _i32]),
    "pf_batch_reset": (_i32, [_vp]),
    "pf_batch_persistent_grid": (_i32, [_vp, _i32, _P(_i32), _P(_i32), _P(_i32)]),
    "pf_batch_advance": (_i32, [_vp, _i64]),
    "pf_batch_last_ms": (_i32, [_vp, _P(_f64)]),
    "pf_batch_write_ce": (_i32, [_vp, _i32, _vp]),
    "pf_nn_conv_forward": (_i32, [_P(PfNnConv), _vp, _vp, _vp, _vp, _vp]),
    "pf_nn_conv_backward": (_i32, [_P(PfNnConv), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pf_nn_conv_backward_noalloc": (_i32, [_P(PfNnConv), _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                           _vp]),
    "pf_nn_mse": (_i32, [_i32, _vp, _vp, _vp, _vp, _vp]),
    "pf_nn_adam": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, C.c_float, C.c_float, C.c_float,
                          C.c_float, _i32, _vp]),
    "pf_nn_to_bf16": (_i32, [_i64, _vp, _vp, _vp]),
    "pf_nn_last_error": (C.c_char_p, []),
    "pf_batch_dft_read": (_i32, [_vp, _i32, _i32, _i32, _P(_f64)]),
    "pf_batch_read_field": (_i32, [_vp, _i32, _i32, _vp]),
}

_LIB = None


def load_library(path: os.PathLike = LIB_PATH) -> C.CDLL:
    """Load and type the library; ``RuntimeError`` when it is missing or stale."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = Path(path)
    if not path.exists():
        raise RuntimeError(
            f"backend='hip' needs the compiled library {path}, which is missing. "
            "Build it with 'python -m photonfdtd_amd.build' (hipcc, gfx950).")
    lib = C.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.pf_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{path} has ABI {lib.pf_abi_version()}, expected {ABI_VERSION}; "
                           "rebuild with 'python -m photonfdtd_amd.build'")
    _LIB = lib
    return lib


def check(lib: C.CDLL, status: int, what: str) -> None:
    """Map a non-zero status to ``RuntimeError("HIP error: ...")`` as the
    reference maps CUDA failures (``rust/src/cuda.rs:267-269``)."""
    if status != 0:
        msg = lib.pf_last_error().decode(errors="replace")
        if status == -1:
            raise ValueError(f"{what}: {msg}")
        raise RuntimeError(f"{what} failed ({status}): {msg}")


def as_ptr(arr, ctype):
    return arr.ctypes.data_as(_P(ctype))
