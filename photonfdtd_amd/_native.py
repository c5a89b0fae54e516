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
ABI_VERSION = 9
PF_FLAG_GENERIC = 1
PF_FLAG_STREAM = 2
PF_FLAG_DENSE64 = 4

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
                ("flags", _i32), ("ce_y", _vp), ("ce_z", _vp), ("t_block", _i32),
                ("tile_rows", _i32)]


class PfView(C.Structure):
    _fields_ = [("ptr", _vp), ("dtype", _i32), ("ndim", _i32), ("shape", _i64 * 6),
                ("strides", _i64 * 6), ("stream", C.c_uint64)]


class PfNnConv(C.Structure):
    _fields_ = [("n", _i32), ("h", _i32), ("w", _i32), ("cin", _i32), ("cout", _i32),
                ("ksize", _i32), ("stride", _i32), ("relu", _i32), ("out_f32", _i32)]


class PfInfo(C.Structure):
    _fields_ = [("path", _i32), ("zpad", _i32), ("zrow", _i32), ("march_nw", _i32),
                ("pml_nw", _i32), ("cpml_lists", _i32), ("tiles_bulk", _i64), ("tiles_cpml", _i64),
                ("cells_bulk", _i64), ("cells_cpml", _i64), ("pair_tiles_bulk", _i64),
                ("pair_tiles_cpml", _i64), ("dft_ring", _i64), ("post_fused", _i64)]


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
    "pf_field_ptr": (_i32, [_vp, _i32, _P(PfView)]),
    "pf_dft_ptr": (_i32, [_vp, _i32, _P(PfView)]),
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


# ---- zero-copy handoff of device arrays to torch (DLPack) -------------------
# pf_field_ptr / pf_dft_ptr describe a device array of the library; to_torch
# wraps it in a DLPack capsule, which torch consumes without a copy. The
# capsule's deleter keeps `owner` (the engine that owns the memory) alive
# until torch frees the tensor.


class _DLDevice(C.Structure):
    _fields_ = [("device_type", C.c_int32), ("device_id", C.c_int32)]


class _DLDataType(C.Structure):
    _fields_ = [("code", C.c_uint8), ("bits", C.c_uint8), ("lanes", C.c_uint16)]


class _DLTensor(C.Structure):
    _fields_ = [("data", _vp), ("device", _DLDevice), ("ndim", C.c_int32),
                ("dtype", _DLDataType), ("shape", _P(_i64)), ("strides", _P(_i64)),
                ("byte_offset", C.c_uint64)]


class _DLManagedTensor(C.Structure):
    pass


_DELETER = C.CFUNCTYPE(None, _P(_DLManagedTensor))
_DLManagedTensor._fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", _vp),
                             ("deleter", _DELETER)]
_KDL_ROCM = 10        # DLPack device type of AMD GPUs
_KDL_FLOAT = 2
_LIVE = {}            # manager_ctx -> (managed tensor, shape, strides, owner)


@_DELETER
def _dl_delete(ptr):
    _LIVE.pop(ptr.contents.manager_ctx, None)


def to_torch(view: PfView, owner, device_index: int = 0):
    """A torch tensor aliasing the device array described by ``view`` (no
    copy). Valid as long as the library keeps the array where it is (see
    pf_field_ptr / pf_dft_ptr); ``owner`` stays referenced until torch frees
    the tensor."""
    import torch
    from torch.utils import dlpack
    nd = int(view.ndim)
    shape = (_i64 * nd)(*view.shape[:nd])
    strides = (_i64 * nd)(*view.strides[:nd])
    m = _DLManagedTensor()
    m.dl_tensor.data = view.ptr
    m.dl_tensor.device = _DLDevice(_KDL_ROCM, int(device_index))
    m.dl_tensor.ndim = nd
    m.dl_tensor.dtype = _DLDataType(_KDL_FLOAT, 32 if view.dtype == PF_F32 else 64, 1)
    m.dl_tensor.shape = shape
    m.dl_tensor.strides = strides
    m.dl_tensor.byte_offset = 0
    key = id(m)
    m.manager_ctx = key
    m.deleter = _dl_delete
    _LIVE[key] = (m, shape, strides, owner)
    new_capsule = C.pythonapi.PyCapsule_New
    new_capsule.restype = C.py_object
    new_capsule.argtypes = [_vp, C.c_char_p, _vp]
    cap = new_capsule(C.addressof(m), b"dltensor", None)
    return dlpack.from_dlpack(cap)
