"""ctypes binding of libmct.so, the C ABI declared in include/mct.h.

The library is required: there is no CPU fallback.  ``lib()`` raises
``RuntimeError`` when it is missing, and every failing call raises
``MctError`` carrying the library's message.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmct.so"

MCT_OK = 0
MCT_ERR_INVALID = -1
MCT_ERR_DEVICE = -2
MCT_ERR_UNSUPPORTED = -3
MCT_ERR_NOMEM = -4

MCT_F32, MCT_F16, MCT_BF16 = 0, 1, 2
MCT_PLAIN, MCT_QUAD, MCT_PAIR, MCT_ROWS = 0, 1, 2, 3
MCT_FILTER_F32_DIRECT, MCT_FILTER_BF16_MFMA, MCT_FILTER_BF16X3_MFMA = 0, 1, 2

MCT_BP_AUTO, MCT_BP_DIRECT, MCT_BP_LDS = 0, 1, 2
MCT_FP_AUTO, MCT_FP_DIRECT, MCT_FP_CELLS = 0, 1, 2

ABI_VERSION = 11


class MctError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mct error {code}: {msg}")
        self.code = code


class Geometry(C.Structure):
    """Mirror of struct mct_geometry (include/mct.h)."""

    _fields_ = [
        ("img_dim", C.c_int32), ("img_dim_z", C.c_int32),
        ("pix", C.c_float), ("vox_h", C.c_float),
        ("cx", C.c_float), ("cy", C.c_float), ("cz", C.c_float),
        ("img_rot", C.c_float), ("view_mode", C.c_int32),
        ("n_views", C.c_int32), ("sid", C.c_float), ("sdd", C.c_float),
        ("total_scan_angle", C.c_float), ("short_scan", C.c_int32),
        ("cone", C.c_int32), ("curved", C.c_int32),
        ("n_u", C.c_int32), ("n_v", C.c_int32), ("n_v_full", C.c_int32),
        ("det_h", C.c_float),
        ("fpj_step", C.c_float), ("fpj_count_mode", C.c_int32),
        ("helical", C.c_int32), ("helical_pitch", C.c_float),
        ("apply_pmatrix", C.c_int32),
        ("angles", C.POINTER(C.c_float)), ("u", C.POINTER(C.c_float)),
        ("v", C.POINTER(C.c_float)), ("pmatrix", C.POINTER(C.c_float)),
        ("center_z", C.POINTER(C.c_float)), ("matrix_a", C.POINTER(C.c_float)),
        ("x_s", C.POINTER(C.c_float)),
    ]


class Helical(C.Structure):
    """Mirror of struct mct_helical (include/mct.h)."""

    _fields_ = [("n_u", C.c_int32), ("n_rows", C.c_int32), ("n_views", C.c_int32),
                ("view_count", C.c_int32), ("det_offset_vertical", C.c_float), ("det_h", C.c_float),
                ("sdd", C.c_float), ("mag", C.c_float), ("dis_per_view", C.c_float)]


class BpJob(C.Structure):
    """Mirror of struct mct_bp_job (include/mct.h)."""

    _fields_ = [("plan", C.c_void_p), ("filt", C.c_void_p), ("dtype", C.c_int32),
                ("layout", C.c_int32), ("vol", C.c_void_p)]


_P = C.c_void_p
_I32, _I64, _F32, _F64, _U64 = C.c_int32, C.c_int64, C.c_float, C.c_double, C.c_uint64

# name -> (restype, argtypes); every symbol include/mct.h declares.
SIGNATURES = {
    "mct_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "mct_destroy": (C.c_int, [_P]),
    "mct_set_stream": (C.c_int, [_P, _P]),
    "mct_synchronize": (C.c_int, [_P]),
    "mct_last_error": (C.c_char_p, []),
    "mct_abi_version": (C.c_int, []),
    "mct_plan_create": (C.c_int, [_P, C.POINTER(Geometry), C.POINTER(_P)]),
    "mct_plan_destroy": (C.c_int, [_P]),
    "mct_weight_sinogram": (C.c_int, [_P, _P, C.POINTER(_F32), _I32]),
    "mct_copy_views": (C.c_int, [_P, _P, _P, _I32, _I32, _I32, _I32, _I32, _P]),
    "mct_weight_views": (C.c_int, [_P, _P, _I32, _I32, C.POINTER(_F32), _I32]),
    "mct_filter_vertical": (C.c_int, [_P, _P, _P, _I32, _I32, _I32, C.POINTER(_F32), _F32]),
    "mct_filter_rows": (C.c_int, [_P, _P, _P, _I64, _I32, C.POINTER(_F32), _F32, _I32]),
    "mct_weight_filter_rows": (C.c_int, [_P, _P, _P, _I32, _I32, C.POINTER(_F32), _I32,
                                         C.POINTER(_F32), _F32, _I32]),
    "mct_weight_filter_pack": (C.c_int, [_P, _P, _P, _I32, _I32, C.POINTER(_F32), _I32,
                                         C.POINTER(_F32), _F32, _I32]),
    "mct_pack_sinogram": (C.c_int, [_P, _P, _I32, _I32, _I32, _P, _I32, _I32]),
    "mct_packed_bytes": (_I64, [_I32, _I32, _I32, _I32, _I32]),
    "mct_backproject_fbp": (C.c_int, [_P, _P, _I32, _I32, _P]),
    "mct_backproject_views": (C.c_int, [_P, _P, _I32, _I32, _I32, _I32, _I32, _P]),
    "mct_set_bp_tile": (C.c_int, [_P, _I32, _I32]),
    "mct_set_bp_kernel": (C.c_int, [_P, _I32]),
    "mct_bp_last_kernel": (C.c_int, [_P]),
    "mct_set_fp_kernel": (C.c_int, [_P, _I32]),
    "mct_fp_last_kernel": (C.c_int, [_P]),
    "mct_set_fp_split": (C.c_int, [_P, _I32]),
    "mct_set_frozen": (C.c_int, [_P, _I32]),
    "mct_backproject_batch": (C.c_int, [_P, C.POINTER(BpJob), _I32]),
    "mct_bp_batch_create": (C.c_int, [_P, C.POINTER(BpJob), _I32, C.POINTER(_P)]),
    "mct_bp_batch_launch": (C.c_int, [_P]),
    "mct_bp_batch_destroy": (C.c_int, [_P]),
    "mct_forward_project": (C.c_int, [_P, _P, _P, _P, _I32]),
    "mct_ray_mask": (C.c_int, [_P, _P]),
    "mct_add_noise": (C.c_int, [_P, _P, _I64, _F32, _U64]),
    "mct_helical_rebin": (C.c_int, [_P, C.POINTER(Helical), _P, _P, _I32, _P]),
    "mct_backproject_ir": (C.c_int, [_P, _P, _P, _P]),
    "mct_dot_f64": (C.c_int, [_P, _P, _P, _I64, _P]),
    "mct_dot_f64_noalloc": (C.c_int, [_P, _P, _P, _I64, _P]),
    "mct_tv_weight": (C.c_int, [_P, _P, _P, _I32, _I32, _I32, _I32, _F64, _P]),
    "mct_dtwd": (C.c_int, [_P, _P, _P, _I32, _I32, _I32, _I32, _F32, _P, _P]),
    "mct_axpby": (C.c_int, [_P, _F32, _P, _F32, _P, _P, _I64]),
    "mct_cg_update": (C.c_int, [_P, _P, _P, _P, _P, _P, _I64]),
    "mct_cg_direction": (C.c_int, [_P, _P, _P, _P, _I64]),
    "mct_sq_diff_f64": (C.c_int, [_P, _P, _P, _P, _I64, _P]),
    "mct_tv_value": (C.c_int, [_P, _P, _P, _I32, _I32, _I32, _I32, _F64, _P]),
}

_lib = None


def load(path: Path | str = LIB_PATH) -> C.CDLL:
    """Loads the library and declares every signature (no GPU call)."""
    path = Path(path)
    if not path.exists():
        raise RuntimeError(
            f"{path} is missing: build it with `python -m reconstruction_program_taichi_amd._build`")
    lib = C.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.mct_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{path}: ABI version {lib.mct_abi_version()} != {ABI_VERSION}")
    return lib


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def check(rc: int) -> None:
    if rc != MCT_OK:
        msg = lib().mct_last_error()
        raise MctError(rc, msg.decode() if msg else "")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
