"""ctypes loader for libbp_oracle.so (test infrastructure only)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "libbp_oracle.so"
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            from .build import build
            build()
        _lib = C.CDLL(str(LIB))
        fp = C.POINTER(C.c_float)
        _lib.bp_oracle_fbp.argtypes = [fp, C.c_int, C.c_int, C.c_int, fp, fp, fp, fp, fp, C.c_int,
                                       C.c_int, C.c_int, C.c_int] + [C.c_float] * 9 + [C.c_int, fp]
        _lib.bp_oracle_fbp.restype = C.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def backproject_fbp(filt, g, z0=0, z1=None):
    """Slices z0..z1-1 of the FP32 reference-order back-projection."""
    from .recon import div_factor, view_table

    f = np.ascontiguousarray(filt, dtype=np.float32)
    nrows, nviews, nu = f.shape
    nx, nz = g["img_dim"], g["img_dim_z"]
    z1 = nz if z1 is None else z1
    c, s, d = view_table(np.asarray(g["angles"], np.float32), g.get("img_rot", 0.0), np.float32)
    u = np.ascontiguousarray(g["u"], dtype=np.float32)
    v = np.ascontiguousarray(g["v"], dtype=np.float32)
    out = np.zeros((z1 - z0, nx, nx), dtype=np.float32)
    div = float(div_factor(g["total_scan_angle"], g.get("short_scan", False), np.float32))
    lib().bp_oracle_fbp(_p(f), nrows, nviews, nu, _p(c), _p(s), _p(d), _p(u), _p(v), nx, nz, z0, z1,
                        g["pix"], g.get("vox_h", 0.0), g.get("cx", 0.0), g.get("cy", 0.0),
                        g.get("cz", 0.0), g["sid"], g["sdd"], g.get("det_h", 1.0), div,
                        int(bool(g.get("cone", False))), _p(out))
    return out
