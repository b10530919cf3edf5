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
        for T, ct in (("f32", C.c_float), ("f64", C.c_double)):
            pt = C.POINTER(ct)
            fn = getattr(_lib, f"fp_oracle_{T}")
            fn.argtypes = ([pt, C.c_int, C.c_int] + [ct] * 5 + [pt, ct, ct, pt, C.c_int, pt, C.c_int, pt,
                           C.c_int, C.c_int, C.c_int, ct, C.c_int, ct, pt])
            fn.restype = C.c_int
            fn = getattr(_lib, f"bpir_oracle_{T}")
            fn.argtypes = ([pt, C.c_int, C.c_int, C.c_int, pt, pt, ct, ct, pt, pt, C.c_int, C.c_int]
                           + [ct] * 5 + [pt, ct, ct, ct, C.c_int, pt])
            fn.restype = C.c_int
        _lib.filter_oracle_f32.argtypes = [fp, C.c_int64, C.c_int, fp, C.c_float, fp]
        _lib.filter_oracle_f32.restype = C.c_int
        _lib.oracle_set_threads.argtypes = [C.c_int]
        _lib.oracle_set_threads.restype = None
        _lib.oracle_max_threads.argtypes = []
        _lib.oracle_max_threads.restype = C.c_int
    return _lib


def set_threads(n: int) -> None:
    """OpenMP threads of the C restatements from now on."""
    lib().oracle_set_threads(int(n))


def max_threads() -> int:
    return int(lib().oracle_max_threads())


def filter_rows(x, taps, scale):
    """FP32 ramp filter of every row of x (last axis), as recon.filter_rows
    with dtype float32."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    n = f.shape[-1]
    t = np.ascontiguousarray(taps, dtype=np.float32)
    assert t.size == 2 * n - 1
    out = np.empty_like(f)
    lib().filter_oracle_f32(_p(f), f.size // n, n, _p(t), float(scale), _p(out))
    return out


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


def _arr(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _pt(a):
    ct = C.c_float if a.dtype == np.float32 else C.c_double
    return a.ctypes.data_as(C.POINTER(ct))


def forward_project(vol, g, oversample=1, ray_mask=None, count_mode=0, dtype=np.float32):
    """recon.forward_project (run_mgfpj.py:463-616) through fp_oracle.c:
    flat or curved detector, cone or fan, no P-matrix; same arithmetic and
    expression order.  vol (nz, ny, nx) -> (rows, views, n_u)."""
    if g.get("matrix_a") is not None:
        raise ValueError("the C forward projector has no P-matrix branch")
    T = np.dtype(dtype).type
    vol = _arr(vol, dtype)
    nz, _, nx = vol.shape
    ang, u, v = _arr(g["angles"], dtype), _arr(g["u"], dtype), _arr(g["v"], dtype)
    nrows, nviews, nuo = v.shape[0], ang.shape[0], u.shape[0]
    czs = g.get("center_z")
    czs = None if czs is None else _arr(czs, dtype)
    out = np.empty((nrows, nviews, nuo), dtype=dtype)
    fn = getattr(lib(), "fp_oracle_f32" if dtype == np.float32 else "fp_oracle_f64")
    fn(_pt(vol), nx, nz, T(g["pix"]), T(g.get("vox_h", 0.0)), T(g.get("cx", 0.0)), T(g.get("cy", 0.0)),
       T(g.get("cz", 0.0)), None if czs is None else _pt(czs), T(g["sid"]), T(g["sdd"]), _pt(ang),
       nviews, _pt(u), nuo, _pt(v), nrows, int(bool(g.get("cone", False))),
       int(bool(g.get("curved", False))), T(g.get("fpj_step", 0.2)), int(count_mode),
       T(g.get("zpv", 0.0)), _pt(out))
    if oversample > 1:
        b = out.reshape(nrows, nviews, nuo // oversample, oversample)
        s = np.zeros(b.shape[:3], dtype=dtype)
        for i in range(oversample):
            s = s + b[..., i]
        out = s / T(oversample)
    if ray_mask is not None:
        out = np.where(np.asarray(ray_mask) == 1, out, T(0))
    return out.astype(dtype)


def backproject_ir(sgm, g, ray_mask=None, dtype=np.float32):
    """recon.backproject_ir (run_mgfbp_iterative_recon_IRN_CGLS.py:372-481)
    through fp_oracle.c: axial view, flat detector, no P-matrix."""
    from .recon import view_table

    if g.get("pmatrix") is not None or g.get("curved", False):
        raise ValueError("the C IR back-projector covers the flat, non-P-matrix geometry")
    T = np.dtype(dtype).type
    f = _arr(sgm, dtype)
    if ray_mask is not None:
        f = _arr(f * np.asarray(ray_mask, dtype=dtype), dtype)
    nrows, nviews, nu = f.shape
    rot = g.get("img_rot", 0.0)
    c, s, _ = view_table(np.asarray(g["angles"], dtype=dtype), rot, dtype)
    c, s = _arr(c, dtype), _arr(s, dtype)
    cr, sr = T(np.cos(np.float64(T(rot)))), T(np.sin(np.float64(T(rot))))
    u, v = _arr(g["u"], dtype), _arr(g["v"], dtype)
    nx, nz = g["img_dim"], g["img_dim_z"]
    czs = g.get("center_z")
    czs = None if czs is None else _arr(czs, dtype)
    out = np.empty((nz, nx, nx), dtype=dtype)
    fn = getattr(lib(), "bpir_oracle_f32" if dtype == np.float32 else "bpir_oracle_f64")
    fn(_pt(f), nrows, nviews, nu, _pt(c), _pt(s), cr, sr, _pt(u), _pt(v), nx, nz, T(g["pix"]),
       T(g.get("vox_h", 0.0)), T(g.get("cx", 0.0)), T(g.get("cy", 0.0)), T(g.get("cz", 0.0)),
       None if czs is None else _pt(czs), T(g["sid"]), T(g["sdd"]), T(g.get("det_h", 1.0)),
       int(bool(g.get("cone", False))), _pt(out))
    return out
