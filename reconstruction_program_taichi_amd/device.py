"""Device plumbing: the C-ABI context and geometry plans over torch tensors.

torch (ROCm) only allocates device memory and supplies the stream; every
computation is a HIP kernel of libmct.so.  All calls are queued on torch's
current stream, so they order with torch's own work on that stream.
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np
import torch

from . import _abi
from ._abi import call
from .geometry import GeometrySpec

DTYPES = {"f32": _abi.MCT_F32, "f16": _abi.MCT_F16, "bf16": _abi.MCT_BF16}
LAYOUTS = {"plain": _abi.MCT_PLAIN, "quad": _abi.MCT_QUAD, "pair": _abi.MCT_PAIR, "rows": _abi.MCT_ROWS}
FILTER_METHODS = {"f32": _abi.MCT_FILTER_F32_DIRECT, "bf16": _abi.MCT_FILTER_BF16_MFMA,
                  "bf16x3": _abi.MCT_FILTER_BF16X3_MFMA}


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


def _f32(t: torch.Tensor) -> int:
    if t.dtype != torch.float32:
        raise ValueError(f"expected float32, got {t.dtype}")
    return _ptr(t)


def _shape(t: torch.Tensor, want: tuple, name: str) -> None:
    """The kernels size every launch from the plan's geometry, so a buffer of
    another shape would be read or written out of bounds: refuse it."""
    if tuple(t.shape) != tuple(want):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, the plan expects {tuple(want)}")


def _packed_bytes(rows: int, views: int, nu: int, dtype: str, layout: str) -> int:
    return int(_abi.lib().mct_packed_bytes(rows, views, nu, DTYPES[dtype], LAYOUTS[layout]))


def _host_f32(a) -> tuple[np.ndarray, C.POINTER(C.c_float)]:
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


class Context:
    """An mct_ctx bound to one device, queued on torch's current stream."""

    def __init__(self, device: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: the reconstruction kernels need a ROCm device")
        self.device = int(device)
        h = C.c_void_p()
        call("mct_create", self.device, C.byref(h))
        self.h = h
        self.bp_kernel = "auto"  # as last set with set_bp_kernel / set_fp_kernel
        self.fp_kernel = "auto"
        self.bind_stream()

    def bind_stream(self, stream: torch.cuda.Stream | None = None) -> None:
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        call("mct_set_stream", self.h, C.c_void_p(s.cuda_stream))

    def synchronize(self) -> None:
        call("mct_synchronize", self.h)

    def copy_views(self, host_sgm: torch.Tensor, view_begin: int, view_count: int,
                   out: torch.Tensor, stream: torch.cuda.Stream | None = None) -> torch.Tensor:
        """Views [view_begin, view_begin + view_count) of a host FP32 sinogram
        (rows, views, n_u) into the device block out (rows, view_count, n_u),
        asynchronously on `stream` (None: the context's stream).  host_sgm
        must stay alive until the copy has run; pinned memory lets it overlap
        kernels."""
        if host_sgm.device.type != "cpu" or host_sgm.dtype != torch.float32 or not host_sgm.is_contiguous():
            raise ValueError("host_sgm must be a contiguous float32 CPU tensor")
        rows, views, nu = host_sgm.shape
        if tuple(out.shape) != (rows, view_count, nu):
            raise ValueError(f"out must have shape {(rows, view_count, nu)}, got {tuple(out.shape)}")
        call("mct_copy_views", self.h, C.c_void_p(stream.cuda_stream) if stream is not None else None,
             C.c_void_p(host_sgm.data_ptr()), rows, views, nu, int(view_begin), int(view_count),
             _f32(out))
        return out

    def set_bp_tile(self, zc: int, batch: int = 0) -> None:
        call("mct_set_bp_tile", self.h, int(zc), int(batch))

    BP_KERNELS = {"auto": _abi.MCT_BP_AUTO, "direct": _abi.MCT_BP_DIRECT, "lds": _abi.MCT_BP_LDS}

    def set_bp_kernel(self, kind: str) -> None:
        """Cone-beam back-projector: 'auto', 'direct' (per-update gathers) or
        'lds' (detector windows staged in LDS)."""
        call("mct_set_bp_kernel", self.h, self.BP_KERNELS[kind])
        self.bp_kernel = kind

    FP_KERNELS = {"auto": _abi.MCT_FP_AUTO, "direct": _abi.MCT_FP_DIRECT,
                  "cells": _abi.MCT_FP_CELLS}

    def set_fp_kernel(self, kind: str) -> None:
        """Forward projector: 'auto', 'direct' (voxel gathers from the volume
        as given) or 'cells' (bilinear cells in two orders, in context
        scratch)."""
        call("mct_set_fp_kernel", self.h, self.FP_KERNELS[kind])
        self.fp_kernel = kind

    def set_fp_split(self, waves_per_cu: int = 0) -> None:
        """Cell projector: split each ray's steps over several wavefronts when a
        launch has fewer than waves_per_cu wavefronts per CU (0: default)."""
        call("mct_set_fp_split", self.h, int(waves_per_cu))

    def freeze(self, frozen: bool = True) -> None:
        """While frozen, calls that would grow, reallocate or rewrite the
        context's scratch or cached tables (which captured graphs address)
        raise MctError instead; calls of the captured sizes and taps run."""
        call("mct_set_frozen", self.h, int(bool(frozen)))

    def fp_last_kernel(self) -> str:
        k = _abi.lib().mct_fp_last_kernel(self.h)
        return {0: "none", 1: "direct", 2: "cells"}.get(k, str(k))

    def bp_last_kernel(self) -> str:
        """Which back-projector the last FBP back-projection on this context ran."""
        k = _abi.lib().mct_bp_last_kernel(self.h)
        return {0: "none", 1: "direct", 2: "lds", 3: "fan", 4: "generic"}.get(k, str(k))

    def close(self) -> None:
        if getattr(self, "h", None):
            _abi.lib().mct_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- context-level operations -------------------------------------------------
    def filter_rows(self, inp: torch.Tensor, out: torch.Tensor, taps: np.ndarray, scale: float,
                    method: str = "f32") -> torch.Tensor:
        n = int(inp.shape[-1])
        if taps.shape[0] != 2 * n - 1:
            raise ValueError(f"expected {2 * n - 1} taps, got {taps.shape[0]}")
        if out.shape != inp.shape:
            raise ValueError("out must have the shape of inp")
        keep, tp = _host_f32(taps)
        call("mct_filter_rows", self.h, _f32(inp), _f32(out), inp.numel() // n, n, tp,
             float(scale), FILTER_METHODS[method])
        return out

    def filter_vertical(self, inp: torch.Tensor, out: torch.Tensor, kern: np.ndarray,
                        det_h: float) -> torch.Tensor:
        rows, views, nu = inp.shape
        if out.shape != inp.shape:
            raise ValueError("out must have the shape of inp")
        if np.asarray(kern).shape[0] != 2 * rows - 1:
            raise ValueError(f"expected {2 * rows - 1} vertical taps, got {np.asarray(kern).shape[0]}")
        keep, kp = _host_f32(kern)
        call("mct_filter_vertical", self.h, _f32(inp), _f32(out), rows, views, nu, kp, float(det_h))
        return out

    def pack(self, inp: torch.Tensor, dtype: str, layout: str) -> torch.Tensor:
        rows, views, nu = inp.shape
        nbytes = _abi.lib().mct_packed_bytes(rows, views, nu, DTYPES[dtype], LAYOUTS[layout])
        if nbytes <= 0:
            raise ValueError(f"cannot pack shape {tuple(inp.shape)} as {layout}")
        out = torch.empty(int(nbytes), dtype=torch.uint8, device=inp.device)
        return self.pack_into(inp, dtype, layout, out)

    def pack_into(self, inp: torch.Tensor, dtype: str, layout: str, out: torch.Tensor) -> torch.Tensor:
        """pack() into a preallocated byte buffer (no allocation: capturable)."""
        rows, views, nu = inp.shape
        nbytes = _abi.lib().mct_packed_bytes(rows, views, nu, DTYPES[dtype], LAYOUTS[layout])
        if nbytes <= 0 or out.numel() * out.element_size() < nbytes or not out.is_contiguous():
            raise ValueError(f"output buffer too small for packing {tuple(inp.shape)} as {layout}")
        call("mct_pack_sinogram", self.h, _f32(inp), rows, views, nu, _ptr(out), DTYPES[dtype],
             LAYOUTS[layout])
        return out

    def _job_array(self, jobs):
        arr = (_abi.BpJob * max(len(jobs), 1))()
        for i, (plan, filt, dtype, layout, vol) in enumerate(jobs):
            if plan.ctx is not self:
                raise ValueError("plan belongs to another context")
            plan._check_packed(filt, dtype, layout, plan.spec.n_views)
            plan._check_vol(vol)
            arr[i] = _abi.BpJob(plan.h.value, _ptr(filt), DTYPES[dtype], LAYOUTS[layout], _f32(vol))
        return arr

    def backproject_batch(self, jobs) -> None:
        """jobs: iterable of (plan, packed filtered sinogram, dtype, layout, vol)."""
        jobs = list(jobs)
        call("mct_backproject_batch", self.h, self._job_array(jobs), len(jobs))

    def prepare_batch(self, jobs) -> "BpBatch":
        """A batched back-projection prepared once (job table on the device)
        and launched many times; BpBatch.launch() only launches the kernel, so
        it can be captured into a graph."""
        jobs = list(jobs)
        h = C.c_void_p()
        call("mct_bp_batch_create", self.h, self._job_array(jobs), len(jobs), C.byref(h))
        return BpBatch(self, h, jobs)

    def helical_rebin(self, sgm: torch.Tensor, z_pos, view_count: int, det_offset_vertical: float,
                      det_h: float, sdd: float, mag: float, dis_per_view: float,
                      out: torch.Tensor | None = None) -> torch.Tensor:
        """Fan sinograms (len(z_pos), view_count, n_u) rebinned from a helical
        cone-beam sinogram (rows, views, n_u) for every slice position."""
        rows, views, nu = sgm.shape
        z, zp = _host_f32(z_pos)
        if out is None:
            out = torch.empty((z.shape[0], view_count, nu), dtype=torch.float32, device=sgm.device)
        if tuple(out.shape) != (z.shape[0], view_count, nu):
            raise ValueError(f"out has shape {tuple(out.shape)}, expected {(z.shape[0], view_count, nu)}")
        h = _abi.Helical(nu, rows, views, int(view_count), det_offset_vertical, det_h, sdd, mag, dis_per_view)
        call("mct_helical_rebin", self.h, C.byref(h), _f32(sgm), zp, int(z.shape[0]), _f32(out))
        return out

    def add_noise(self, sgm: torch.Tensor, photons: float, seed: int) -> torch.Tensor:
        call("mct_add_noise", self.h, _f32(sgm), sgm.numel(), float(photons), int(seed) & (2**64 - 1))
        return sgm

    def dot(self, a: torch.Tensor, b: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        call("mct_dot_f64", self.h, _f32(a), _f32(b), a.numel(), _ptr(out))
        return out

    def sq_diff(self, a, b, mb, out):
        call("mct_sq_diff_f64", self.h, _f32(a), _f32(b), _ptr(mb), a.numel(), _ptr(out))
        return out

    def tv_weight(self, x, prior, beta: float, out, couple_z: bool = True):
        nz, ny, nx = x.shape
        call("mct_tv_weight", self.h, _f32(x), _ptr(prior), nz, ny, nx, int(couple_z), float(beta),
             _f32(out))
        return out

    def tv_value(self, x, prior, beta: float, out, couple_z: bool = True):
        nz, ny, nx = x.shape
        call("mct_tv_value", self.h, _f32(x), _ptr(prior), nz, ny, nx, int(couple_z), float(beta),
             _ptr(out))
        return out

    def dtwd(self, x, wr, scale: float, base, out, couple_z: bool = True):
        nz, ny, nx = x.shape
        call("mct_dtwd", self.h, _f32(x), _f32(wr), nz, ny, nx, int(couple_z), float(scale),
             _ptr(base), _f32(out))
        return out

    def axpby(self, a: float, x, b: float, y, out):
        call("mct_axpby", self.h, float(a), _f32(x), float(b), _f32(y), _f32(out), x.numel())
        return out

    def cg_update(self, x, r, d, q, s):
        call("mct_cg_update", self.h, _f32(x), _f32(r), _f32(d), _f32(q), _ptr(s), x.numel())

    def cg_direction(self, d, r, s):
        call("mct_cg_direction", self.h, _f32(d), _f32(r), _ptr(s), d.numel())


class BpBatch:
    """A prepared batched back-projection (mct_bp_batch_*).  Keeps the jobs'
    plans and tensors alive as long as it exists."""

    def __init__(self, ctx: Context, h: C.c_void_p, jobs: list):
        self.ctx, self.h, self._jobs = ctx, h, jobs

    def launch(self) -> None:
        call("mct_bp_batch_launch", self.h)

    def close(self) -> None:
        if getattr(self, "h", None):
            _abi.lib().mct_bp_batch_destroy(self.h)
            self.h = None
            self._jobs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """An mct_plan: a geometry uploaded once, with its per-view tables."""

    def __init__(self, ctx: Context, spec: GeometrySpec):
        self.ctx = ctx
        self.spec = spec
        self._fallback_warned: set = set()
        st = spec.to_struct()
        h = C.c_void_p()
        call("mct_plan_create", ctx.h, C.byref(st), C.byref(h))
        self.h = h

    def close(self) -> None:
        if getattr(self, "h", None):
            _abi.lib().mct_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- kernel fallbacks the caller should know about ------------------------------
    def _warn_once(self, key: str, msg: str) -> None:
        if key not in getattr(self, "_fallback_warned", set()):
            self._fallback_warned = getattr(self, "_fallback_warned", set()) | {key}
            warnings.warn(msg, RuntimeWarning, stacklevel=3)

    def _report_fallback_bp(self, layout: str) -> None:
        """The automatic choice runs the LDS-staged cone back-projector unless
        its plan tables do not fit (window table over 16 GiB, or a detector
        window over the LDS budget): then the direct kernel runs, about 1.8x
        slower.  Said once per plan."""
        s = self.spec
        if (s.cone and layout in ("quad", "rows") and self.ctx.bp_kernel == "auto"
                and self.ctx.bp_last_kernel() == "direct"):
            self._warn_once("bp", "cone back-projection ran the direct-gather kernel: the LDS-staged "
                                  "kernel's tables do not fit this geometry (window table over 16 GiB "
                                  "or detector window over the LDS budget); expect about 1.8x the time")

    def _report_fallback_fp(self) -> None:
        """The automatic choice runs the record forward projector unless its
        records do not fit 32-bit offsets (about 4 GiB): then the direct
        kernel runs, several times slower.  Said once per plan."""
        if self.ctx.fp_kernel == "auto" and self.ctx.fp_last_kernel() == "direct":
            self._warn_once("fp", "forward projection ran the direct kernel: the volume's records do not "
                                  "fit 32-bit offsets (about 4 GiB); expect several times the time")

    # ---- shape checks against the plan's geometry (before any device call) --------
    def _check_sgm(self, t: torch.Tensor, name: str = "sinogram", views: int | None = None,
                   nu: int | None = None) -> None:
        s = self.spec
        _shape(t, (s.n_v, s.n_views if views is None else views, s.n_u if nu is None else nu), name)

    def _check_vol(self, t: torch.Tensor) -> None:
        s = self.spec
        _shape(t, (s.img_dim_z, s.img_dim, s.img_dim), "volume")

    def _check_packed(self, t: torch.Tensor, dtype: str, layout: str, views: int) -> None:
        s = self.spec
        need = _packed_bytes(s.n_v, views, s.n_u, dtype, layout)
        if need <= 0:
            raise ValueError(f"no {layout} packing for ({s.n_v}, {views}, {s.n_u})")
        have = t.numel() * t.element_size()
        if have < need or not t.is_contiguous():
            raise ValueError(f"packed sinogram holds {have} bytes, the plan needs {need} contiguous bytes")

    def _check_views(self, view_begin: int, view_count: int) -> None:
        if view_begin < 0 or view_count < 0 or view_begin + view_count > self.spec.n_views:
            raise ValueError(f"views [{view_begin}, {view_begin + view_count}) outside the plan's "
                             f"{self.spec.n_views} views")

    def weight(self, sgm: torch.Tensor, bh_coef=None) -> torch.Tensor:
        self._check_sgm(sgm)
        if bh_coef is None or len(bh_coef) == 0:
            call("mct_weight_sinogram", self.h, _f32(sgm), None, 0)
        else:
            keep, cp = _host_f32(bh_coef)
            call("mct_weight_sinogram", self.h, _f32(sgm), cp, len(keep))
        return sgm

    def backproject_fbp(self, filt: torch.Tensor, dtype: str, layout: str,
                        vol: torch.Tensor) -> torch.Tensor:
        self._check_packed(filt, dtype, layout, self.spec.n_views)
        self._check_vol(vol)
        call("mct_backproject_fbp", self.h, _ptr(filt), DTYPES[dtype], LAYOUTS[layout], _f32(vol))
        self._report_fallback_bp(layout)
        return vol

    def weight_views(self, sgm: torch.Tensor, view_begin: int, bh_coef=None) -> torch.Tensor:
        """weight() of a block (rows, view_count, n_u) holding views
        [view_begin, view_begin + view_count) of the scan."""
        n = int(sgm.shape[1])
        self._check_sgm(sgm, views=n)
        self._check_views(view_begin, n)
        if bh_coef is None or len(bh_coef) == 0:
            call("mct_weight_views", self.h, _f32(sgm), int(view_begin), n, None, 0)
        else:
            keep, cp = _host_f32(bh_coef)
            call("mct_weight_views", self.h, _f32(sgm), int(view_begin), n, cp, len(keep))
        return sgm

    def weight_filter(self, sgm: torch.Tensor, out: torch.Tensor, taps: np.ndarray, scale: float,
                      method: str, view_begin: int = 0, bh_coef=None) -> torch.Tensor:
        """weight_views(sgm, view_begin, bh_coef) then the context's
        filter_rows(sgm, out, taps, scale, method) in one pass, sgm left
        unweighted; MFMA methods only (bf16, bf16x3)."""
        rows, n, nu = sgm.shape
        self._check_sgm(sgm, views=n)
        self._check_sgm(out, "out", views=n)
        self._check_views(view_begin, n)
        if taps.shape[0] != 2 * nu - 1:
            raise ValueError(f"expected {2 * nu - 1} taps, got {taps.shape[0]}")
        keep_t, tp = _host_f32(taps)
        if bh_coef is None or len(bh_coef) == 0:
            call("mct_weight_filter_rows", self.h, _f32(sgm), _f32(out), int(view_begin), n, None, 0,
                 tp, float(scale), FILTER_METHODS[method])
        else:
            keep, cp = _host_f32(bh_coef)
            call("mct_weight_filter_rows", self.h, _f32(sgm), _f32(out), int(view_begin), n, cp,
                 len(keep), tp, float(scale), FILTER_METHODS[method])
        return out

    def weight_filter_pack(self, sgm: torch.Tensor, packed: torch.Tensor, taps: np.ndarray, scale: float,
                           method: str, view_begin: int = 0, bh_coef=None) -> torch.Tensor:
        """weight_filter(sgm, ...) then the context's pack(..., "f32", "rows")
        into packed, in one pass and without the FP32 filtered sinogram (the
        filter writes the ROWS taps itself); MFMA methods only, n_u - 1 not a
        multiple of 64."""
        rows, n, nu = sgm.shape
        self._check_sgm(sgm, views=n)
        self._check_views(view_begin, n)
        self._check_packed(packed, "f32", "rows", n)
        if taps.shape[0] != 2 * nu - 1:
            raise ValueError(f"expected {2 * nu - 1} taps, got {taps.shape[0]}")
        keep_t, tp = _host_f32(taps)
        if bh_coef is None or len(bh_coef) == 0:
            call("mct_weight_filter_pack", self.h, _f32(sgm), _ptr(packed), int(view_begin), n, None, 0,
                 tp, float(scale), FILTER_METHODS[method])
        else:
            keep, cp = _host_f32(bh_coef)
            call("mct_weight_filter_pack", self.h, _f32(sgm), _ptr(packed), int(view_begin), n, cp,
                 len(keep), tp, float(scale), FILTER_METHODS[method])
        return packed

    def backproject_views(self, filt: torch.Tensor, dtype: str, layout: str, view_begin: int,
                          view_count: int, vol: torch.Tensor, finish: bool = False) -> torch.Tensor:
        """Adds the back-projection of views [view_begin, view_begin +
        view_count) (filt: their packed block) to vol; finish on the last
        block clears the truncation marks."""
        self._check_views(view_begin, view_count)
        self._check_packed(filt, dtype, layout, view_count)
        self._check_vol(vol)
        call("mct_backproject_views", self.h, _ptr(filt), DTYPES[dtype], LAYOUTS[layout],
             int(view_begin), int(view_count), int(bool(finish)), _f32(vol))
        self._report_fallback_bp(layout)
        return vol

    def forward_project(self, vol: torch.Tensor, out: torch.Tensor, ray_mask=None,
                        oversample: int = 1) -> torch.Tensor:
        s = self.spec
        if oversample < 1 or s.n_u % oversample:
            raise ValueError(f"n_u ({s.n_u}) is not a multiple of oversample ({oversample})")
        self._check_vol(vol)
        self._check_sgm(out, "out", nu=s.n_u // oversample)
        if ray_mask is not None:
            self._check_sgm(ray_mask, "ray_mask", nu=s.n_u // oversample)
        call("mct_forward_project", self.h, _f32(vol), _f32(out),
             None if ray_mask is None else _f32(ray_mask), int(oversample))
        self._report_fallback_fp()
        return out

    def ray_mask(self, out: torch.Tensor) -> torch.Tensor:
        self._check_sgm(out, "ray_mask")
        call("mct_ray_mask", self.h, _f32(out))
        return out

    def backproject_ir(self, sgm: torch.Tensor, ray_mask, vol: torch.Tensor) -> torch.Tensor:
        self._check_sgm(sgm)
        if ray_mask is not None:
            self._check_sgm(ray_mask, "ray_mask")
        self._check_vol(vol)
        call("mct_backproject_ir", self.h, _f32(sgm), None if ray_mask is None else _f32(ray_mask),
             _f32(vol))
        return vol
