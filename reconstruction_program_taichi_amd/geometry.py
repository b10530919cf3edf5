"""Host-side geometry arrays and filter taps, in the reference's FP32 order.

The reference fills these with small Taichi kernels in FP32
(run_mgfbp.py:877-944, run_mgfpj.py:451-461); here NumPy float32 arithmetic
reproduces the same operations in the same order, so the arrays the GPU
kernels receive are the ones the reference's kernels would see.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi

PI = 3.1415926536  # the reference's constant (run_mgfbp.py:16), not math.pi
f32 = np.float32


def angle_array(view_num: int, scan_angle: float, offset: float) -> np.ndarray:
    """beta_i = scan / V * i + offset in FP32 (run_mgfbp.py:930-934 with
    offset = ImageRotation; run_mgfpj.py:451-455 with offset = StartAngle)."""
    i = np.arange(view_num, dtype=f32)
    return (f32(scan_angle) / f32(view_num)) * i + f32(offset)


def det_pos_array(count: int, count_actual: int, width: float, offset: float,
                  begin: int = 0) -> np.ndarray:
    """x_i = (i + begin - (count-1)/2) * width + offset in FP32
    (GenerateDectPixPosArray run_mgfbp.py:937-944; run_mgfpj.py:457-461 is the
    begin = 0, count_actual = count case)."""
    i = (np.arange(count_actual, dtype=np.int64) + int(begin)).astype(f32)
    return (i - f32((count - 1) / 2.0)) * f32(width) + f32(offset)


def hamming_taps(n: int, w: float, t: float, sdd: float, curved: bool = False,
                 dbt: bool = False) -> np.ndarray:
    """Ramp / Hamming-cosine blend taps h[i], i in [0, 2n-1), n_i = i - (n-1)
    (GenerateHammingKernel run_mgfbp.py:877-914), FP32."""
    nn = np.arange(2 * n - 1, dtype=np.int64) - (n - 1)
    w, t, sdd = f32(w), f32(t), f32(sdd)
    pi = f32(PI)
    h = np.zeros(2 * n - 1, dtype=f32)
    odd = (nn % 2) != 0
    h[nn == 0] = t / (f32(4) * w * w)
    if curved:
        tv = nn[odd].astype(f32) * w / sdd
        s = tv - tv ** 3 / f32(3) / f32(2) / f32(1) + tv ** 5 / f32(5) / f32(4) / f32(3) / f32(2) / f32(1)
        h[odd] = -t / (pi * pi * (sdd ** 2) * s ** 2)
    else:
        h[odd] = -t / (pi * pi * (nn[odd].astype(f32) ** 2) * (w ** 2))
    sgn = np.where(nn % 2 == 0, f32(1), f32(-1))
    a = (f32(1) + f32(2) * nn.astype(f32))
    b = (f32(1) - f32(2) * nn.astype(f32))
    cos_part = (sgn / (f32(2) * pi * w * w) * (f32(1) / a + f32(1) / b)
                - f32(1) / (pi * pi * w * w) * (f32(1) / a / a + f32(1) / b / b))
    h = (h + (f32(1) - t) * cos_part).astype(f32)
    if dbt:
        k_t = f32(0.02) / (f32(2) * w)
        add = np.empty_like(h)
        add[nn == 0] = k_t ** 2
        tv = nn[nn != 0].astype(f32) * w * k_t * pi
        add[nn != 0] = (np.sin(tv) ** 2) / (tv ** 2) * (k_t ** 2)
        h = (h + add).astype(f32)
    return h


def gaussian_taps(n: int, w: float, delta: float) -> np.ndarray:
    """Normalised Gaussian taps (GenerateGassianKernel run_mgfbp.py:917-927)."""
    nn = (np.arange(2 * n - 1, dtype=np.int64) - (n - 1)).astype(f32)
    g = np.exp(-nn * nn / f32(2) / f32(delta) / f32(delta)).astype(f32)
    return (g / np.float32(g.sum(dtype=np.float64)) / f32(w)).astype(f32)


def gaussian_apodized_ramp_taps(n: int, w: float, delta: float, sdd: float,
                                curved: bool = False, dbt: bool = False) -> np.ndarray:
    """Ramp (t = 1) convolved with the Gaussian taps, times w, truncated to the
    central 2n-1 (ConvolveKernelAndKernel run_mgfbp.py:1048-1060 with
    InitializeReconKernel run_mgfbp.py:1266-1276)."""
    ramp = hamming_taps(n, w, 1.0, sdd, curved, dbt).astype(np.float64)
    gauss = gaussian_taps(n, w, delta).astype(np.float64)
    full = np.convolve(gauss, ramp)  # length 4n-3; out[i] = sum_j g[j] r[i - (j - (n-1))]
    return (full[n - 1:3 * n - 2] * float(w)).astype(f32)


def shepp_logan_taps(n: int, w: float) -> np.ndarray:
    """Shepp-Logan taps h[k] = -2 / (pi^2 w^2 (4k^2 - 1)) (Shepp & Logan 1974;
    Kak & Slaney ch. 3).  The reference has no such filter; this is an
    additive option ('SheppLoganFilter')."""
    k = (np.arange(2 * n - 1, dtype=np.int64) - (n - 1)).astype(np.float64)
    return (-2.0 / (np.pi ** 2 * float(w) ** 2 * (4.0 * k * k - 1.0))).astype(f32)


@dataclass
class GeometrySpec:
    """Everything struct mct_geometry carries, with NumPy float32 arrays."""

    img_dim: int
    img_dim_z: int
    pix: float
    vox_h: float
    n_views: int
    sid: float
    sdd: float
    total_scan_angle: float
    angles: np.ndarray
    u: np.ndarray
    v: np.ndarray
    cx: float = 0.0
    cy: float = 0.0
    cz: float = 0.0
    img_rot: float = 0.0
    view_mode: int = 1
    short_scan: bool = False
    cone: bool = False
    curved: bool = False
    n_v_full: int = 0
    det_h: float = 0.0
    fpj_step: float = 0.2
    fpj_count_mode: int = 0
    helical: bool = False
    helical_pitch: float = 0.0
    pmatrix: np.ndarray | None = None
    center_z: np.ndarray | None = None
    # forward-projection P-matrix branch: A = inv(P[:, :3]) and the source
    # x_s per view (run_mgfpj_v2.py:218-227)
    matrix_a: np.ndarray | None = None
    x_s: np.ndarray | None = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_u(self) -> int:
        return int(self.u.shape[0])

    @property
    def n_v(self) -> int:
        return int(self.v.shape[0])

    def to_struct(self) -> _abi.Geometry:
        g = _abi.Geometry()
        g.img_dim, g.img_dim_z = int(self.img_dim), int(self.img_dim_z)
        g.pix, g.vox_h = float(self.pix), float(self.vox_h)
        g.cx, g.cy, g.cz = float(self.cx), float(self.cy), float(self.cz)
        g.img_rot, g.view_mode = float(self.img_rot), int(self.view_mode)
        g.n_views = int(self.n_views)
        g.sid, g.sdd = float(self.sid), float(self.sdd)
        g.total_scan_angle = float(self.total_scan_angle)
        g.short_scan, g.cone, g.curved = int(bool(self.short_scan)), int(bool(self.cone)), int(bool(self.curved))
        g.n_u, g.n_v = self.n_u, self.n_v
        g.n_v_full = int(self.n_v_full or self.n_v)
        g.det_h = float(self.det_h)
        g.fpj_step, g.fpj_count_mode = float(self.fpj_step), int(self.fpj_count_mode)
        g.helical, g.helical_pitch = int(bool(self.helical)), float(self.helical_pitch)
        g.apply_pmatrix = int(self.pmatrix is not None)
        self._keep = []

        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=f32)
            self._keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_float))

        if self.angles.shape[0] != self.n_views:
            raise ValueError(f"angles has {self.angles.shape[0]} entries for {self.n_views} views")
        g.angles, g.u, g.v = arr(self.angles), arr(self.u), arr(self.v)
        g.pmatrix = arr(self.pmatrix)
        g.center_z = arr(self.center_z)
        g.matrix_a = arr(self.matrix_a)
        g.x_s = arr(self.x_s)
        return g


def div_factor(total_scan_angle: float, short_scan: bool) -> float:
    """Redundancy factor of run_mgfbp.py:1072-1084 in FP32."""
    pi = f32(PI)
    a = f32(abs(f32(total_scan_angle)))
    rounds = np.floor(a / (pi * f32(2)))
    if short_scan:
        remain = a - rounds * f32(2) * pi
        if a < f32(2) * pi:
            return 1.0
        if remain < pi:
            return float(f32(1) / (rounds * f32(2)))
        if pi < remain < f32(2) * pi:
            return float(f32(1) / (rounds * f32(2) + f32(1)))
        return 1.0
    return float(f32(1) / (rounds * f32(2)))
