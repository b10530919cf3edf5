"""CPU oracle: NumPy restatement of the reference's reconstruction hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module; only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the
checker.  Every function cites the reference file:line it restates
(SEU-CT-Recon/Reconstruction_program_taichi, Taichi kernels executed here as
plain array code).

Two arithmetic modes:
  * dtype=np.float64 - the accuracy oracle: same discretisation, FP64
    throughout, cos/sin/sqrt in FP64;
  * dtype=np.float32 - "reference order": FP32 with the reference's
    expression order (NumPy never fuses multiply-add), so mask decisions
    (floor / range tests) land where the reference's FP32 code puts them.
    Per-view cos/sin are FP64 values rounded once to FP32.

How each function is pinned (tests/test_oracle.py): closed forms (taps,
weights, chord lengths, analytic ellipse integrals, FBP of a disc),
library routines (np.convolve, dense Toeplitz products, scipy.sparse
difference operators, np.linalg.solve), published known-answer vectors
(Philox4x32-10) and brute force on tiny grids.  Reference behaviours that
only an import of Taichi could pin (fast-math rounding of intrinsics,
ti.randn's stream) stay unpinned and are covered by tolerances.
"""
from __future__ import annotations

import numpy as np

PI = 3.1415926536  # run_mgfbp.py:16


def _t(dtype):
    return np.dtype(dtype).type


# ---------------------------------------------------------------- geometry
def angles(view_num, scan, offset, dtype=np.float32):
    """run_mgfbp.py:930-934 / run_mgfpj.py:451-455."""
    T = _t(dtype)
    i = np.arange(view_num).astype(dtype)
    return (T(scan) / T(view_num)) * i + T(offset)


def det_positions(count, count_actual, width, offset, begin=0, dtype=np.float32):
    """run_mgfbp.py:937-944."""
    T = _t(dtype)
    i = (np.arange(count_actual) + begin).astype(dtype)
    return (i - T((count - 1) / 2.0)) * T(width) + T(offset)


def view_table(ang, rot, dtype=np.float32):
    """cos/sin of (angle - rot) and dbeta (run_mgfbp.py:1105-1111, 1122)."""
    T = _t(dtype)
    a = np.asarray(ang, dtype=dtype)
    ar = (a - T(rot)).astype(dtype)
    c = np.cos(ar.astype(np.float64)).astype(dtype)
    s = np.sin(ar.astype(np.float64)).astype(dtype)
    nv = a.shape[0]
    d = np.empty(nv, dtype=dtype)
    d[:-1] = np.abs(a[1:] - a[:-1])
    d[-1] = np.abs(a[nv - 1] - a[0]) / T(nv - 1)
    return c, s, d


def div_factor(scan, short_scan, dtype=np.float32):
    """run_mgfbp.py:1072-1084."""
    T = _t(dtype)
    pi = T(PI)
    a = T(abs(T(scan)))
    rounds = np.floor(a / (pi * T(2)))
    if short_scan:
        remain = a - rounds * T(2) * pi
        if a < T(2) * pi:
            return T(1)
        if remain < pi:
            return T(1) / (rounds * T(2))
        if pi < remain < T(2) * pi:
            return T(1) / (rounds * T(2) + T(1))
        return T(1)
    return T(1) / (rounds * T(2))


# ------------------------------------------------------------- filter taps
def ramlak_taps(n, w, dtype=np.float64):
    """GenerateHammingKernel with t = 1, flat detector (run_mgfbp.py:887-898):
    h0 = 1/(4w^2); even n -> 0; odd n -> -1/(pi^2 n^2 w^2)."""
    T = _t(dtype)
    nn = (np.arange(2 * n - 1) - (n - 1)).astype(dtype)
    h = np.zeros(2 * n - 1, dtype=dtype)
    h[n - 1] = T(1) / (T(4) * T(w) * T(w))
    odd = (np.arange(2 * n - 1) - (n - 1)) % 2 != 0
    h[odd] = T(-1) / (T(PI) * T(PI) * nn[odd] ** 2 * T(w) ** 2)
    return h


# ----------------------------------------------------------------- weighting
def weight_sinogram(sgm, u, v, ang, sdd, scan, short_scan, curved=False, bh=None,
                    dtype=np.float32):
    """BHCorrection (run_mgfbp.py:946-954) then WeightSgm (:956-1011).
    sgm: (rows, views, n_u); returns a new array."""
    T = _t(dtype)
    p = np.asarray(sgm, dtype=dtype).copy()
    if bh is not None and len(bh) > 0:
        acc = np.zeros_like(p)
        pw = p.copy()
        for c in np.asarray(bh, dtype=dtype):
            acc = acc + c * pw
            pw = pw * p
        p = acc
    U = np.asarray(u, dtype=dtype)[None, None, :]
    V = np.asarray(v, dtype=dtype)[:, None, None]
    sdd = T(sdd)
    if curved:
        p = p * sdd * np.cos((T(-1) * U / sdd).astype(np.float64)).astype(dtype) * sdd / np.sqrt(sdd * sdd + V * V)
    else:
        p = (p * sdd * sdd) / np.sqrt(sdd * sdd + U * U + V * V)
    if short_scan:
        pi = T(PI)
        scan = T(scan)
        asc = T(abs(scan))
        rounds = np.floor(asc / (pi * T(2)))
        remain = asc - rounds * pi * T(2)
        A = np.asarray(ang, dtype=dtype)
        beta = np.abs(A - A[0])[None, :, None]
        direc = asc / scan
        if curved:
            gamma = ((T(-1) * U) / sdd) * direc
        else:
            gamma = np.arctan2((T(-1) * U).astype(np.float64), float(sdd)).astype(dtype) * direc
        gmax = remain - pi
        beta, gamma = np.broadcast_arrays(beta, gamma)
        w = np.zeros(beta.shape, dtype=dtype)
        if remain <= pi and rounds == 0:
            w[:] = 1
        elif remain <= pi and rounds >= 1:
            m1 = (0 <= beta) & (beta < remain)
            m2 = (remain <= beta) & (beta < T(2) * pi * rounds)
            m3 = (T(2) * pi * rounds <= beta) & (beta <= T(2) * pi * rounds + remain)
            w = np.where(m1, beta / remain, w)
            w = np.where(m2 & ~m1, T(1), w)
            w = np.where(m3 & ~m1 & ~m2, (T(2) * pi * rounds + remain - beta) / remain, w)
        elif pi < remain <= T(2) * pi:
            b1 = gmax - T(2) * gamma
            b2 = pi * (T(2) * rounds + T(1)) - T(2) * gamma
            b3 = pi * (T(2) * rounds + T(1)) + gmax
            m1 = (0 <= beta) & (beta < b1)
            m2 = (b1 <= beta) & (beta < b2)
            m3 = (b2 <= beta) & (beta <= b3)
            with np.errstate(divide="ignore", invalid="ignore"):
                s1 = np.sin((pi / T(2) * beta / b1).astype(np.float64)).astype(dtype)
                s3 = np.sin((pi / T(2) * (pi + gmax - (beta - pi * T(2) * rounds)) / (gmax + T(2) * gamma)).astype(np.float64)).astype(dtype)
            w = np.where(m1, s1 * s1, w)
            w = np.where(m2 & ~m1, T(1), w)
            w = np.where(m3 & ~m1 & ~m2, s3 * s3, w)
        else:
            w[:] = 1
        p = p * w
    return p.astype(dtype)


# -------------------------------------------------------------- ramp filter
def filter_rows(x, taps, scale, dtype=np.float64):
    """ConvolveSgmAndKernel horizontal pass (run_mgfbp.py:1033-1046):
    out[r,k] = scale * sum_m x[r,m] taps[k+n-1-m], summed in ascending m."""
    x = np.asarray(x, dtype=dtype)
    n = x.shape[-1]
    rows = x.reshape(-1, n)
    tp = np.asarray(taps, dtype=dtype)
    if dtype == np.float64:
        out = np.stack([np.convolve(r, tp)[n - 1:2 * n - 1] for r in rows]) if rows.shape[0] else rows.copy()
    else:
        acc = np.zeros_like(rows)
        k = np.arange(n)
        for m in range(n):
            acc = acc + rows[:, m:m + 1] * tp[k + n - 1 - m][None, :]
        out = acc
    return (out * _t(dtype)(scale)).reshape(x.shape).astype(dtype)


def filter_vertical(x, kern, det_h, dtype=np.float64):
    """Vertical Gaussian of run_mgfbp.py:1019-1031 (|i-n| <= 10)."""
    x = np.asarray(x, dtype=dtype)
    rows = x.shape[0]
    out = np.zeros_like(x)
    kern = np.asarray(kern, dtype=dtype)
    for i in range(rows):
        acc = np.zeros_like(x[0])
        for n in range(rows):
            if -10 <= i - n <= 10:
                acc = acc + x[n] * kern[i + rows - 1 - n]
        out[i] = acc * _t(dtype)(det_h)
    return out


# ---------------------------------------------------------- back-projection
def _pm_project(pm, j, x, y, z, T):
    """P-matrix branch of run_mgfbp.py:1133-1141 and :1165-1168 (also
    run_mgfbp_iterative_recon_IRN_CGLS.py:430-458): the detector indexes of
    voxel (x, y, z) in view j, summed in the reference's order."""
    P = [T(p) for p in pm[12 * j:12 * j + 12]]
    mag = T(1) / (P[8] * x + P[9] * y + P[10] * z + P[11] * T(1))
    uidx = (P[0] * x + P[1] * y + P[2] * z + P[3] * T(1)) * mag
    vidx = (P[4] * x + P[5] * y + P[6] * z + P[7] * T(1)) * mag
    return mag, uidx, vidx


def _fbp_view_terms(filt, g, dtype):
    """The view loop of Mgfbp.BackProjectionPixelDriven (run_mgfbp.py:1107-1187),
    axial view, P-matrix branch when g["pmatrix"] (12 per view) is given: for
    each view j in order, (ok, term) over the (nz, ny, nx) grid, ok = the voxel
    passes view j's u (and, cone beam, v) range tests, term = its weighted
    contribution sid/L^2 * value * dbeta_j * div_factor (meaningless where not
    ok).  filt: (rows, views, n_u)."""
    T = _t(dtype)
    f = np.asarray(filt, dtype=dtype)
    nx, nz = g["img_dim"], g["img_dim_z"]
    nu = f.shape[2]
    nrows = f.shape[0]
    pix, vh, sid, sdd = T(g["pix"]), T(g.get("vox_h", 0.0)), T(g["sid"]), T(g["sdd"])
    cx, cy, cz = T(g.get("cx", 0.0)), T(g.get("cy", 0.0)), T(g.get("cz", 0.0))
    rot = g.get("img_rot", 0.0)
    cone = bool(g.get("cone", False))
    curved = bool(g.get("curved", False))
    ang = np.asarray(g["angles"], dtype=dtype)
    u = np.asarray(g["u"], dtype=dtype)
    v = np.asarray(g["v"], dtype=dtype)
    c, s, db = view_table(ang, rot, dtype)
    div = div_factor(g["total_scan_angle"], g.get("short_scan", False), dtype)
    cr = T(np.cos(np.float64(T(rot))))
    sr = T(np.sin(np.float64(T(rot))))
    half, halfz = T((nx - 1) / 2.0), T((nz - 1) / 2.0)
    ix = np.arange(nx).astype(dtype)
    iz = np.arange(nz).astype(dtype)
    xa = (pix * (ix - half) + cx)[None, None, :]
    ya = ((-pix) * (ix - half) + cy)[None, :, None]
    x = xa * cr + ya * sr
    y = (-xa) * sr + ya * cr
    z = ((iz - halfz) * vh + cz)[:, None, None]
    x, y, z = np.broadcast_arrays(x, y, z)
    u0, du = u[0], u[1] - u[0]
    if cone:
        v0, dv = v[0], v[1] - v[0]
        adv = np.abs(dv)
        dh = T(g["det_h"])
    slice_idx = np.broadcast_to(np.arange(nz)[:, None, None], x.shape)
    pm = g.get("pmatrix")
    if pm is not None:
        pm = np.asarray(pm, dtype=dtype)
    for j in range(ang.shape[0]):
        L = sid - x * c[j] - y * s[j]
        if pm is not None:
            mag, uidx, vidx_pm = _pm_project(pm, j, x, y, z, T)
        else:
            mag = sdd / L
            yr = (-x) * s[j] + y * c[j]
            if curved:
                uu = (sdd * np.arctan2(yr.astype(np.float64), L.astype(np.float64))).astype(dtype)
            else:
                uu = mag * yr
            uidx = (uu - u0) / du
        ok = ~((uidx < 0) | (uidx + T(1) > T(nu - 1)))
        uf = np.clip(np.floor(uidx), 0, nu - 2).astype(np.int64)
        ru = uidx - uf.astype(dtype)
        if curved:
            q = x * s[j] - y * c[j]
            w = T(1) / ((L * L) + q * q)
        else:
            w = T(1) / (L * L)
        if cone:
            vidx = vidx_pm if pm is not None else (mag * z - v0) / dh * adv / dv
            vff = np.floor(vidx)
            ok &= ~((vff < 0) | (vff + T(1) > T(nrows - 1)))
            vf = np.clip(vff, 0, nrows - 2).astype(np.int64)
            rv = vidx - vf.astype(dtype)
            p0 = f[vf, j, uf] * (T(1) - ru) + f[vf, j, uf + 1] * ru
            p1 = f[vf + 1, j, uf] * (T(1) - ru) + f[vf + 1, j, uf + 1] * ru
            val = (T(1) - rv) * p0 + rv * p1
        else:
            val = (T(1) - ru) * f[slice_idx, j, uf] + ru * f[slice_idx, j, uf + 1]
        yield ok, (sid * w) * val * db[j] * div


def backproject_fbp(filt, g, dtype=np.float32, return_mask=False):
    """Mgfbp.BackProjectionPixelDriven (run_mgfbp.py:1062-1187), axial view,
    with the P-matrix branch when g["pmatrix"] (12 per view) is given: the
    terms summed in view order; a voxel that fails a range test in any view
    is 0.  filt: (rows, views, n_u).  g: dict of geometry scalars and arrays
    (angles, u, v).  Returns (nz, ny, nx)."""
    T = _t(dtype)
    acc = alive = None
    for ok, term in _fbp_view_terms(filt, g, dtype):
        if acc is None:
            acc = np.zeros(term.shape, dtype=dtype)
            alive = np.ones(term.shape, dtype=bool)
        alive &= ok
        acc = acc + term
    out = np.where(alive, acc, T(0)).astype(dtype)
    return (out, alive) if return_mask else out


def backproject_fbp_v2(filt, g, vol=None, dtype=np.float32):
    """The view-by-view FBP of Mgfbp_v2 (run_mgfbp_v2.py:71-94): each view's
    BackProjectionPixelDriven (:238-364) adds its term to the volume, a view
    whose range test fails sets the voxel to -10000 (:312-313, 332-333) and
    later views keep adding; SetTruncatedRegionToZero (:232-236) then sets
    every voxel below -5000 to 0.  vol: the volume the views are added to
    (zeros when None).  Returns (nz, ny, nx)."""
    T = _t(dtype)
    acc = None if vol is None else np.array(vol, dtype=dtype)
    for ok, term in _fbp_view_terms(filt, g, dtype):
        if acc is None:
            acc = np.zeros(term.shape, dtype=dtype)
        acc = np.where(ok, acc + term, T(-10000)).astype(dtype)
    return np.where(acc < T(-5000), T(0), acc).astype(dtype)


def backproject_ir(sgm, g, ray_mask=None, dtype=np.float32):
    """Mgfbp_ir.BackProjectionPixelDrivenPerView over all views plus
    SetTruncatedRegionToZero (run_mgfbp_iterative_recon_IRN_CGLS.py:372-481),
    axial view, P-matrix branch when g["pmatrix"] is given.
    sgm: (rows, views, n_u)."""
    T = _t(dtype)
    f = np.asarray(sgm, dtype=dtype)
    if ray_mask is not None:
        f = f * np.asarray(ray_mask, dtype=dtype)
    nx, nz = g["img_dim"], g["img_dim_z"]
    nu, nrows = f.shape[2], f.shape[0]
    pix, vh, sid, sdd = T(g["pix"]), T(g.get("vox_h", 0.0)), T(g["sid"]), T(g["sdd"])
    cx, cy, cz = T(g.get("cx", 0.0)), T(g.get("cy", 0.0)), T(g.get("cz", 0.0))
    rot = g.get("img_rot", 0.0)
    cone = bool(g.get("cone", False))
    ang = np.asarray(g["angles"], dtype=dtype)
    u = np.asarray(g["u"], dtype=dtype)
    v = np.asarray(g["v"], dtype=dtype)
    c, s, _ = view_table(ang, rot, dtype)
    cr = T(np.cos(np.float64(T(rot))))
    sr = T(np.sin(np.float64(T(rot))))
    half, halfz = T((nx - 1) / 2.0), T((nz - 1) / 2.0)
    ix = np.arange(nx).astype(dtype)
    iz = np.arange(nz).astype(dtype)
    xa = (pix * (ix - half) + cx)[None, None, :]
    ya = ((-pix) * (ix - half) + cy)[None, :, None]
    x = xa * cr + ya * sr
    y = (-xa) * sr + ya * cr
    zb = ((iz - halfz) * vh)[:, None, None]
    x, y, zb = np.broadcast_arrays(x, y, zb)
    acc = np.zeros(x.shape, dtype=dtype)
    flag = np.ones(x.shape, dtype=bool)
    u0, du = u[0], u[1] - u[0]
    if cone:
        v0, dv = v[0], v[1] - v[0]
        adv = np.abs(dv)
        dh = T(g["det_h"])
    czs = g.get("center_z")
    slice_idx = np.broadcast_to(np.arange(nz)[:, None, None], x.shape)
    pm = g.get("pmatrix")
    if pm is not None:
        pm = np.asarray(pm, dtype=dtype)
    for j in range(ang.shape[0]):
        z = zb + (T(czs[j]) if czs is not None else cz)
        L = sid - x * c[j] - y * s[j]
        dxs = x - sid * c[j]
        dys = y - sid * s[j]
        den = np.where(np.abs(dxs) > np.abs(dys), np.abs(dxs), np.abs(dys))
        df = np.sqrt(dxs * dxs + dys * dys + z * z) / den
        if pm is not None:
            mag, uidx, vidx_pm = _pm_project(pm, j, x, y, z, T)
        else:
            mag = sdd / L
            uu = mag * ((-x) * s[j] + y * c[j])
            uidx = (uu - u0) / du
        bad = (uidx < 0) | (uidx + T(1) > T(nu - 1))
        flag &= ~bad
        uf = np.clip(np.floor(uidx), 0, nu - 2).astype(np.int64)
        ru = uidx - uf.astype(dtype)
        if cone:
            vidx = vidx_pm if pm is not None else (mag * z - v0) / dh * adv / dv
            vff = np.floor(vidx)
            okv = ~((vff < 0) | (vff + T(1) > T(nrows - 1)))
            vf = np.clip(vff, 0, nrows - 2).astype(np.int64)
            rv = vidx - vf.astype(dtype)
            p0 = f[vf, j, uf] * (T(1) - ru) + f[vf, j, uf + 1] * ru
            p1 = f[vf + 1, j, uf] * (T(1) - ru) + f[vf + 1, j, uf + 1] * ru
            val = np.where(okv, (T(1) - rv) * p0 + rv * p1, T(0))
        else:
            val = (T(1) - ru) * f[slice_idx, j, uf] + ru * f[slice_idx, j, uf + 1]
        acc = acc + np.where(bad, T(0), val * df * pix)
    return np.where(flag, acc, T(0)).astype(dtype)


# ------------------------------------------------------- helical rebinning
def helical_rebin(sgm, z_pos, view_count, off_v, det_h, sdd, mag, dz, dtype=np.float32):
    """Mgfbp_helical.SinogramRebinning (run_mgfbp_helical.py:165-198) for
    every slice position: sgm (rows, views, n_u) -> (len(z_pos), view_count,
    n_u).  Slice z takes views k = floor(k0 - view_count/2 + 1/2) + i with
    k0 = (z - off_v/mag)/dz; in view k it interpolates linearly between the
    detector rows around v_pos = (z - k dz) mag - off_v (row index
    -v_pos/det_h + (rows-1)/2, clamped to the first / last row) and scales by
    sdd / sqrt(sdd^2 + v_pos^2); views outside [0, views-2] give 0.  The
    view window is an index decision of the reference's FP32 kernel and is
    taken in FP32 in both modes (a slice position on a tie would otherwise
    pick another window than the kernel and the slice's image rotation)."""
    T = _t(dtype)
    F = np.float32
    p = np.asarray(sgm, dtype=dtype)
    rows, views, nu = p.shape
    off_v, det_h, sdd, mag, dz = T(off_v), T(det_h), T(sdd), T(mag), T(dz)
    out = np.zeros((len(z_pos), view_count, nu), dtype=dtype)
    for zi, zf in enumerate(np.asarray(z_pos, dtype=dtype)):
        k0 = (F(zf) - F(off_v) / F(mag)) / F(dz)
        begin = T(np.floor(k0 - F(0.5) * F(view_count) + F(0.5)))
        for i in range(view_count):
            view = begin + T(i)
            if not (0 <= view <= views - 2):
                continue
            k = int(view)
            v_pos = (zf - view * dz) * mag - off_v
            v_idx = (-v_pos) / det_h + (T(rows) - T(1)) / T(2)
            if v_idx >= rows - 1:
                val = p[rows - 1, k]
            elif v_idx <= 0:
                val = p[0, k]
            else:
                fl = np.floor(v_idx)
                w = v_idx - fl
                r = int(fl)
                val = p[r, k] * (T(1) - w) + p[r + 1, k] * w
            out[zi, i] = val * (sdd / np.sqrt(sdd * sdd + v_pos * v_pos))
    return out


# ------------------------------------------------------- forward projection
def fp_scalars(g, count_mode=0, dtype=np.float32):
    """Scalars of run_mgfpj.py:486-527 (count_mode 1: run_mgfpj_v3.py:252)."""
    T = _t(dtype)
    nx, nz = g["img_dim"], g["img_dim_z"]
    pix, vh = T(g["pix"]), T(g.get("vox_h", 0.0))
    dim = T(nx) * pix
    dimz = T(nz) * vh
    D = np.sqrt(T(2) * (dim * dim) + dimz * dimz)
    sid = T(g["sid"])
    lmin = sid - D / T(2)
    lmax = sid + D / T(2)
    diag = np.sqrt(T(2) * (pix * pix) + vh * vh)
    span = D if count_mode == 1 else (lmax - lmin)
    count = int(np.floor(span / (T(g.get("fpj_step", 0.2)) * diag)))
    x0 = -(T(nx) - T(1)) / T(2) * pix + T(g.get("cx", 0.0))
    y0 = -(T(nx) - T(1)) / T(2) * (-pix) + T(g.get("cy", 0.0))
    zbase = -(T(nz) - T(1)) / T(2) * vh
    return dict(lmin=lmin, diag=diag, count=count, x0=x0, y0=y0, zbase=zbase)


def helical_z_per_view(ang, v, n_v_full, pitch, sdd, sid, dtype=np.float32):
    """Gantry travel per view of a helical scan (run_mgfpj.py:518-523): the
    scan angle estimated from the first and last view angles and the view
    count, over 2 pi, gives the laps; pitch x (detector height at the
    isocentre: |v[1] - v[0]| x n_v_full x sid / sdd) per lap."""
    T = _t(dtype)
    ang = np.asarray(ang, dtype=dtype)
    v = np.asarray(v, dtype=dtype)
    nv = ang.shape[0]
    tsa = np.abs(ang[nv - 1] - ang[0]) / T(nv - 1) * T(nv)
    laps = tsa / (T(PI) * T(2))
    return T(T(pitch) * (laps / T(nv)) * (np.abs(v[1] - v[0]) * T(n_v_full)) / (T(sdd) / T(sid)))


def _pm_rays(g, ang, Ui, R, J, dtype):
    """P-matrix branch of the forward projector (run_mgfpj_v2.py:334-355,
    run_mgfpj_v3.py:286-307, run_mgfbp_iterative_recon_IRN_CGLS.py:598-622):
    sample l of ray (u index, row, view j) sits at
    s * (A_j [u, v, 1]) + x_s_j with s = l / |detector element - source|,
    z shifted by the helical offset, then rotated by angles[0].  Returns
    None without g["matrix_a"] (9 per view) and g["x_s"] (3 per view), else
    a function of (l, dist, zs) giving (x_rot, y_rot, z_rot)."""
    if g.get("matrix_a") is None:
        return None
    T = _t(dtype)
    nv = ang.shape[0]
    A = np.asarray(g["matrix_a"], dtype=dtype).reshape(nv, 9)[J]
    xs = np.asarray(g["x_s"], dtype=dtype).reshape(nv, 3)[J]
    fu, fv = Ui.astype(dtype), R.astype(dtype)
    w = [A[..., 3 * r] * fu + A[..., 3 * r + 1] * fv + A[..., 3 * r + 2] * T(1) for r in range(3)]
    c0 = T(np.cos(np.float64(ang[0])))
    s0 = T(np.sin(np.float64(ang[0])))

    def at(l, dist, zs):
        om = l / dist
        xp = om * w[0] + xs[..., 0]
        yp = om * w[1] + xs[..., 1]
        zp = om * w[2] + xs[..., 2] + zs
        return xp * c0 - yp * s0, yp * c0 + xp * s0, zp

    return at


def forward_project(vol, g, oversample=1, ray_mask=None, count_mode=0, dtype=np.float32,
                    samples_in=False):
    """Mgfpj.ForwardProjectionBilinear + BinSinogram (run_mgfpj.py:463-616),
    all rows and views; u already oversampled (len = n_u * oversample); the
    P-matrix rays of _pm_rays when g carries matrix_a and x_s.
    vol: (nz, ny, nx).  Returns (rows, views, n_u)."""
    T = _t(dtype)
    vol = np.asarray(vol, dtype=dtype)
    nz, _, nx = vol.shape
    sc = fp_scalars(g, count_mode, dtype)
    sid, sdd = T(g["sid"]), T(g["sdd"])
    pix, vh = T(g["pix"]), T(g.get("vox_h", 0.0))
    cone, curved = bool(g.get("cone", False)), bool(g.get("curved", False))
    ang = np.asarray(g["angles"], dtype=dtype)
    u = np.asarray(g["u"], dtype=dtype)
    v = np.asarray(g["v"], dtype=dtype)
    nrows, nviews, nuo = v.shape[0], ang.shape[0], u.shape[0]
    zpv = T(g.get("zpv", 0.0))
    czs = g.get("center_z")
    cz = T(g.get("cz", 0.0))
    # rays: (rows, views, nuo)
    R, J, Ui = np.meshgrid(np.arange(nrows), np.arange(nviews), np.arange(nuo), indexing="ij")
    if curved:
        gp = (-u[Ui]) / sdd
        dx = (-sdd * np.cos(gp.astype(np.float64)).astype(dtype) + sid)
        dy = -sdd * np.sin(gp.astype(np.float64)).astype(dtype)
    else:
        dx = np.full(Ui.shape, -(sdd - sid), dtype=dtype)
        dy = u[Ui]
    zs = zpv * J.astype(dtype)
    dz = v[R] + zs
    ex, ey, ez = dx - sid, dy - T(0), dz - zs
    dist = np.sqrt(ex * ex + ey * ey + ez * ez)
    ux, uy, uz = ex / dist, ey / dist, ez / dist
    ca = np.cos(ang.astype(np.float64)).astype(dtype)[J]
    sa = np.sin(ang.astype(np.float64)).astype(dtype)[J]
    z0 = sc["zbase"] + (np.asarray(czs, dtype=dtype)[J] if czs is not None else cz)
    acc = np.zeros(Ui.shape, dtype=dtype)
    nin = 0
    step, diag, lmin = T(g.get("fpj_step", 0.2)), sc["diag"], sc["lmin"]
    x0, y0 = sc["x0"], sc["y0"]
    flat = vol.reshape(-1)
    pm_ray = _pm_rays(g, ang, Ui, R, J, dtype)
    for k in range(sc["count"]):
        l = T(k) * step * diag + lmin
        if pm_ray is not None:
            xr, yr, z = pm_ray(l, dist, zs)
        else:
            x = sid + ux * l
            y = T(0) + uy * l
            z = zs + uz * l
            xr = x * ca - y * sa
            yr = y * ca + x * sa
        xi = np.floor((xr - x0) / pix).astype(np.int64)
        yi = np.floor((yr - y0) / (-pix)).astype(np.int64)
        ok = (xi >= 0) & (xi + 1 < nx) & (yi >= 0) & (yi + 1 < nx)
        if not ok.any():
            continue
        xic, yic = np.clip(xi, 0, nx - 2), np.clip(yi, 0, nx - 2)
        xw = (xr - (xic.astype(dtype) * pix + x0)) / pix
        yw = (yr - (yic.astype(dtype) * (-pix) + y0)) / (-pix)
        if cone:
            zi = np.floor((z - z0) / vh).astype(np.int64)
            ok &= (zi >= 0) & (zi + 1 < nz)
            zic = np.clip(zi, 0, nz - 2)
            zw = (z - (zic.astype(dtype) * vh + z0)) / vh
            base = (zic * nx + yic) * nx + xic
            pl = nx * nx
            lo = ((T(1) - xw) * (T(1) - yw) * flat[base] + xw * (T(1) - yw) * flat[base + 1]
                  + (T(1) - xw) * yw * flat[base + nx] + xw * yw * flat[base + nx + 1])
            hi = ((T(1) - xw) * (T(1) - yw) * flat[base + pl] + xw * (T(1) - yw) * flat[base + pl + 1]
                  + (T(1) - xw) * yw * flat[base + pl + nx] + xw * yw * flat[base + pl + nx + 1])
            val = ((T(1) - zw) * lo + zw * hi) * step * diag
        else:
            base = (R * nx + yic) * nx + xic
            val = ((T(1) - xw) * (T(1) - yw) * flat[base] + xw * (T(1) - yw) * flat[base + 1]
                   + (T(1) - xw) * yw * flat[base + nx] + xw * yw * flat[base + nx + 1]) * step * diag
        acc = acc + np.where(ok, val, T(0))
        nin += int(ok.sum())
    out = acc
    if oversample > 1:
        b = out.reshape(nrows, nviews, nuo // oversample, oversample)
        s = np.zeros(b.shape[:3], dtype=dtype)
        for i in range(oversample):
            s = s + b[..., i]
        out = s / T(oversample)
    if ray_mask is not None:
        out = np.where(np.asarray(ray_mask) == 1, out, T(0))
    out = out.astype(dtype)
    return (out, nin) if samples_in else out


def ray_mask(g, count_mode=0, dtype=np.float32):
    """GenEffectiveMapForwardProjectionAgent
    (run_mgfbp_iterative_recon_IRN_CGLS.py:679-833): 0 when a sample inside
    the image and its inscribed disc leaves the z range (cone beam); P-matrix
    rays as in forward_project."""
    T = _t(dtype)
    nx, nz = g["img_dim"], g["img_dim_z"]
    ang = np.asarray(g["angles"], dtype=dtype)
    u = np.asarray(g["u"], dtype=dtype)
    v = np.asarray(g["v"], dtype=dtype)
    nrows, nviews, nuo = v.shape[0], ang.shape[0], u.shape[0]
    if not g.get("cone", False):
        return np.ones((nrows, nviews, nuo), dtype=dtype)
    sc = fp_scalars(g, count_mode, dtype)
    sid, sdd = T(g["sid"]), T(g["sdd"])
    pix, vh = T(g["pix"]), T(g["vox_h"])
    R, J, Ui = np.meshgrid(np.arange(nrows), np.arange(nviews), np.arange(nuo), indexing="ij")
    dx = np.full(Ui.shape, -(sdd - sid), dtype=dtype)
    dy = u[Ui]
    zs = T(g.get("zpv", 0.0)) * J.astype(dtype)
    dz = v[R] + zs
    ex, ey, ez = dx - sid, dy, dz - zs
    dist = np.sqrt(ex * ex + ey * ey + ez * ez)
    ux, uy, uz = ex / dist, ey / dist, ez / dist
    ca = np.cos(ang.astype(np.float64)).astype(dtype)[J]
    sa = np.sin(ang.astype(np.float64)).astype(dtype)[J]
    czs = g.get("center_z")
    z0 = sc["zbase"] + (np.asarray(czs, dtype=dtype)[J] if czs is not None else T(g.get("cz", 0.0)))
    m = np.ones(Ui.shape, dtype=bool)
    half = T(nx) / T(2)
    step, diag, lmin = T(g.get("fpj_step", 0.2)), sc["diag"], sc["lmin"]
    pm_ray = _pm_rays(g, ang, Ui, R, J, dtype)
    for k in range(sc["count"]):
        l = T(k) * step * diag + lmin
        if pm_ray is not None:
            xr, yr, z = pm_ray(l, dist, zs)
        else:
            x = sid + ux * l
            y = uy * l
            z = zs + uz * l
            xr = x * ca - y * sa
            yr = y * ca + x * sa
        xi = np.floor((xr - sc["x0"]) / pix)
        yi = np.floor((yr - sc["y0"]) / (-pix))
        ok = (xi >= 0) & (xi + 1 < nx) & (yi >= 0) & (yi + 1 < nx)
        disc = (xi - half) ** 2 + (yi - half) ** 2 < half ** 2
        zi = np.floor((z - z0) / vh)
        m &= ~(ok & disc & ((zi < 0) | (zi + 1 >= nz)))
    return m.astype(dtype)


# -------------------------------------------------------------- noise RNG
M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85


def philox4x32_10(ctr, key):
    """Philox4x32 with 10 rounds (Salmon et al., SC'11).  ctr: (..., 4)
    uint32, key: (..., 2) uint32 arrays; returns (..., 4) uint32."""
    c = [np.asarray(ctr[..., i], dtype=np.uint64) for i in range(4)]
    k = [np.asarray(key[..., i], dtype=np.uint64) for i in range(2)]
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        c = [hi1 ^ c[1] ^ k[0], lo1, hi0 ^ c[3] ^ k[1], lo0]
        k = [(k[0] + np.uint64(W0)) & mask, (k[1] + np.uint64(W1)) & mask]
    return np.stack([x.astype(np.uint32) for x in c], axis=-1)


def normals(n, seed):
    """Standard normals for element indexes 0..n-1: Box-Muller on the first
    two words of Philox(counter = index, key = seed), FP64."""
    i = np.arange(n, dtype=np.uint64)
    ctr = np.zeros((n, 4), dtype=np.uint32)
    ctr[:, 0] = (i & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    ctr[:, 1] = (i >> np.uint64(32)).astype(np.uint32)
    key = np.zeros((n, 2), dtype=np.uint32)
    key[:, 0] = seed & 0xFFFFFFFF
    key[:, 1] = (seed >> 32) & 0xFFFFFFFF
    r = philox4x32_10(ctr, key)
    u1 = (r[:, 0].astype(np.float64) + 1.0) * (1.0 / 4294967296.0)
    u2 = r[:, 1].astype(np.float64) * (1.0 / 4294967296.0)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * 3.141592653589793 * u2)


def add_noise(sgm, photons, seed):
    """AddPossionNoise (run_mgfpj.py:643-654) with the counter-based normals."""
    p = np.asarray(sgm, dtype=np.float32)
    g = normals(p.size, seed).reshape(p.shape)
    t = np.float32(photons) * np.exp(-p)
    t = t + g.astype(np.float32) * np.sqrt(t)
    t = np.where(t <= 0, np.float32(1e-6), t)
    return np.log(np.float32(photons) / t).astype(np.float32)


# ---------------------------------------------------------- CG / TV pieces
def _diffs(x):
    """Forward differences of run_mgfbp_iterative_recon_IRN_CGLS.py:235-241."""
    x = np.asarray(x, dtype=np.float64)
    ux = np.zeros_like(x)
    uy = np.zeros_like(x)
    uz = np.zeros_like(x)
    ux[:, :-1, :] = x[:, 1:, :] - x[:, :-1, :]
    uy[:, :, :-1] = x[:, :, 1:] - x[:, :, :-1]
    if x.shape[0] > 2:
        uz[:-1, :, :] = x[1:, :, :] - x[:-1, :, :]
    return ux, uy, uz


def tv_weight(x, beta):
    """GenerateWR (:234-243)."""
    ux, uy, uz = _diffs(x)
    return 1.0 / np.sqrt(ux ** 2 + uy ** 2 + uz ** 2 + beta)


def tv_value(x, beta):
    """TVPenaltyVal (:245-254)."""
    ux, uy, uz = _diffs(x)
    return float(np.sum(np.sqrt(ux * ux + uy * uy + uz * uz + beta)))


def dtwd(x, wr):
    """Dt_W_D (:267-286)."""
    ux, uy, uz = _diffs(x)
    ux, uy, uz = wr * ux, wr * uy, wr * uz
    uxx = np.zeros_like(ux)
    uyy = np.zeros_like(uy)
    uzz = np.zeros_like(uz)
    uxx[:, 1:-1, :] = ux[:, 1:-1, :] - ux[:, 0:-2, :]
    uyy[:, :, 1:-1] = uy[:, :, 1:-1] - uy[:, :, 0:-2]
    if x.shape[0] > 2:
        uzz[1:-1, :, :] = uz[1:-1, :, :] - uz[0:-2, :, :]
    return -uxx - uyy - uzz


def cg_solve(apply_A, b, x0, iters):
    """The CG recursion of TikhonovSol (:188-217): r = d = b - A x0; then
    alpha = <r,r>/<d,Ad>, x += alpha d, r -= alpha Ad, beta = <r',r'>/<r,r>,
    d = r + beta d; no convergence test."""
    x = np.array(x0, dtype=np.float64)
    r = b - apply_A(x)
    d = r.copy()
    for _ in range(iters):
        q = apply_A(d)
        rr = float(np.sum(r * r))
        alpha = rr / float(np.sum(d * q))
        x = x + alpha * d
        r = r - alpha * q
        beta = float(np.sum(r * r)) / rr
        d = r + beta * d
    return x


def ir_solve(sgm, g, x0, lam, beta_tv, n_irn, n_iter, couple_z=True, prior=None, alpha=0.0,
             dtype=np.float64, projectors=None):
    """Mgfbp_ir.MainFunction / TikhonovSol (run_mgfbp_iterative_recon_IRN_CGLS.py:
    157-232) and, with a prior, Mgfbp_ir_piccs (…_piccs.py:83-150), with the
    oracle's projectors: ray map, P^T b, then 1 + n_irn reweighted CG passes of
    n_iter iterations.  couple_z=False drops the z difference (independent
    slices).  projectors: (forward_project, backproject_ir) with this module's
    signatures (e.g. the C restatements of oracle/cbp.py, for large cases);
    None uses this module's.  Returns x."""
    fwd, bwd = projectors if projectors is not None else (forward_project, backproject_ir)
    sgm = np.asarray(sgm, dtype=dtype)
    rmask = ray_mask(g, dtype=dtype)
    nz = x0.shape[0]
    zc = couple_z and nz > 2

    def tvw(x):
        if zc or nz <= 2:
            return tv_weight(x, beta_tv)
        return np.stack([tv_weight(x[k:k + 1], beta_tv)[0] for k in range(nz)])

    def dtwd_(x, w):
        if zc or nz <= 2:
            return dtwd(x, w)
        return np.stack([dtwd(x[k:k + 1], w[k:k + 1])[0] for k in range(nz)])

    def fpbp(x):
        fp = fwd(x, g, ray_mask=rmask, dtype=dtype)
        return bwd(fp, g, dtype=dtype).astype(np.float64)

    ratio = sgm.size / x0.size
    bpb = bwd(sgm, g, ray_mask=rmask, dtype=dtype).astype(np.float64)
    x = np.asarray(x0 if prior is None else prior, dtype=np.float64)
    c1 = lam * (1 - alpha) * ratio
    c2 = lam * alpha * ratio
    w = np.ones_like(x)
    wp = np.ones_like(x)
    for p in range(1 + n_irn):
        if p > 0:
            w = tvw(x)
            if prior is not None:
                wp = tvw(x - prior)

        def A(v):
            out = fpbp(v) + c1 * dtwd_(v, w)
            if prior is not None:
                out = out + c2 * dtwd_(v, wp)
            return out

        b = bpb + (c2 * dtwd_(np.asarray(prior, np.float64), wp) if prior is not None else 0.0)
        x = cg_solve(A, b, x, n_iter)
    return x
