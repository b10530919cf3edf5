"""Pins the CPU oracle (oracle/recon.py, oracle/bp_oracle.c) against what the
mathematics fixes: closed forms, library routines, published known-answer
vectors and brute force.  The reference ships no tests or data, so these are
the pins (each chosen so that a dropped term, a sign or index error or a
transposed operand fails).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import cbp, recon
from reconstruction_program_taichi_amd import geometry as geo
from tests._cases import cone_case, fan_case, matrix_a_for, oracle_dict, pmatrix_for, rel_l2


# ------------------------------------------------------------------ taps
def _response(h, w, f):
    n = (np.arange(h.shape[0]) - (h.shape[0] - 1) // 2) * w
    return np.real(np.sum(h[None, :] * np.exp(-2j * np.pi * f[:, None] * n[None, :]), axis=1)) * w


def test_ramlak_taps_frequency_response_is_ramp():
    n, w = 512, 0.4
    h = recon.ramlak_taps(n, w, np.float64)
    f = np.linspace(0.05, 0.95, 19) / (2 * w)
    r = _response(h, w, f)
    assert np.allclose(r, np.abs(f), rtol=2e-3, atol=0)
    assert abs(_response(h, w, np.array([0.0]))[0]) < 1e-3 / w  # no DC response


@pytest.mark.parametrize("t", [1.0, 0.54, 0.0])
def test_hamming_taps_frequency_response(t):
    """GenerateHammingKernel blends the ramp with the ramp times a cosine
    window: response |f| (t + (1-t) cos(pi f w)) on |f| < 1/(2w)."""
    n, w = 512, 0.4
    h = geo.hamming_taps(n, w, t, 1000.0).astype(np.float64)
    f = np.linspace(0.05, 0.9, 18) / (2 * w)
    r = _response(h, w, f)
    expect = np.abs(f) * (t + (1 - t) * np.cos(np.pi * f * w))
    assert np.allclose(r, expect, rtol=5e-3, atol=2e-3 / (2 * w))


def test_shepp_logan_taps_frequency_response():
    """Shepp-Logan taps: the ramp times a sinc window, |sin(pi f w)| / (pi w)
    on |f| < 1/(2w) (Shepp & Logan 1974)."""
    n, w = 512, 0.4
    h = geo.shepp_logan_taps(n, w).astype(np.float64)
    f = np.linspace(0.05, 0.95, 19) / (2 * w)
    r = _response(h, w, f)
    assert np.allclose(r, np.abs(np.sin(np.pi * f * w)) / (np.pi * w), rtol=3e-3, atol=0)


def test_hamming_t1_equals_ramlak_oracle():
    h32 = geo.hamming_taps(300, 0.4, 1.0, 1000.0)
    assert np.array_equal(h32, recon.ramlak_taps(300, 0.4, np.float32))


def test_gaussian_apodized_ramp_is_convolution():
    n, w, d = 64, 0.5, 1.5
    full = geo.gaussian_apodized_ramp_taps(n, w, d, 1000.0)
    g = geo.gaussian_taps(n, w, d).astype(np.float64)
    r = geo.hamming_taps(n, w, 1.0, 1000.0).astype(np.float64)
    brute = np.array([sum(g[j] * r[i - (j - (n - 1))] for j in range(2 * n - 1)
                          if 0 <= i - (j - (n - 1)) <= 2 * n - 2) for i in range(2 * n - 1)]) * w
    assert np.allclose(full, brute, rtol=1e-5, atol=1e-7)


# ---------------------------------------------------------------- filter
def test_filter_rows_equals_dense_toeplitz_and_convolve():
    rng = np.random.default_rng(0)
    n = 57
    x = rng.standard_normal((5, n))
    taps = rng.standard_normal(2 * n - 1)
    T = np.array([[taps[k + n - 1 - m] for k in range(n)] for m in range(n)])
    ref = x @ T * 0.7
    assert np.allclose(recon.filter_rows(x, taps, 0.7, np.float64), ref, rtol=1e-12, atol=1e-12)
    got32 = recon.filter_rows(x.astype(np.float32), taps.astype(np.float32), 0.7, np.float32)
    assert rel_l2(got32, ref) < 1e-6


# ------------------------------------------------------------- weighting
def _parker(beta, gamma, scan):
    """Parker weight through the oracle: unit sinogram, cosine weight divided out."""
    sdd = 1000.0
    u = np.array([-np.tan(gamma) * sdd], np.float64)
    ang = np.array([0.0, beta], np.float64)
    s = np.ones((1, 2, 1))
    w = recon.weight_sinogram(s, u, np.zeros(1), ang, sdd, scan, True, dtype=np.float64)
    cosw = sdd * sdd / np.sqrt(sdd * sdd + u[0] ** 2)
    return w[0, 1, 0] / cosw


def test_parker_weights_are_complementary():
    rng = np.random.default_rng(1)
    gmax = 0.25
    scan = np.pi + gmax  # recon.PI differs from pi in the 11th digit; use the reference's PI
    scan = recon.PI + gmax
    for _ in range(50):
        g = rng.uniform(-gmax / 2 * 0.95, gmax / 2 * 0.95)
        b = rng.uniform(0, recon.PI - 2 * g)
        w1 = _parker(b, g, scan)
        w2 = _parker(b + recon.PI + 2 * g, -g, scan)
        assert abs(w1 + w2 - 1.0) < 1e-6


def test_full_scan_weight_is_cosine_only():
    spec = fan_case()
    s = np.ones((spec.n_v, spec.n_views, spec.n_u), np.float32)
    w = recon.weight_sinogram(s, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle, False,
                              dtype=np.float64)
    u = spec.u.astype(np.float64)
    assert np.allclose(w[0, 3], spec.sdd ** 2 / np.sqrt(spec.sdd ** 2 + u ** 2))


# -------------------------------------------------------- back-projection
def _bp_brute(filt, g, v2=False, vol=None):
    """Scalar loops of run_mgfbp.py:1086-1182 on a tiny grid (FP32 scalars).
    v2: the view-by-view form of run_mgfbp_v2.py:71-92, 232-364 instead (start
    from vol, a failed range test sets -10000 and the loop goes on, voxels
    below -5000 end as 0)."""
    f32 = np.float32
    nx, nz = g["img_dim"], g["img_dim_z"]
    nrows, nviews, nu = filt.shape
    c, s, d = recon.view_table(g["angles"], 0.0, np.float32)
    div = recon.div_factor(g["total_scan_angle"], False, np.float32)
    u, v = g["u"], g["v"]
    out = np.zeros((nz, nx, nx), np.float32)
    for iz in range(nz):
        for iy in range(nx):
            for ix in range(nx):
                x = f32(g["pix"]) * (f32(ix) - f32((nx - 1) / 2.0))
                y = (-f32(g["pix"])) * (f32(iy) - f32((nx - 1) / 2.0))
                z = (f32(iz) - f32((nz - 1) / 2.0)) * f32(g["vox_h"]) + f32(g.get("cz", 0.0))
                acc = f32(0) if vol is None else f32(vol[iz, iy, ix])
                for j in range(nviews):
                    L = f32(g["sid"]) - x * c[j] - y * s[j]
                    mag = f32(g["sdd"]) / L
                    ui = (mag * ((-x) * s[j] + y * c[j]) - u[0]) / (u[1] - u[0])
                    if ui < 0 or ui + f32(1) > f32(nu - 1):
                        if v2:
                            acc = f32(-10000)
                            continue
                        acc = None
                        break
                    uf = int(np.floor(ui))
                    ru = ui - f32(uf)
                    vi = (mag * z - v[0]) / f32(g["det_h"]) * np.abs(v[1] - v[0]) / (v[1] - v[0])
                    vf = int(np.floor(vi))
                    if vf < 0 or vf + 1 > nrows - 1:
                        if v2:
                            acc = f32(-10000)
                            continue
                        acc = None
                        break
                    rv = vi - f32(vf)
                    p0 = filt[vf, j, uf] * (f32(1) - ru) + filt[vf, j, uf + 1] * ru
                    p1 = filt[vf + 1, j, uf] * (f32(1) - ru) + filt[vf + 1, j, uf + 1] * ru
                    acc = acc + (f32(g["sid"]) * (f32(1) / (L * L))) * ((f32(1) - rv) * p0 + rv * p1) * d[j] * div
                if v2 and acc < f32(-5000):
                    acc = f32(0)
                out[iz, iy, ix] = 0 if acc is None else acc
    return out


def test_bp_oracle_matches_brute_force_tiny():
    spec = cone_case(nx=6, nz=5, nv=8, nu=20, nrows=10, pix=2.0, vh=2.0, w=1.6, h=1.6)
    g = oracle_dict(spec)
    rng = np.random.default_rng(2)
    f = rng.standard_normal((spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    a = recon.backproject_fbp(f, g, np.float32)
    b = _bp_brute(f, g)
    assert np.array_equal(a, b)


def test_bp_v2_oracle_matches_brute_force_tiny():
    """The view-by-view FBP: equal to the one-pass FBP on ordinary data; with
    values large enough for the -10000 mark and the -5000 clearing to show,
    equal to the scalar loop of run_mgfbp_v2.py (a voxel masked in an early
    view only ends as -10000 plus its later views; sums below -5000 end as 0;
    the views add to the volume they are given)."""
    spec = cone_case(nx=6, nz=5, nv=8, nu=20, nrows=10, pix=2.0, vh=2.0, w=1.6, h=1.6, sid=60.0,
                     sdd=100.0)
    g = oracle_dict(spec)
    rng = np.random.default_rng(5)
    f = rng.standard_normal((spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    one, alive = recon.backproject_fbp(f, g, np.float32, return_mask=True)
    assert 0 < alive.sum() < alive.size  # some voxels leave the detector
    assert np.array_equal(recon.backproject_fbp_v2(f, g), one)
    assert np.array_equal(recon.backproject_fbp_v2(f, g), _bp_brute(f, g, v2=True))
    big = (f + 4.0) * np.float32(4e5)  # positive: masked voxels climb back above -5000
    a = recon.backproject_fbp_v2(big, g)
    assert np.array_equal(a, _bp_brute(big, g, v2=True))
    assert np.any((a != 0) & ~alive)
    neg = -big  # negative: every sum drops below -5000
    b = recon.backproject_fbp_v2(neg, g)
    assert np.array_equal(b, _bp_brute(neg, g, v2=True))
    full = recon.backproject_fbp(neg, g, np.float32)
    assert np.array_equal(b, np.where(full < -5000, np.float32(0), full)) and not b.any()
    vol = rng.uniform(-1, 1, one.shape).astype(np.float32)
    c = recon.backproject_fbp_v2(f, g, vol=vol)
    assert np.array_equal(c, _bp_brute(f, g, v2=True, vol=vol))
    assert not np.array_equal(c, one)


@pytest.mark.parametrize("kind", ["cone", "fan"])
def test_c_oracle_matches_numpy_oracle(kind):
    spec = cone_case() if kind == "cone" else fan_case()
    g = oracle_dict(spec)
    rng = np.random.default_rng(3)
    f = rng.standard_normal((spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    assert np.array_equal(recon.backproject_fbp(f, g, np.float32), cbp.backproject_fbp(f, g))


def _disc_fbp(nviews=360, nu=256, nx=64):
    """FBP (FP64 oracle) of the analytic fan-beam sinogram of a uniform disc."""
    R, mu = 10.0, 0.02
    pix, w, sid, sdd = 0.5, 0.25, 200.0, 400.0
    spec = fan_case(nx=nx, nv=nviews, nu=nu, pix=pix, w=w, sid=sid, sdd=sdd, rows=1)
    u = spec.u.astype(np.float64)
    gam = np.arctan2(u, sdd)
    dd = sid * np.abs(np.sin(gam))  # distance of each ray from the isocentre
    chord = 2 * np.sqrt(np.clip(R * R - dd * dd, 0, None)) * mu
    sgm = np.broadcast_to(chord[None, None, :], (1, nviews, nu)).astype(np.float64)
    ws = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, sdd, spec.total_scan_angle, False,
                               dtype=np.float64)
    fl = recon.filter_rows(ws, recon.ramlak_taps(nu, w), w, np.float64)
    g = oracle_dict(spec)
    return recon.backproject_fbp(fl, g, np.float64), mu, R, pix, nx


def test_fbp_of_disc_reproduces_disc():
    img, mu, R, pix, nx = _disc_fbp()
    c = (nx - 1) / 2.0
    yy, xx = np.mgrid[0:nx, 0:nx]
    r = np.hypot((xx - c) * pix, (yy - c) * pix)
    inner = img[0][r < 0.7 * R]
    outer = img[0][(r > 1.3 * R) & (r < 14.0)]
    assert abs(inner.mean() / mu - 1) < 0.02
    assert inner.std() / mu < 0.03
    assert np.abs(outer).mean() / mu < 0.03


@pytest.mark.parametrize("rows", [1, 5, 40])
def test_filter_vertical_is_a_truncated_convolution(rows):
    """The vertical Gaussian pass equals np.convolve along the row axis with
    the taps outside |offset| <= 10 zeroed, times the row height; a Gaussian
    normalised to 1/h leaves a constant column constant on rows at least 10
    away from both ends."""
    h = 0.7
    kern = geo.gaussian_taps(rows, h, 1.5).astype(np.float64)
    rng = np.random.default_rng(rows)
    x = rng.standard_normal((rows, 3, 4))
    kt = kern.copy()
    off = np.arange(2 * rows - 1) - (rows - 1)
    kt[np.abs(off) > 10] = 0
    want = np.apply_along_axis(lambda c: np.convolve(c, kt)[rows - 1:2 * rows - 1], 0, x) * h
    assert np.allclose(recon.filter_vertical(x, kern, h), want, atol=1e-12)
    const = recon.filter_vertical(np.ones((rows, 1, 1)), kern, h)[:, 0, 0]
    interior = const[10:rows - 10]
    assert np.allclose(interior, 1.0, atol=1e-6)


def test_fbp_of_offcentre_disc_reproduces_it_in_place():
    """Analytic fan-beam line integrals of an off-centre disc (distance of each
    source-to-element line from the disc centre), FBP'd by the oracle: the
    disc comes back at its position, not mirrored or rotated."""
    nx, nviews, nu = 64, 360, 256
    R, mu, pix, w, sid, sdd = 5.0, 0.02, 0.5, 0.25, 200.0, 400.0
    cxy = np.array([6.0, -4.0])  # object coordinates, mm
    spec = fan_case(nx=nx, nv=nviews, nu=nu, pix=pix, w=w, sid=sid, sdd=sdd, rows=1)
    u = spec.u.astype(np.float64)
    sgm = np.zeros((1, nviews, nu))
    for j, b in enumerate(spec.angles.astype(np.float64)):
        cb, sb = np.cos(b), np.sin(b)
        S = sid * np.array([cb, sb])
        D = np.stack([(sid - sdd) * cb - u * sb, (sid - sdd) * sb + u * cb], axis=1)
        e = D - S
        d = np.abs(e[:, 0] * (cxy[1] - S[1]) - e[:, 1] * (cxy[0] - S[0])) / np.hypot(e[:, 0], e[:, 1])
        sgm[0, j] = 2 * np.sqrt(np.clip(R * R - d * d, 0, None)) * mu
    ws = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, sdd, spec.total_scan_angle, False,
                               dtype=np.float64)
    fl = recon.filter_rows(ws, recon.ramlak_taps(nu, w), w, np.float64)
    img = recon.backproject_fbp(fl, oracle_dict(spec), np.float64)[0]
    xs, ys = _image_xy(nx, pix)
    X, Y = np.meshgrid(xs, ys)
    r = np.hypot(X - cxy[0], Y - cxy[1])
    assert abs(img[r < 0.6 * R].mean() / mu - 1) < 0.03
    for mirror in ([-1, 1], [1, -1], [-1, -1]):
        rm = np.hypot(X - mirror[0] * cxy[0], Y - mirror[1] * cxy[1])
        assert np.abs(img[rm < 0.6 * R]).mean() / mu < 0.05


# ---------------------------------------------------- forward projection
def test_fp_of_uniform_disc_equals_chord_length():
    """FP (FP64 oracle) of a slice holding a centred uniform disc: chord
    length times mu, within the sampling/interpolation error."""
    nx, pix, R, mu = 96, 0.5, 18.0, 0.02
    c = (nx - 1) / 2.0
    yy, xx = np.mgrid[0:nx, 0:nx]
    ss = 4
    img = np.zeros((nx, nx))
    for a in range(ss):
        for b in range(ss):
            img += np.hypot((xx - c + (a + 0.5) / ss - 0.5) * pix, (yy - c + (b + 0.5) / ss - 0.5) * pix) <= R
    img = img / ss ** 2 * mu
    spec = fan_case(nx=nx, nv=12, nu=160, pix=pix, w=0.35, sid=150.0, sdd=300.0, rows=1)
    g = oracle_dict(spec)
    sgm = recon.forward_project(img[None], g, dtype=np.float64)
    u = spec.u.astype(np.float64)
    dd = 150.0 * np.abs(np.sin(np.arctan2(u, 300.0)))
    chord = 2 * np.sqrt(np.clip(R * R - dd * dd, 0, None)) * mu
    inside = dd < R - 1.0
    err = np.abs(sgm[0][:, inside] - chord[None, inside]) / (chord[None, inside])
    assert err.max() < 0.02


def test_fp_oracle_fp32_close_to_fp64():
    spec = cone_case(nx=24, nz=12, nv=10, nu=40, nrows=16)
    g = oracle_dict(spec)
    rng = np.random.default_rng(4)
    vol = rng.uniform(0, 1, (12, 24, 24))
    a = recon.forward_project(vol.astype(np.float32), g, dtype=np.float32)
    b = recon.forward_project(vol, g, dtype=np.float64)
    assert rel_l2(a, b) < 1e-5


def test_fp_is_linear():
    spec = fan_case(nx=32, nv=8, nu=50)
    g = oracle_dict(spec)
    rng = np.random.default_rng(5)
    a, b = rng.uniform(0, 1, (2, 32, 32)), rng.uniform(0, 1, (2, 32, 32))
    fa = recon.forward_project(a, g, dtype=np.float64)
    fb = recon.forward_project(b, g, dtype=np.float64)
    fab = recon.forward_project(2 * a - 3 * b, g, dtype=np.float64)
    assert np.allclose(fab, 2 * fa - 3 * fb, atol=1e-10)


def _image_xy(nx, pix):
    """Object coordinates of pixel centres: x grows with the column index,
    y falls with the row index (run_mgfbp.py:1086-1104)."""
    h = (nx - 1) / 2.0
    i = np.arange(nx)
    return pix * (i - h), -pix * (i - h)


def _project_point(x, y, beta, sid, sdd):
    """Flat-detector coordinate of the ray through object point (x, y) for a
    gantry at angle beta: source at sid (cos b, sin b), detector line at
    distance sdd; the point seen from the rotated frame is R(-b) (x, y)."""
    xg = x * np.cos(beta) + y * np.sin(beta)
    yg = -x * np.sin(beta) + y * np.cos(beta)
    return sdd * yg / (sid - xg)


def test_fp_of_offcentre_blob_peaks_at_its_projection():
    """The FP oracle's sinogram of an off-centre blob is centred, in every
    view, on the detector coordinate where the blob's centre projects: pins
    the rotation direction, the detector orientation and the image axes."""
    nx, pix = 64, 0.5
    spec = fan_case(nx=nx, nv=24, nu=200, pix=pix, w=0.4, sid=200.0, sdd=400.0, rows=1)
    g = oracle_dict(spec)
    ix, iy, s = 47, 18, 1.5
    yy, xx = np.mgrid[0:nx, 0:nx]
    img = np.exp(-((xx - ix) ** 2 + (yy - iy) ** 2) / (2 * s * s))
    sgm = recon.forward_project(img[None], g, dtype=np.float64)[0]
    xs, ys = _image_xy(nx, pix)
    u = spec.u.astype(np.float64)
    for j, beta in enumerate(spec.angles.astype(np.float64)):
        want = _project_point(xs[ix], ys[iy], beta, 200.0, 400.0)
        got = float((sgm[j] * u).sum() / sgm[j].sum())
        assert abs(got - want) < 0.25 * 0.4, (j, got, want)


def test_bp_ir_of_a_ray_pair_is_the_band_between_the_rays():
    """Unfiltered IR back-projection of a sinogram holding ones at detector
    elements k, k+1 of one view: every pixel strictly between the two rays
    (lines from the source through the element centres) receives the chord
    length of its source ray through a pixel column, pix / max(|cos|, |sin|);
    pixels beyond the neighbouring rays receive nothing; pixels that leave the
    detector in some view are zero (full-scan FOV disc)."""
    nx, pix, sid, sdd, w = 64, 0.5, 200.0, 400.0, 0.5
    spec = fan_case(nx=nx, nv=8, nu=120, pix=pix, w=w, sid=sid, sdd=sdd, rows=1)
    g = oracle_dict(spec)
    j, k = 3, 52
    sgm = np.zeros((1, 8, 120))
    sgm[0, j, k] = sgm[0, j, k + 1] = 1.0
    bp = recon.backproject_ir(sgm, g, dtype=np.float64)[0]
    beta = float(spec.angles[j])
    u = spec.u.astype(np.float64)
    xs, ys = _image_xy(nx, pix)
    X, Y = np.meshgrid(xs, ys)  # [row][col]
    up = _project_point(X, Y, beta, sid, sdd)
    r_fov = sid * np.sin(np.arctan(np.abs(u).max() / sdd))
    rad = np.hypot(X, Y)
    lo, hi = min(u[k], u[k + 1]), max(u[k], u[k + 1])
    inside = (up > lo + 0.05 * w) & (up < hi - 0.05 * w) & (rad < r_fov - 2 * pix)
    dxs, dys = X - sid * np.cos(beta), Y - sid * np.sin(beta)
    ang = np.arctan2(dys, dxs)
    chord = pix / np.maximum(np.abs(np.cos(ang)), np.abs(np.sin(ang)))
    assert inside.sum() > 20
    assert np.allclose(bp[inside], chord[inside], rtol=1e-9)
    outside = (up < min(u[k - 1], u[k + 2]) - 0.05 * w) | (up > max(u[k - 1], u[k + 2]) + 0.05 * w)
    assert np.all(bp[outside] == 0)
    ones = recon.backproject_ir(np.ones((1, 8, 120)), g, dtype=np.float64)[0]
    ups = np.stack([_project_point(X, Y, b, sid, sdd) for b in spec.angles.astype(np.float64)])
    # the range test keeps (u - u[0]) / du in [0, n_u - 2] (run_mgfbp_iterative_
    # recon_IRN_CGLS.py:439-441): the interval between the last two elements
    # counts as off the detector
    ulo, uhi = sorted((u[0], u[-2]))
    margin = np.minimum(ups - ulo, uhi - ups).min(axis=0)  # < 0: off the detector
    assert np.all(ones[margin < -0.1 * w] == 0)
    assert np.all(ones[margin > 0.1 * w] > 0)
    assert (margin < -0.1 * w).sum() > 100


def test_ray_mask_matches_cylinder_intersection():
    """GenEffectiveMapForwardProjectionAgent's cone-beam map against exact
    geometry: a ray is masked iff its segment inside the image's inscribed
    cylinder leaves the z range between the first and last slice centres.
    Rays within a pixel (in-plane) or a slice (in z) of either boundary are
    not classified."""
    nx, nz, pix, vh, h = 32, 8, 0.5, 0.5, 0.5
    sid, sdd = 150.0, 300.0
    spec = cone_case(nx=nx, nz=nz, nv=6, nu=60, nrows=40, pix=pix, vh=vh, w=0.5, h=h, sid=sid,
                     sdd=sdd)
    g = oracle_dict(spec)
    m = recon.ray_mask(g, dtype=np.float64)
    R = nx * pix / 2
    z_lo, z_hi = -(nz - 1) / 2 * vh, (nz - 1) / 2 * vh
    u = spec.u.astype(np.float64)
    v = spec.v.astype(np.float64)
    # gantry frame: source (sid, 0, 0), element (sid - sdd, u, v); the cylinder
    # is centred on the rotation axis, so the view angle does not matter
    U, V = np.meshgrid(u, v)  # [row][col]
    dx, dy = -sdd, U
    a = dx * dx + dy * dy
    b = 2 * sid * dx
    def zrange(rad):
        c = sid * sid - rad * rad
        disc = b * b - 4 * a * c
        hit = disc > 0
        sq = np.sqrt(np.where(hit, disc, 0))
        t1, t2 = (-b - sq) / (2 * a), (-b + sq) / (2 * a)
        z1, z2 = t1 * V, t2 * V
        return hit, np.minimum(z1, z2), np.maximum(z1, z2)
    def kept(rad, dz):  # dz > 0 widens the allowed z range
        hit, zmin, zmax = zrange(rad)
        return ~(hit & ((zmin < z_lo - dz) | (zmax > z_hi + dz)))

    # robust: the same verdict for the smaller cylinder with the wider z range
    # and for the larger cylinder with the narrower one
    small_wide, big_narrow = kept(R - 2 * pix, vh), kept(R + 2 * pix, -vh)
    sure = small_wide == big_narrow
    want = small_wide
    assert sure.mean() > 0.6 and want[sure].any() and (~want[sure]).any()
    for jv in range(spec.n_views):
        assert np.array_equal(m[:, jv][sure] == 1, want[sure]), jv


# ------------------------------------------------------------------ RNG
def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors (Salmon et al. 2011, Random123)."""
    cases = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in cases:
        got = recon.philox4x32_10(np.array([ctr], np.uint32), np.array([key], np.uint32))[0]
        assert [int(x) for x in got] == want


def test_normals_statistics():
    g = recon.normals(200000, seed=7)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01


# ---------------------------------------------------------- CG / TV
def test_cg_recursion_solves_spd_system():
    rng = np.random.default_rng(6)
    M = rng.standard_normal((30, 30))
    A = M @ M.T + 30 * np.eye(30)
    b = rng.standard_normal(30)
    x = recon.cg_solve(lambda v: A @ v, b, np.zeros(30), 40)
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)


def _diff_op(n):
    """Forward difference with a zero last row (the reference's D)."""
    D = sp.lil_matrix((n, n))
    for i in range(n - 1):
        D[i, i], D[i, i + 1] = -1, 1
    return D.tocsr()


@pytest.mark.parametrize("nz", [1, 2, 5])
def test_dtwd_matches_sparse_operator_interior(nz):
    ny, nx = 7, 6
    rng = np.random.default_rng(nz)
    x = rng.standard_normal((nz, ny, nx))
    wr = rng.uniform(0.5, 2, (nz, ny, nx))
    I = lambda n: sp.identity(n, format="csr")
    Dy = sp.kron(sp.kron(I(nz), _diff_op(ny)), I(nx))
    Dx = sp.kron(sp.kron(I(nz), I(ny)), _diff_op(nx))
    W = sp.diags(wr.ravel())
    ops = [Dy, Dx]
    if nz > 2:
        ops.append(sp.kron(sp.kron(_diff_op(nz), I(ny)), I(nx)))
    full = sum((D.T @ (W @ (D @ x.ravel()))) for D in ops).reshape(x.shape)
    got = recon.dtwd(x, wr)
    # The reference's operator equals D^T W D except that each axis' first and
    # last index receive no contribution from that axis.
    sl = (slice(1, -1) if nz > 2 else slice(None), slice(1, -1), slice(1, -1))
    assert np.allclose(got[sl], full[sl], atol=1e-12)
    if nz <= 2:  # row y = 0 only sees the x-axis term
        xonly = (Dx.T @ (W @ (Dx @ x.ravel()))).reshape(x.shape)
        assert np.allclose(got[:, 0, 1:-1], xonly[:, 0, 1:-1], atol=1e-12)


def test_tv_weight_and_value():
    rng = np.random.default_rng(8)
    x = rng.standard_normal((3, 4, 5))
    beta = 1e-3
    ux = np.zeros_like(x)
    ux[:, :-1] = np.diff(x, axis=1)
    uy = np.zeros_like(x)
    uy[:, :, :-1] = np.diff(x, axis=2)
    uz = np.zeros_like(x)
    uz[:-1] = np.diff(x, axis=0)
    mag = np.sqrt(ux ** 2 + uy ** 2 + uz ** 2 + beta)
    assert np.allclose(recon.tv_weight(x, beta), 1 / mag)
    assert np.isclose(recon.tv_value(x, beta), mag.sum())
    # two slices: no z coupling (reference: only when Nz > 2)
    x2 = x[:2]
    mag2 = np.sqrt(ux[:2] ** 2 + uy[:2] ** 2 + beta)
    assert np.allclose(recon.tv_weight(x2, beta), 1 / mag2)


# ------------------------------------------------------------ P-matrix
def _pm_spec(rot=0.0, off_v=0.0):
    # pitches and offsets exact in FP32, so u and v are exactly linear in the index
    spec = cone_case(nx=24, nz=12, nv=24, nu=60, nrows=30, pix=0.5, vh=0.5, w=0.5, h=0.5,
                     sid=150.0, sdd=300.0, off_v=off_v)
    if rot:
        spec.img_rot = rot
        spec.angles = geo.angle_array(spec.n_views, spec.total_scan_angle, rot)
    return spec


@pytest.mark.parametrize("rot", [0.0, 0.375])
def test_pmatrix_bp_equals_standard_bp(rot):
    """The P-matrix branch of the back-projector with the P-matrices of the
    standard circular geometry (built from first principles) reproduces the
    standard branch: same masks, same values (FP64).  A swapped P row, a
    wrong sign or a transposed matrix moves every voxel's projection."""
    spec = _pm_spec(rot, off_v=0.75)
    g = oracle_dict(spec)
    rng = np.random.default_rng(11)
    filt = rng.standard_normal((spec.n_v, spec.n_views, spec.n_u))
    ref, alive = recon.backproject_fbp(filt, g, np.float64, return_mask=True)
    gp = dict(g, pmatrix=pmatrix_for(spec))
    got, alive_p = recon.backproject_fbp(filt, gp, np.float64, return_mask=True)
    assert np.array_equal(alive, alive_p) and alive.mean() > 0.3
    assert rel_l2(got, ref) <= 1e-9
    ir = recon.backproject_ir(filt, g, dtype=np.float64)
    ir_p = recon.backproject_ir(filt, gp, dtype=np.float64)
    assert rel_l2(ir_p, ir) <= 1e-9 and np.array_equal(ir != 0, ir_p != 0)


@pytest.mark.parametrize("rot", [0.0, 0.375])
def test_pmatrix_fp_and_mask_equal_standard(rot):
    """Forward projection and ray mask through the P-matrix rays
    (s A_j [u, v, 1] + x_s_j, rotated by angles[0]) with A, x_s of the
    standard geometry equal the standard rays (FP64)."""
    spec = _pm_spec(rot)
    spec.fpj_step = 0.5
    g = oracle_dict(spec)
    vol = np.random.default_rng(12).uniform(0, 1, (spec.img_dim_z, spec.img_dim, spec.img_dim))
    # the forward projector's frame: rotation by angles[0] happens after A
    A, xs = matrix_a_for(pmatrix_for(spec, rot=float(spec.angles[0])), spec.n_views)
    gp = dict(g, matrix_a=A, x_s=xs)
    ref = recon.forward_project(vol, g, dtype=np.float64)
    got = recon.forward_project(vol, gp, dtype=np.float64)
    assert rel_l2(got, ref) <= 1e-9 and (ref > 0).mean() > 0.3
    m = recon.ray_mask(g, dtype=np.float64)
    mp = recon.ray_mask(gp, dtype=np.float64)
    assert np.array_equal(m, mp) and 0.05 < (m == 0).mean() < 0.95


def test_pmatrix_fp32_reference_order_close_to_standard():
    """FP32 reference order: the P-matrix branch differs from the standard
    one only by rounding of the matrices and the operation order."""
    spec = _pm_spec()
    g = oracle_dict(spec)
    filt = np.random.default_rng(13).standard_normal((spec.n_v, spec.n_views, spec.n_u))
    ref, alive = recon.backproject_fbp(filt, g, np.float32, return_mask=True)
    gp = dict(g, pmatrix=pmatrix_for(spec).astype(np.float32))
    got, alive_p = recon.backproject_fbp(filt, gp, np.float32, return_mask=True)
    both = alive & alive_p
    assert (alive != alive_p).mean() < 1e-3
    assert rel_l2(got[both], ref[both]) <= 1e-5


# ------------------------------------------------------------ helical rebinning
HEL = dict(rows=16, views=400, nu=24, det_h=1.1, sdd=800.0, mag=1.6, dz=0.05, off_v=0.7)


def _helical_rows_at_isocentre(rows, views, det_h, mag, dz, off_v):
    """Height at the isocentre of the ray to detector row r in view k: the
    source and detector rise by dz per view; row r sits at
    v = ((rows-1)/2 - r) det_h above the detector centre (row 0 on top), plus
    the vertical offset, and the isocentre sees heights divided by mag."""
    r = np.arange(rows)[:, None]
    k = np.arange(views)[None, :]
    v = ((rows - 1) / 2.0 - r) * det_h
    return k * dz + (v + off_v) / mag


def test_helical_rebin_recovers_a_function_of_height():
    """Helical projections of an object whose line integrals depend only on
    the height where the ray crosses the isocentre (a * z + b, divided by the
    cone-angle factor sdd / |d - s|): rebinning to slice z returns a * z + b
    in every view that sees z, and 0 in views past the scan."""
    h = HEL
    zi = _helical_rows_at_isocentre(h["rows"], h["views"], h["det_h"], h["mag"], h["dz"], h["off_v"])
    vrow = ((h["rows"] - 1) / 2.0 - np.arange(h["rows"]))[:, None] * h["det_h"]
    cosg = h["sdd"] / np.sqrt(h["sdd"] ** 2 + vrow ** 2)
    a, b = 0.3, 2.0
    sgm = np.broadcast_to(((a * zi + b) / cosg)[:, :, None], (h["rows"], h["views"], h["nu"])).copy()
    z = np.array([2.0, 5.5, 9.25, 18.0, 19.6])
    vc = 120
    reb = recon.helical_rebin(sgm, z, vc, h["off_v"], h["det_h"], h["sdd"], h["mag"], h["dz"], np.float64)
    for n, zz in enumerate(z):
        k0 = (zz - h["off_v"] / h["mag"]) / h["dz"]
        ks = np.floor(k0 - vc / 2 + 0.5) + np.arange(vc)
        # the rebinned views are the view_count consecutive views whose centre
        # lies within one view below k0 (floor(k0 - view_count/2 + 1/2))
        assert k0 - 1 < ks.mean() <= k0 + 1e-9
        v_pos = (zz - ks * h["dz"]) * h["mag"] - h["off_v"]
        inside = (ks >= 0) & (ks <= h["views"] - 2) & (np.abs(v_pos) < ((h["rows"] - 1) / 2.0) * h["det_h"])
        past = (ks < 0) | (ks > h["views"] - 2)
        assert inside.sum() > 20
        # the interpolated cone factor differs from the exact one by < 1e-6 here
        assert np.allclose(reb[n][inside], a * zz + b, rtol=1e-5)
        assert np.all(reb[n][past] == 0)


def test_helical_rebin_clamps_to_the_edge_rows():
    """Heights beyond the detector's reach take the first / last row."""
    h = HEL
    sgm = np.zeros((h["rows"], h["views"], 4))
    sgm[0] = 1.0
    sgm[-1] = 3.0
    z = np.array([10.0])
    reb = recon.helical_rebin(sgm, z, 360, h["off_v"], h["det_h"], h["sdd"], h["mag"], h["dz"], np.float64)
    k0 = (10.0 - h["off_v"] / h["mag"]) / h["dz"]
    ks = np.floor(k0 - 180 + 0.5) + np.arange(360)
    v_pos = (10.0 - ks * h["dz"]) * h["mag"] - h["off_v"]
    w = h["sdd"] / np.sqrt(h["sdd"] ** 2 + v_pos ** 2)
    seen = (ks >= 0) & (ks <= h["views"] - 2)
    top = seen & (v_pos >= ((h["rows"] - 1) / 2.0) * h["det_h"])  # row index <= 0: row 0
    bottom = seen & (v_pos <= -((h["rows"] - 1) / 2.0) * h["det_h"])
    assert top.any() and bottom.any()
    assert np.allclose(reb[0, top, 0], 1.0 * w[top]) and np.allclose(reb[0, bottom, 0], 3.0 * w[bottom])


# ------------------------------------- C restatements of FP and the IR BP
def _fp_cases():
    cone = cone_case(nx=20, nz=10, nv=7, nu=26, nrows=12, pix=0.5, vh=0.6, w=0.8, h=1.1, sid=60.0,
                     sdd=120.0)
    cone.cz = 0.3
    curved = fan_case(nx=24, nv=9, nu=40, pix=0.5, w=0.5, sid=60.0, sdd=120.0, rows=3)
    curved.curved = True
    helical = cone_case(nx=16, nz=12, nv=6, nu=20, nrows=8, pix=0.5, vh=0.5, w=0.8, h=0.9, sid=60.0,
                        sdd=120.0)
    percz = cone_case(nx=16, nz=8, nv=5, nu=20, nrows=8, pix=0.5, vh=0.5, w=0.8, h=0.9, sid=60.0,
                      sdd=120.0)
    percz.center_z = np.linspace(-0.6, 0.8, 5).astype(np.float32)
    gs = []
    for spec, extra in ((cone, {}), (curved, {}), (helical, {"zpv": 0.37}), (percz, {})):
        g = oracle_dict(spec)
        g.update(extra)
        gs.append(g)
    return gs


@pytest.mark.parametrize("case", range(4))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_c_forward_projector_matches_numpy_oracle(case, dtype):
    g = _fp_cases()[case]
    rng = np.random.default_rng(40 + case)
    vol = rng.uniform(0, 1, (g["img_dim_z"], g["img_dim"], g["img_dim"]))
    for count_mode in (0, 1):
        a = cbp.forward_project(vol, g, count_mode=count_mode, dtype=dtype)
        b = recon.forward_project(vol, g, count_mode=count_mode, dtype=dtype)
        assert a.shape == b.shape and np.abs(b).max() > 1
        if dtype == np.float64:
            assert np.allclose(a, b, rtol=1e-12, atol=1e-12)
        else:
            assert rel_l2(a, b) < 1e-6


def test_c_forward_projector_oversample_and_mask():
    g = _fp_cases()[0]
    rng = np.random.default_rng(45)
    vol = rng.uniform(0, 1, (g["img_dim_z"], g["img_dim"], g["img_dim"]))
    m = (rng.uniform(0, 1, (12, 7, 13)) > 0.3).astype(np.float64)
    g2 = dict(g, u=geo.det_pos_array(26, 26, -0.4, 0.0))
    a = cbp.forward_project(vol, g2, oversample=2, ray_mask=m, dtype=np.float64)
    b = recon.forward_project(vol, g2, oversample=2, ray_mask=m, dtype=np.float64)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12) and (a == 0).any()


@pytest.mark.parametrize("kind", ["cone", "fan", "percz", "rot"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_c_ir_backprojector_matches_numpy_oracle(kind, dtype):
    if kind == "fan":
        spec = fan_case(nx=24, nv=16, nu=40, rows=3)
    elif kind == "rot":
        spec = fan_case(nx=24, nv=16, nu=40, rows=2, rot_deg=17.0)
    else:
        spec = cone_case(nx=20, nz=10, nv=12, nu=36, nrows=16)
        if kind == "percz":
            spec.center_z = np.linspace(-0.8, 0.5, 12).astype(np.float32)
    g = oracle_dict(spec)
    rng = np.random.default_rng(50)
    sgm = rng.standard_normal((spec.n_v, spec.n_views, spec.n_u))
    mask = (rng.uniform(0, 1, sgm.shape) > 0.2).astype(np.float64)
    a = cbp.backproject_ir(sgm, g, ray_mask=mask, dtype=dtype)
    b = recon.backproject_ir(sgm, g, ray_mask=mask, dtype=dtype)
    assert (b == 0).any() and np.abs(b).max() > 0.1
    if dtype == np.float64:
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12)
    else:
        assert rel_l2(a, b) < 1e-6
        assert np.array_equal(a == 0, b == 0)


def test_ir_solve_with_c_projectors_equals_numpy_projectors():
    spec = fan_case(nx=16, nv=20, nu=30, rows=3)
    g = oracle_dict(spec)
    rng = np.random.default_rng(60)
    sgm = rng.uniform(0, 1, (3, 20, 30))
    x0 = rng.uniform(0, 0.1, (3, 16, 16))
    a = recon.ir_solve(sgm, g, x0, 1e-2, 1e-4, 1, 3)
    b = recon.ir_solve(sgm, g, x0, 1e-2, 1e-4, 1, 3, projectors=(cbp.forward_project, cbp.backproject_ir))
    assert rel_l2(b, a) < 1e-11


def test_c_filter_equals_numpy_oracle_bit_for_bit():
    """The C restatement of the FP32 ramp filter (CPU baseline leg) sums in
    the numpy oracle's order: identical FP32 results, on 1 and on all
    threads."""
    rng = np.random.default_rng(70)
    x = rng.standard_normal((3, 5, 64)).astype(np.float32)
    taps = recon.ramlak_taps(64, 0.4)
    want = recon.filter_rows(x, taps, 0.4, np.float32)
    n0 = cbp.max_threads()
    try:
        for t in (1, n0):
            cbp.set_threads(t)
            assert np.array_equal(cbp.filter_rows(x, taps, 0.4), want)
    finally:
        cbp.set_threads(n0)
