"""Oracle pins for the branches that plain restatement leaves open: the cone
and curved weighting, the curved back-projection and forward projection, the
cone forward projector's trilinear sampling, the noise model, and the IR /
PICCS normal equations.

Each pin is a function of the oracle module, so tests/test_oracle_mutations.py
can run the same pin against copies of the oracle with one plausible mistake
planted and check that it fails.  Every expected value comes from an
independent route: vector geometry (tests/_analytic.py), analytic line
integrals, a library interpolator (scipy.ndimage.map_coordinates), explicit
matrices and sparse difference operators, or sampling statistics.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
from scipy import ndimage

from oracle import recon
from reconstruction_program_taichi_amd import geometry as geo
from tests import _analytic as an
from tests._cases import cone_case, fan_case, oracle_dict, rel_l2

PI = geo.PI


# ----------------------------------------------------------- weighting
def pin_cone_weight_is_ray_cosine(R):
    """WeightSgm (run_mgfbp.py:962-970) multiplies each sample by sdd times
    the cosine of the angle between its ray and the central ray.  The cosine
    comes from the source and element positions (flat plane or arc of radius
    sdd around the source) as a normalised dot product, for rows with v != 0."""
    sid, sdd = 300.0, 500.0
    u = np.linspace(-150.0, 140.0, 23)
    v = np.linspace(-90.0, 120.0, 11)
    ang = np.array([0.0, 0.3], np.float64)
    ones = np.ones((v.size, 2, u.size))
    for curved in (False, True):
        gx, gy, gz = an.element_gantry(u[None, :], v[:, None], sid, sdd, curved)
        ray = np.stack([gx - sid, gy, gz], axis=-1)
        cx, cy, cz = an.element_gantry(np.zeros(1), np.zeros(1), sid, sdd, curved)
        central = np.array([cx[0] - sid, cy[0], cz[0]])
        cosang = (ray @ central) / (np.linalg.norm(ray, axis=-1) * np.linalg.norm(central))
        w = R.weight_sinogram(ones, u, v, ang, sdd, 2 * PI, False, curved=curved, dtype=np.float64)
        assert np.allclose(w[:, 1, :], sdd * cosang, rtol=1e-12), curved


# --------------------------------------------- FDK of a z-invariant object
def _fdk_cylinder(R, curved):
    """Cone-beam FDK (weight -> ramp -> back-project) of the analytic
    projections of an infinite z-parallel cylinder.  FDK is exact for objects
    that do not vary along z, so every slice, at every cone angle, must
    reproduce the cylinder's value."""
    nx, nz, pix, vh = 40, 12, 0.6, 4.0
    nv, nu, nrows, w, h = 180, 96, 48, 0.5, 2.4
    sid, sdd = 100.0, 200.0
    spec = cone_case(nx=nx, nz=nz, nv=nv, nu=nu, nrows=nrows, pix=pix, vh=vh, w=w, h=h,
                     sid=sid, sdd=sdd)
    spec.curved = curved
    cx, cy, r, mu = 2.0, -1.5, 6.0, 0.02
    S, E = an.rays(spec.angles, spec.u, spec.v, sid, sdd, curved)
    sgm = an.cylinder_integrals(S, E, cx, cy, r, mu)
    ws = R.weight_sinogram(sgm, spec.u, spec.v, spec.angles, sdd, spec.total_scan_angle, False,
                           curved=curved, dtype=np.float64)
    taps = geo.hamming_taps(nu, w, 1.0, sdd, curved=curved).astype(np.float64)
    fl = R.filter_rows(ws, taps, w, np.float64)
    vol = R.backproject_fbp(fl, oracle_dict(spec), np.float64)
    xs, ys = an.pixel_centres(nx, pix)
    X, Y = np.meshgrid(xs, ys)
    inner = np.hypot(X - cx, Y - cy) < 0.6 * r
    outer = (np.hypot(X - cx, Y - cy) > 1.4 * r) & (np.hypot(X, Y) < 9.0)
    return vol, inner, outer, mu


def pin_fdk_of_z_invariant_cylinder(R, curved=False):
    vol, inner, outer, mu = _fdk_cylinder(R, curved)
    for k in range(vol.shape[0]):
        m = vol[k][inner].mean() / mu
        assert abs(m - 1) < 0.002, (k, m)
        assert np.abs(vol[k][outer]).mean() / mu < 0.02, k


# ----------------------------------------------------- curved detector
def pin_curved_fbp_of_offcentre_disc(R):
    """Fan FBP on an arc detector of the analytic projections of an
    off-centre disc: the disc comes back at its place (not mirrored) with
    its value, using the arc taps (run_mgfbp.py:892-896), the arc weighting
    (:965-967) and the arc back-projection (:1128-1129)."""
    nx, nv, nu = 64, 360, 256
    rad, mu, pix, w, sid, sdd = 5.0, 0.02, 0.5, 0.25, 100.0, 200.0
    cxy = (6.0, -4.0)
    spec = fan_case(nx=nx, nv=nv, nu=nu, pix=pix, w=w, sid=sid, sdd=sdd, rows=1)
    spec.curved = True
    S, E = an.rays(spec.angles, spec.u, spec.v, sid, sdd, curved=True)
    sgm = an.cylinder_integrals(S, E, cxy[0], cxy[1], rad, mu)
    ws = R.weight_sinogram(sgm, spec.u, spec.v, spec.angles, sdd, spec.total_scan_angle, False,
                           curved=True, dtype=np.float64)
    taps = geo.hamming_taps(nu, w, 1.0, sdd, curved=True).astype(np.float64)
    img = R.backproject_fbp(R.filter_rows(ws, taps, w, np.float64), oracle_dict(spec), np.float64)[0]
    xs, ys = an.pixel_centres(nx, pix)
    X, Y = np.meshgrid(xs, ys)
    assert abs(img[np.hypot(X - cxy[0], Y - cxy[1]) < 0.6 * rad].mean() / mu - 1) < 0.02
    for mx, my in ((-1, 1), (1, -1), (-1, -1)):
        far = np.hypot(X - mx * cxy[0], Y - my * cxy[1]) < 0.6 * rad
        assert np.abs(img[far]).mean() / mu < 0.05, (mx, my)


def pin_curved_fp_blob_on_arc(R):
    """The arc-detector forward projection of an off-centre blob is centred,
    in every view, at sdd times the angle of the blob centre's ray from the
    central ray (run_mgfpj.py:531-535)."""
    nx, pix, sid, sdd, w = 64, 0.5, 100.0, 200.0, 0.3
    spec = fan_case(nx=nx, nv=24, nu=200, pix=pix, w=w, sid=sid, sdd=sdd, rows=1)
    spec.curved = True
    ix, iy, s = 47, 18, 1.5
    yy, xx = np.mgrid[0:nx, 0:nx]
    img = np.exp(-((xx - ix) ** 2 + (yy - iy) ** 2) / (2 * s * s))
    sgm = R.forward_project(img[None], oracle_dict(spec), dtype=np.float64)[0]
    xs, ys = an.pixel_centres(nx, pix)
    u = spec.u.astype(np.float64)
    for j, beta in enumerate(spec.angles.astype(np.float64)):
        want = an.arc_projection(xs[ix], ys[iy], beta, sid, sdd)
        got = float((sgm[j] * u).sum() / sgm[j].sum())
        assert abs(got - want) < 0.25 * w, (j, got, want)


# --------------------------------------------------- cone forward projector
def _fp_samples(spec, step=0.2):
    """Sample positions of run_mgfpj.py:509-567 computed from the vector
    geometry: l_k = l_min + k * step * diag for k < floor((l_max - l_min) /
    (step * diag)), l_min/max = sid -/+ D/2, D the volume diagonal; returned
    as fractional voxel indexes (slice, row, column) plus the step length."""
    nx, nz, pix, vh = spec.img_dim, spec.img_dim_z, spec.pix, spec.vox_h
    D = np.sqrt(2 * (nx * pix) ** 2 + (nz * vh) ** 2)
    diag = np.sqrt(2 * pix * pix + vh * vh)
    count = int(np.floor(D / (step * diag)))
    S, E = an.rays(spec.angles, spec.u, spec.v, spec.sid, spec.sdd, spec.curved)
    e = (E - S) / np.linalg.norm(E - S, axis=-1, keepdims=True)
    ls = spec.sid - D / 2 + np.arange(count) * step * diag
    P = S[..., None, :] + e[..., None, :] * ls[:, None]
    x0 = -(nx - 1) / 2 * pix + spec.cx
    y0 = (nx - 1) / 2 * pix + spec.cy
    z0 = -(nz - 1) / 2 * vh + spec.cz
    fx = (P[..., 0] - x0) / pix
    fy = (P[..., 1] - y0) / (-pix)
    fz = (P[..., 2] - z0) / vh
    return fz, fy, fx, step * diag


def pin_cone_fp_is_trilinear_line_sum(R):
    """Cone forward projection of a smooth random volume (zero on its outer
    voxel shell) equals the left Riemann sum of scipy's trilinear
    interpolation (map_coordinates, order 1) at the sample positions: pins
    the z interpolation weights, the axis order and the step length."""
    spec = cone_case(nx=20, nz=10, nv=6, nu=30, nrows=14, pix=0.5, vh=0.7, w=0.9, h=1.4,
                     sid=60.0, sdd=120.0)
    spec.cz = 0.35
    rng = np.random.default_rng(21)
    vol = ndimage.gaussian_filter(rng.uniform(0, 1, (10, 20, 20)), 1.5)
    vol[[0, -1]] = 0
    vol[:, [0, -1]] = 0
    vol[:, :, [0, -1]] = 0
    fz, fy, fx, dl = _fp_samples(spec)
    vals = ndimage.map_coordinates(vol, [fz.ravel(), fy.ravel(), fx.ravel()], order=1,
                                   mode="constant", cval=0.0).reshape(fz.shape)
    want = vals.sum(axis=-1) * dl
    got = R.forward_project(vol, oracle_dict(spec), dtype=np.float64)
    assert want.max() > 1.0
    assert np.allclose(got, want, rtol=1e-9, atol=1e-9)


def pin_cone_fp_of_ball_is_chord(R):
    """Cone forward projection of a voxelised uniform ball against the
    analytic 3D chord length (within voxelisation and sampling error), for
    rays at cone angles up to 9 degrees."""
    nx, nz, pix, vh = 48, 40, 0.5, 0.5
    spec = cone_case(nx=nx, nz=nz, nv=4, nu=70, nrows=30, pix=pix, vh=vh, w=0.8, h=1.2,
                     sid=100.0, sdd=200.0)
    c, r, mu = (1.5, -2.0, 1.0), 7.0, 0.02
    vol = an.supersampled(nx, pix, lambda X, Y, Z: ((X - c[0]) ** 2 + (Y - c[1]) ** 2
                                                   + (Z - c[2]) ** 2) <= r * r, ss=3, nz=nz, vh=vh) * mu
    S, E = an.rays(spec.angles, spec.u, spec.v, spec.sid, spec.sdd)
    want = an.ball_integrals(S, E, c, r, mu)
    got = R.forward_project(vol, oracle_dict(spec), dtype=np.float64)
    deep = want > 0.6 * 2 * r * mu
    assert deep.sum() > 200
    err = np.abs(got[deep] - want[deep]) / want[deep]
    assert err.max() < 0.02, err.max()


# ------------------------------------------------------------------ noise
def pin_noise_statistics(R):
    """AddPossionNoise (run_mgfpj.py:643-654): the transmitted count
    N' = I0 exp(-out) has mean and variance N = I0 exp(-p) (Gaussian
    approximation of Poisson); counts that fall to <= 0 become 1e-6, so
    their output is log(I0 / 1e-6), with the probability P(g < -sqrt(N))."""
    I0, p = 1.0e4, 1.0
    out = R.add_noise(np.full((1, 500, 400), p, np.float32), I0, seed=5)
    n = I0 * np.exp(-out.astype(np.float64))
    N = I0 * np.exp(-p)
    k = n.size
    assert abs(n.mean() - N) < 4 * np.sqrt(N / k)
    assert abs(n.var() / N - 1) < 0.02
    I0, p = 1.0e5, np.log(1.0e5 / 0.25)  # N = 0.25: P(N' <= 0) = P(g < -0.5) = 0.3085
    out = R.add_noise(np.full((1, 400, 500), p, np.float32), I0, seed=9)
    clamp = np.float32(np.log(np.float32(I0) / np.float32(1e-6)))
    frac = float(np.mean(out == clamp))
    assert abs(frac - 0.3085) < 0.005, frac


# -------------------------------------------------------- IR / PICCS
def _ir_problem():
    spec = fan_case(nx=8, nv=10, nu=16, pix=1.0, w=1.6, sid=60.0, sdd=120.0, rows=1)
    g = oracle_dict(spec)
    rng = np.random.default_rng(31)
    sgm = rng.uniform(0.5, 1.5, (1, 10, 16))
    x0 = rng.uniform(0.0, 0.2, (1, 8, 8))
    return spec, g, sgm, x0


def _matrices(R, g, nz, n, nrows, nv, nu):
    """The forward projector F and the IR back-projector B as explicit
    matrices (columns = responses to unit vectors)."""
    F = np.stack([R.forward_project(e.reshape(nz, n, n), g, dtype=np.float64).ravel()
                  for e in np.eye(nz * n * n)], axis=1)
    B = np.stack([R.backproject_ir(e.reshape(nrows, nv, nu), g, dtype=np.float64).ravel()
                  for e in np.eye(nrows * nv * nu)], axis=1)
    return F, B


def _diff(n):
    D = sp.lil_matrix((n, n))
    for i in range(n - 1):
        D[i, i], D[i, i + 1] = -1, 1
    return D.tocsr()


def _tv_operator(shape, w):
    """Dt_W_D (…IRN_CGLS.py:267-286) as a matrix: per axis, D^T W D with the
    axis' first and last index receiving nothing from that axis; the z axis
    only when there are more than two slices."""
    nz, ny, nx = shape
    I = lambda k: sp.identity(k, format="csr")
    W = sp.diags(np.asarray(w).ravel())
    out = None
    axes = [(sp.kron(sp.kron(I(nz), _diff(ny)), I(nx)), 1), (sp.kron(sp.kron(I(nz), I(ny)), _diff(nx)), 2)]
    if nz > 2:
        axes.append((sp.kron(sp.kron(_diff(nz), I(ny)), I(nx)), 0))
    for D, ax in axes:
        keep = np.ones(shape, bool)
        idx = [slice(None)] * 3
        idx[ax] = 0
        keep[tuple(idx)] = False
        idx[ax] = -1
        keep[tuple(idx)] = False
        term = sp.diags(keep.ravel().astype(float)) @ (D.T @ W @ D)
        out = term if out is None else out + term
    return out.toarray()


def _tv_weight(x, beta):
    """GenerateWR (…IRN_CGLS.py:234-243) from np.diff."""
    ux = np.zeros_like(x)
    uy = np.zeros_like(x)
    uz = np.zeros_like(x)
    ux[:, :-1] = np.diff(x, axis=1)
    uy[:, :, :-1] = np.diff(x, axis=2)
    if x.shape[0] > 2:
        uz[:-1] = np.diff(x, axis=0)
    return 1 / np.sqrt(ux * ux + uy * uy + uz * uz + beta)


def _cg(A, b, x0, iters):
    """Textbook CG (the recursion of TikhonovSol, …IRN_CGLS.py:188-217)."""
    x = x0.copy()
    r = b - A @ x
    d = r.copy()
    for _ in range(iters):
        q = A @ d
        rr = r @ r
        a = rr / (d @ q)
        x = x + a * d
        r = r - a * q
        d = r + (r @ r) / rr * d
    return x


def pin_ir_solve_is_cg_on_the_normal_equations(R):
    """Mgfbp_ir (…IRN_CGLS.py:127-129, 157-217): CG on
    (B F + lambda * ratio * DtWD) x = B b with ratio = #sinogram elements /
    #image voxels, one Tikhonov pass (W = 1) then IRN passes with W from the
    last iterate; F, B and DtWD as explicit matrices."""
    spec, g, sgm, x0 = _ir_problem()
    F, B = _matrices(R, g, 1, 8, 1, 10, 16)
    ratio = (1 * 10 * 16) / (1 * 8 * 8)
    BF = B @ F
    L1 = _tv_operator(x0.shape, np.ones(x0.shape))
    lam = 0.5 * np.linalg.norm(BF, 2) / (ratio * np.linalg.norm(L1, 2))
    beta, iters = 1e-4, 5
    b = B @ sgm.ravel()
    x = _cg(BF + lam * ratio * L1, b, x0.ravel(), iters)
    W = _tv_weight(x.reshape(x0.shape), beta)
    x = _cg(BF + lam * ratio * _tv_operator(x0.shape, W), b, x, iters)
    got = R.ir_solve(sgm, g, x0, lam, beta, 1, iters)
    assert rel_l2(got.ravel(), x) < 1e-9


def pin_piccs_is_cg_on_its_normal_equations(R):
    """Mgfbp_ir_piccs (…piccs.py:83-150): start from the prior; CG on
    (B F + l(1-a) r DtW D + l a r DtW_p D) x = B b + l a r DtW_p D prior, with
    W from x and W_p from x - prior after the first pass."""
    spec, g, sgm, x0 = _ir_problem()
    prior = x0 + 0.05 * np.random.default_rng(32).standard_normal(x0.shape)
    F, B = _matrices(R, g, 1, 8, 1, 10, 16)
    ratio = 160 / 64
    BF = B @ F
    L1 = _tv_operator(x0.shape, np.ones(x0.shape))
    lam = 0.5 * np.linalg.norm(BF, 2) / (ratio * np.linalg.norm(L1, 2))
    alpha, beta, iters = 0.5, 1e-4, 5
    c1, c2 = lam * (1 - alpha) * ratio, lam * alpha * ratio
    x = prior.ravel()
    Bb = B @ sgm.ravel()
    W = Wp = np.ones(x0.shape)
    for p in range(2):
        if p:
            W = _tv_weight(x.reshape(x0.shape), beta)
            Wp = _tv_weight(x.reshape(x0.shape) - prior, beta)
        Lp = _tv_operator(x0.shape, Wp)
        A = BF + c1 * _tv_operator(x0.shape, W) + c2 * Lp
        x = _cg(A, Bb + c2 * Lp @ prior.ravel(), x, iters)
    got = R.ir_solve(sgm, g, x0, lam, beta, 1, iters, prior=prior, alpha=alpha)
    assert rel_l2(got.ravel(), x) < 1e-9
    # alpha = 0: PICCS is CGLS-TV started from the prior
    a0 = R.ir_solve(sgm, g, x0, lam, beta, 1, iters, prior=prior, alpha=0.0)
    cg = R.ir_solve(sgm, g, prior, lam, beta, 1, iters)
    assert rel_l2(a0, cg) < 1e-12


# -------------------------------------------------------- integration
def pin_fbp_of_fp_reproduces_phantom(R):
    """FBP(FP(phantom)) ~ phantom: the oracle's fan forward projection of a
    three-ellipse phantom, weighted, ramp-filtered and back-projected, comes
    back within the discretisation error inside the field of view."""
    nx, pix = 64, 0.5
    spec = fan_case(nx=nx, nv=360, nu=200, pix=pix, w=0.35, sid=150.0, sdd=300.0, rows=1)
    ells = [((0.0, 0.0), (12.0, 9.0), 0.3, 0.02), ((4.0, -3.0), (3.0, 5.0), 1.0, 0.01),
            ((-5.0, 4.0), (2.5, 2.5), 0.0, -0.008)]

    def inside(X, Y, e):
        (x0, y0), (a, b), t, _ = e
        xr = (X - x0) * np.cos(t) + (Y - y0) * np.sin(t)
        yr = -(X - x0) * np.sin(t) + (Y - y0) * np.cos(t)
        return (xr / a) ** 2 + (yr / b) ** 2 <= 1
    img = sum(an.supersampled(nx, pix, lambda X, Y, e=e: inside(X, Y, e)) * e[3] for e in ells)
    g = oracle_dict(spec)
    sgm = R.forward_project(img[None], g, dtype=np.float64)
    ws = R.weight_sinogram(sgm, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle, False,
                           dtype=np.float64)
    rec = R.backproject_fbp(R.filter_rows(ws, R.ramlak_taps(spec.n_u, 0.35), 0.35, np.float64), g,
                            np.float64)[0]
    xs, ys = an.pixel_centres(nx, pix)
    X, Y = np.meshgrid(xs, ys)
    fov = np.hypot(X, Y) < 14.0
    assert rel_l2(rec[fov], img[fov]) < 0.12
    smooth = ndimage.uniform_filter(rec, 5)[fov]
    assert rel_l2(smooth, ndimage.uniform_filter(img, 5)[fov]) < 0.03


# ---------------------------------------------------------------- tests
def test_cone_weight_is_ray_cosine():
    pin_cone_weight_is_ray_cosine(recon)


@pytest.mark.parametrize("curved", [False, True])
def test_fdk_of_z_invariant_cylinder(curved):
    pin_fdk_of_z_invariant_cylinder(recon, curved)


def test_curved_fbp_of_offcentre_disc():
    pin_curved_fbp_of_offcentre_disc(recon)


def test_curved_fp_blob_on_arc():
    pin_curved_fp_blob_on_arc(recon)


def test_cone_fp_is_trilinear_line_sum():
    pin_cone_fp_is_trilinear_line_sum(recon)


def test_cone_fp_of_ball_is_chord():
    pin_cone_fp_of_ball_is_chord(recon)


def test_noise_statistics():
    pin_noise_statistics(recon)


def test_ir_solve_is_cg_on_the_normal_equations():
    pin_ir_solve_is_cg_on_the_normal_equations(recon)


def test_piccs_is_cg_on_its_normal_equations():
    pin_piccs_is_cg_on_its_normal_equations(recon)


def test_fbp_of_fp_reproduces_phantom():
    pin_fbp_of_fp_reproduces_phantom(recon)


# ------------------------------------------------------ taps (product)
def test_curved_taps_closed_form():
    """Arc-detector ramp taps (run_mgfbp.py:892-896): odd n give
    -1/(pi^2 sdd^2 sin^2(n w / sdd)), the equiangular ramp (Kak & Slaney
    eq. 3.?? with gamma = n w / sdd), up to the fifth-order sine series the
    reference uses; even n give 0 and n = 0 gives 1/(4 w^2)."""
    n, w, sdd = 300, 0.4, 200.0
    h = geo.hamming_taps(n, w, 1.0, sdd, curved=True).astype(np.float64)
    k = np.arange(2 * n - 1) - (n - 1)
    odd = k % 2 != 0
    g = k[odd] * w / sdd
    want = -1 / (PI ** 2 * sdd ** 2 * np.sin(g) ** 2)
    assert np.allclose(h[odd], want, rtol=2e-6 + np.abs(g).max() ** 6 / 5040)
    assert np.all(h[(~odd) & (k != 0)] == 0)
    assert np.isclose(h[k == 0][0], 1 / (4 * w * w), rtol=1e-6)
    # small angles: the flat taps
    flat = geo.hamming_taps(n, w, 1.0, sdd).astype(np.float64)
    near = odd & (np.abs(k) < 5)
    assert np.allclose(h[near], flat[near], rtol=1e-4)


def test_dbt_taps_add_a_triangle_response():
    """DBTMode (run_mgfbp.py:903-912) adds k_t^2 sinc^2(pi n w k_t) with
    k_t = 0.02 / (2 w): its spectrum (times w) is the triangle
    k_t * max(0, 1 - |f| / k_t)."""
    n, w = 2048, 0.4
    base = geo.hamming_taps(n, w, 1.0, 1000.0).astype(np.float64)
    dbt = geo.hamming_taps(n, w, 1.0, 1000.0, dbt=True).astype(np.float64)
    add = dbt - base
    kt = 0.02 / (2 * w)
    f = np.linspace(0, 1.5, 31) * kt
    k = (np.arange(add.size) - (add.size - 1) // 2) * w
    resp = np.real(np.sum(add[None, :] * np.exp(-2j * np.pi * f[:, None] * k[None, :]), axis=1)) * w
    want = kt * np.clip(1 - f / kt, 0, None)
    assert np.allclose(resp, want, atol=0.01 * kt)


def pin_helical_travel_is_pitch_times_isocentre_coverage_per_lap(R):
    """Over one lap of views the helical z travel is pitch x the detector's
    height at the isocentre (n rows x row pitch x sid / sdd); two laps of
    views at the same spacing move the same per view."""
    for nv_lap, pitch, rows, dv, sid, sdd in [(360, 1.0, 16, 0.5, 500.0, 1000.0),
                                              (90, 0.7, 64, 0.25, 600.0, 1100.0)]:
        for laps in (1, 2):
            nv = nv_lap * laps
            ang = np.arange(nv, dtype=np.float64) * (2 * np.pi / nv_lap)
            v = (np.arange(rows) - (rows - 1) / 2) * dv
            zpv = float(R.helical_z_per_view(ang, v, rows, pitch, sdd, sid, dtype=np.float64))
            want = pitch * rows * dv * sid / sdd
            assert abs(zpv * nv_lap - want) <= 1e-9 * want


def pin_helical_fp_view_equals_circular_fp_of_the_shifted_volume(R):
    """View j of a helical cone forward projection (source and detector moved
    by j zpv in z) equals view j of the circular projection of the volume
    moved by -j zpv (the volume centre cz lowered by j zpv)."""
    spec = cone_case(nx=14, nz=10, nv=6, nu=20, nrows=12)
    spec.fpj_step = 0.5
    g = oracle_dict(spec)
    rng = np.random.default_rng(81)
    vol = rng.uniform(0, 1, (10, 14, 14))
    zpv = 0.37
    hel = R.forward_project(vol, dict(g, zpv=zpv), dtype=np.float64)
    for j in (0, 2, 5):
        gj = dict(g, angles=np.asarray(g["angles"])[j:j + 1], cz=g["cz"] - j * zpv)
        circ = R.forward_project(vol, gj, dtype=np.float64)[:, 0]
        assert np.allclose(hel[:, j], circ, rtol=1e-9, atol=1e-9 * np.abs(circ).max())
    assert np.abs(hel[:, 5] - R.forward_project(vol, g, dtype=np.float64)[:, 5]).max() > 1e-3


def test_helical_travel_is_pitch_times_isocentre_coverage_per_lap():
    pin_helical_travel_is_pitch_times_isocentre_coverage_per_lap(recon)


def test_helical_fp_view_equals_circular_fp_of_the_shifted_volume():
    pin_helical_fp_view_equals_circular_fp_of_the_shifted_volume(recon)
