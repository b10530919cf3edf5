"""Independent ray geometry and analytic line integrals for the oracle pins.

Nothing here calls the oracle.  The conventions are the reference's, stated
once:

* gantry frame: the source sits at (sid, 0, zs); a flat detector element at
  (sid - sdd, u, v); an arc (curved) element at arc length u on the circle of
  radius sdd around the source, (sid - sdd cos(u/sdd), sdd sin(u/sdd), v)
  (run_mgfpj.py:531-539);
* view j turns the gantry frame into the object frame by +beta_j:
  (x, y) -> (x cos b - y sin b, y cos b + x sin b) (run_mgfpj.py:563-567), so
  an object point seen from view j sits at R(-b)(x, y) in the gantry frame
  (run_mgfbp.py:1122-1131).
"""
from __future__ import annotations

import numpy as np


def element_gantry(u, v, sid, sdd, curved):
    """Gantry-frame positions of detector elements, broadcast over u and v."""
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    if curved:
        gx = sid - sdd * np.cos(u / sdd)
        gy = sdd * np.sin(u / sdd)
    else:
        gx = np.full(np.broadcast(u, v).shape, sid - sdd, np.float64)
        gy = u + 0 * v
    return np.broadcast_arrays(gx, gy, v + 0 * u)


def rays(angles, u, v, sid, sdd, curved=False):
    """Object-frame source S and element E of every ray, shape (rows, views,
    cols, 3): rows follow v, cols follow u."""
    ang = np.asarray(angles, np.float64)
    u = np.asarray(u, np.float64)
    v = np.asarray(v, np.float64)
    gx, gy, gz = element_gantry(u[None, :], v[:, None], sid, sdd, curved)  # (rows, cols)
    c, s = np.cos(ang)[None, :, None], np.sin(ang)[None, :, None]
    ex = gx[:, None, :] * c - gy[:, None, :] * s
    ey = gy[:, None, :] * c + gx[:, None, :] * s
    ez = np.broadcast_to(gz[:, None, :], ex.shape)
    E = np.stack([ex, ey, ez], axis=-1)
    S = np.stack(np.broadcast_arrays(sid * c, sid * s, 0 * c), axis=-1)
    S = np.broadcast_to(S, E.shape)
    return S, E


def cylinder_integrals(S, E, cx, cy, r, mu):
    """mu times the length of each line S->E inside the infinite z-parallel
    cylinder of radius r around (cx, cy): the in-plane chord scaled by the
    ray's 3D / in-plane length ratio."""
    e = E - S
    exy = np.hypot(e[..., 0], e[..., 1])
    d = np.abs(e[..., 0] * (cy - S[..., 1]) - e[..., 1] * (cx - S[..., 0])) / exy
    chord_xy = 2 * np.sqrt(np.clip(r * r - d * d, 0, None))
    return mu * chord_xy * np.linalg.norm(e, axis=-1) / exy


def ball_integrals(S, E, c, r, mu):
    """mu times the chord of each line S->E through the ball (c, r)."""
    e = E - S
    e = e / np.linalg.norm(e, axis=-1, keepdims=True)
    w = np.asarray(c, np.float64) - S
    d2 = np.sum(w * w, axis=-1) - np.sum(w * e, axis=-1) ** 2
    return mu * 2 * np.sqrt(np.clip(r * r - d2, 0, None))


def arc_projection(x, y, beta, sid, sdd):
    """Arc-length coordinate on a curved detector of the ray through object
    point (x, y) at view angle beta: sdd times the ray's angle from the
    central ray."""
    xg = x * np.cos(beta) + y * np.sin(beta)
    yg = -x * np.sin(beta) + y * np.cos(beta)
    return sdd * np.arctan2(yg, sid - xg)


def pixel_centres(nx, pix, cx=0.0, cy=0.0):
    """Object coordinates of the pixel centres along x (columns) and y (rows):
    x grows with the column index, y falls with the row index
    (run_mgfbp.py:1086-1093)."""
    h = (nx - 1) / 2.0
    i = np.arange(nx)
    return pix * (i - h) + cx, -pix * (i - h) + cy


def supersampled(nx, pix, inside, ss=4, nz=1, vh=0.0):
    """Fraction of each pixel (and, with nz > 1, each voxel of height vh)
    that the predicate inside(x, y[, z]) covers, from ss points per axis."""
    off = (np.arange(ss) + 0.5) / ss - 0.5
    xs, ys = pixel_centres(nx, pix)
    if nz == 1 and vh == 0.0:
        acc = np.zeros((nx, nx))
        for a in off:
            for b in off:
                X, Y = np.meshgrid(xs + a * pix, ys - b * pix)
                acc += inside(X, Y)
        return acc / ss ** 2
    zs = (np.arange(nz) - (nz - 1) / 2.0) * vh
    acc = np.zeros((nz, nx, nx))
    for a in off:
        for b in off:
            for c in off:
                Z, Y, X = np.meshgrid(zs + c * vh, ys - b * pix, xs + a * pix, indexing="ij")
                acc += inside(X, Y, Z)
    return acc / ss ** 3
