"""Seeded ellipse / ellipsoid phantoms and their analytic projections.

Synthetic inputs for benchmarks and the serving stream (the reference ships
no data).  Phantom parameters are drawn with numpy.random.default_rng(seed):
centres uniform in a disc of radius 0.5 * r_fov, semi-axes U[5, 60] mm
in-plane (U[5, 30] mm in z), rotation about z U[0, pi), attenuation
U[0, 0.02] mm^-1, additive.  Analytic projections use the scanner geometry
of the reference's forward projector (run_mgfpj.py:474-551): the object
rotates by the view angle, the source sits at (sid, 0, 0) and detector
element (u, v) at (-(sdd - sid), u, v).
"""
from __future__ import annotations

import numpy as np
import torch


def random_ellipsoids(n: int, seed: int, r_fov: float, z_semi=(5.0, 30.0), z_half=0.0):
    """(n, 8) float64 rows: cx, cy, cz, a, b, c, phi, mu."""
    rng = np.random.default_rng(seed)
    r = 0.5 * r_fov * np.sqrt(rng.uniform(0, 1, n))
    t = rng.uniform(0, 2 * np.pi, n)
    cx, cy = r * np.cos(t), r * np.sin(t)
    cz = rng.uniform(-z_half, z_half, n) if z_half > 0 else np.zeros(n)
    a = rng.uniform(5, 60, n)
    b = rng.uniform(5, 60, n)
    c = rng.uniform(z_semi[0], z_semi[1], n) if z_half > 0 else np.full(n, 1e9)
    phi = rng.uniform(0, np.pi, n)
    mu = rng.uniform(0, 0.02, n)
    return np.stack([cx, cy, cz, a, b, c, phi, mu], axis=1)


def voxelize(ell: np.ndarray, nx: int, nz: int, pix: float, vh: float, clip_r: float | None = None,
             supersample: int = 2, device="cuda", cz0: float = 0.0) -> torch.Tensor:
    """Point-sampled phantom on the reference's grid ([z][y][x]; x = p(i-(N-1)/2),
    y = -p(j-(N-1)/2), z = h(k-(Nz-1)/2) + cz0) with supersample^2 in-plane
    sub-samples; float32."""
    dev = torch.device(device)
    e = torch.tensor(ell, dtype=torch.float64, device=dev)
    ss = supersample
    off = (torch.arange(ss, dtype=torch.float64, device=dev) + 0.5) / ss - 0.5
    i = torch.arange(nx, dtype=torch.float64, device=dev)
    xs = (pix * ((i[:, None] + off[None, :]) - (nx - 1) / 2.0)).reshape(-1)
    ys = (-pix * ((i[:, None] + off[None, :]) - (nx - 1) / 2.0)).reshape(-1)
    out = torch.zeros((nz, nx, nx), dtype=torch.float32, device=dev)
    for k in range(nz):
        z = vh * (k - (nz - 1) / 2.0) + cz0 if nz > 1 else cz0
        Y, X = torch.meshgrid(ys, xs, indexing="ij")
        acc = torch.zeros_like(X)
        for row in e:
            cx, cy, cz, a, b, c, phi, mu = row.tolist()
            dz = (z - cz) / c
            if abs(dz) >= 1:
                continue
            cp, sp = np.cos(phi), np.sin(phi)
            xr = (X - cx) * cp + (Y - cy) * sp
            yr = -(X - cx) * sp + (Y - cy) * cp
            inside = (xr / a) ** 2 + (yr / b) ** 2 + dz * dz <= 1
            acc = acc + inside.to(acc.dtype) * mu
        if clip_r is not None:
            acc = acc * ((X * X + Y * Y) <= clip_r * clip_r).to(acc.dtype)
        out[k] = acc.reshape(nx, ss, nx, ss).mean(dim=(1, 3)).to(torch.float32)
    return out


def project(ell: np.ndarray, angles: np.ndarray, u: np.ndarray, v: np.ndarray, sid: float,
            sdd: float, device="cuda", chunk_views: int = 64) -> torch.Tensor:
    """Analytic line integrals (rows, views, n_u) float32 of the ellipsoids
    (cz, c as given; c = 1e9 gives a cylinder, i.e. a 2-D ellipse)."""
    dev = torch.device(device)
    ang = torch.tensor(np.asarray(angles, dtype=np.float64), device=dev)
    U = torch.tensor(np.asarray(u, dtype=np.float64), device=dev)
    V = torch.tensor(np.asarray(v, dtype=np.float64), device=dev)
    nr, nv, nu = V.shape[0], ang.shape[0], U.shape[0]
    out = torch.empty((nr, nv, nu), dtype=torch.float32, device=dev)
    # ray in the rotating object frame: source S, direction D (unit)
    ex = torch.full((nr, 1, nu), -(sdd - sid) - sid, dtype=torch.float64, device=dev)
    ey = U[None, None, :].expand(nr, 1, nu)
    ez = V[:, None, None].expand(nr, 1, nu)
    dist = torch.sqrt(ex * ex + ey * ey + ez * ez)
    dx0, dy0, dz0 = ex / dist, ey / dist, ez / dist
    for j0 in range(0, nv, chunk_views):
        a = ang[j0:j0 + chunk_views][None, :, None]
        ca, sa = torch.cos(a), torch.sin(a)
        sx, sy = sid * ca, sid * sa  # rotated source (sid, 0) by +angle
        dx = dx0 * ca - dy0 * sa
        dy = dy0 * ca + dx0 * sa
        dz = dz0.expand_as(dx)
        acc = torch.zeros(dx.shape, dtype=torch.float64, device=dev)
        for row in ell:
            cx, cy, cz, ea, eb, ec, phi, mu = [float(t) for t in row]
            cp, sp = np.cos(phi), np.sin(phi)
            px, py, pz = sx - cx, sy - cy, -cz
            qx = (px * cp + py * sp) / ea
            qy = (-px * sp + py * cp) / eb
            qz = torch.full_like(qx, pz / ec)
            rx = (dx * cp + dy * sp) / ea
            ry = (-dx * sp + dy * cp) / eb
            rz = dz / ec
            A = rx * rx + ry * ry + rz * rz
            B = 2 * (qx * rx + qy * ry + qz * rz)
            Cc = qx * qx + qy * qy + qz * qz - 1
            disc = B * B - 4 * A * Cc
            chord = torch.sqrt(torch.clamp(disc, min=0)) / A
            acc = acc + torch.where(disc > 0, chord, torch.zeros_like(chord)) * mu
        out[:, j0:j0 + chunk_views, :] = acc.to(torch.float32)
    return out
