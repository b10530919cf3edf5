"""Small seeded geometries shared by the tests (scaled-down BASELINE cases)."""
from __future__ import annotations

import numpy as np

from reconstruction_program_taichi_amd import geometry as geo

PI = geo.PI


def fan_case(nx=64, nv=90, nu=100, pix=0.6, w=0.8, sid=250.0, sdd=500.0, rows=2, scan_deg=360.0,
             rot_deg=0.0, off=0.0):
    scan = scan_deg / 180.0 * PI
    rot = rot_deg / 180.0 * PI
    ang = geo.angle_array(nv, scan, rot)
    u = geo.det_pos_array(nu, nu, -w, -off, 0)
    v = geo.det_pos_array(rows, rows, 0, 0, 0)
    short = abs(scan % (2 * PI)) >= (0.01 / 180 * PI)
    spec = geo.GeometrySpec(img_dim=nx, img_dim_z=rows, pix=pix, vox_h=0.0, n_views=nv, sid=sid,
                            sdd=sdd, total_scan_angle=scan, angles=ang, u=u, v=v, img_rot=rot,
                            short_scan=short, cone=False, n_v_full=rows)
    return spec


def cone_case(nx=48, nz=24, nv=72, nu=96, nrows=40, pix=0.5, vh=0.5, w=0.6, h=0.6, sid=300.0,
              sdd=500.0, top_row=False, off_v=0.0, cz=None):
    scan = 2 * PI
    ang = geo.angle_array(nv, scan, 0.0)
    u = geo.det_pos_array(nu, nu, -w, 0.0, 0)
    v = geo.det_pos_array(nrows, nrows, (-1 if top_row else 1) * h, off_v, 0)
    if cz is None:
        cz = off_v * sid / sdd
    spec = geo.GeometrySpec(img_dim=nx, img_dim_z=nz, pix=pix, vox_h=vh, n_views=nv, sid=sid,
                            sdd=sdd, total_scan_angle=scan, angles=ang, u=u, v=v, cz=cz,
                            short_scan=False, cone=True, n_v_full=nrows, det_h=h)
    return spec


def pmatrix_for(spec, rot=None):
    """FP64 P-matrices (12 per view) of spec's circular flat-detector
    geometry, built from first principles: view j has its source at
    S = sid (cos b, sin b, 0) with b = angles[j] - rot, the detector normal
    n = -(cos b, sin b, 0) at distance sdd, columns along e = (-sin b, cos b, 0)
    with u[i] = u[0] + i du and rows along z with v[k] = v[0] + k dv.  Then
    A = [du e, dv e_z, sdd n + u[0] e + v[0] e_z] and P = [A^-1 | -A^-1 S]
    maps a point of the back-projector's frame to (u index, v index)."""
    rot = spec.img_rot if rot is None else rot
    ang = np.asarray(spec.angles, np.float64)
    u = np.asarray(spec.u, np.float64)
    v = np.asarray(spec.v, np.float64)
    du = u[1] - u[0]
    dv = (v[1] - v[0]) if v.shape[0] > 1 else 1.0
    out = []
    for b in ang - rot:
        c, s = np.cos(b), np.sin(b)
        S = np.array([spec.sid * c, spec.sid * s, 0.0])
        n = -np.array([c, s, 0.0])
        e = np.array([-s, c, 0.0])
        ez = np.array([0.0, 0.0, 1.0])
        A = np.column_stack([du * e, dv * ez, spec.sdd * n + u[0] * e + v[0] * ez])
        Ai = np.linalg.inv(A)
        out.append(np.hstack([Ai, -Ai @ S[:, None]]).reshape(12))
    return np.concatenate(out)


def matrix_a_for(pm, nv):
    """A_j = inv(P_j[:, :3]) and x_s_j = -A_j P_j[:, 3] in FP64 (the inverse
    form the forward projector's P-matrix branch reads)."""
    P = np.asarray(pm, np.float64).reshape(nv, 3, 4)
    A = np.linalg.inv(P[:, :, :3])
    xs = -np.einsum("vij,vj->vi", A, P[:, :, 3])
    return A.reshape(-1), xs.reshape(-1)


def oracle_dict(spec):
    return dict(img_dim=spec.img_dim, img_dim_z=spec.img_dim_z, pix=spec.pix, vox_h=spec.vox_h,
                sid=spec.sid, sdd=spec.sdd, angles=spec.angles, u=spec.u, v=spec.v,
                det_h=spec.det_h, cone=spec.cone, curved=spec.curved, cx=spec.cx, cy=spec.cy,
                cz=spec.cz, img_rot=spec.img_rot, total_scan_angle=spec.total_scan_angle,
                short_scan=spec.short_scan, fpj_step=spec.fpj_step, center_z=spec.center_z)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
