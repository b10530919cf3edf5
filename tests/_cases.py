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
