"""Projection-matrix (P-matrix) geometry, host side.

A P-matrix P_j (3 x 4, 12 floats per view, row major) maps a point X of the
image frame to detector indexes: (u, v) = (P_j[0] X~, P_j[1] X~) / P_j[2] X~
with X~ = (x, y, z, 1).  Its inverse form is A_j = inv(P_j[:, :3]) (columns
e_u, e_v and detector-origin-minus-source) and the source x_s = -A_j P_j[:, 3].

Reference: Mgfbp.ChangePMatrix_SourceTrajectory, ChangePMatrix_PMatrixPixelSize
and ChangePMatrix_VerticalReconRange (run_mgfbp.py:721-874), called from
Mgfbp.__init__ (run_mgfbp.py:551-627), and Mgfpj.GenerateMatrixAFromPMatrix
(run_mgfpj_v2.py:218-227).  The normalisation is host algebra on 3 x 3
matrices; the element types follow the reference's NumPy flow (the P-matrix
array is float32, so matrix inverses run in float32 and the trajectory fits in
float64), because the normalised matrices are the geometry the device kernels
then use.

Reproduced as the reference computes it: the pixel-size step shifts the
detector origin by 0.5 * (e_v - e_v0) + 0.5 * (e_u - e_v0) (run_mgfbp.py:843),
i.e. by half of (e_u - e_v0) even when the P-matrix and scan pixel sizes are
equal.
"""
from __future__ import annotations

import math

import numpy as np

from .config import PI, XuLeastSquareSolution


def _views(pm, view_num: int) -> np.ndarray:
    """(view_num, 3, 4) float32 copy of a flat 12-per-view array."""
    return np.asarray(pm, dtype=np.float32).reshape(view_num, 3, 4).copy()


def _from_inverse(a_inv: np.ndarray, x_s: np.ndarray) -> np.ndarray:
    """P = [A^-1 | -A^-1 x_s] for one view (A^-1 float32, x_s (3, 1))."""
    return np.append(a_inv, -np.matmul(a_inv, x_s), axis=1)


def matrix_a_and_source(pm, view_num: int) -> tuple[np.ndarray, np.ndarray]:
    """GenerateMatrixAFromPMatrix (run_mgfpj_v2.py:218-227): per view
    A = inv(P[:, :3]) (9 floats, row major) and x_s = -A P[:, 3] (3 floats),
    float32 arrays of 9 * view_num and 3 * view_num."""
    P = _views(pm, view_num)
    A = np.empty((view_num, 3, 3), dtype=np.float32)
    xs = np.empty((view_num, 3), dtype=np.float32)
    for j in range(view_num):
        A[j] = np.linalg.inv(P[j, :, 0:3])
        xs[j] = -np.matmul(A[j], P[j, :, 3])
    return A.reshape(-1), xs.reshape(-1)


def _rotation_y(t):
    return np.array([[math.cos(t), 0, -math.sin(t)], [0, 1, 0], [math.sin(t), 0, math.cos(t)]])


def _rotation_x(t):
    return np.array([[1, 0, 0], [0, math.cos(t), -math.sin(t)], [0, math.sin(t), math.cos(t)]])


def _rotation_z(t):
    return np.array([[math.cos(t), math.sin(t), 0], [-math.sin(t), math.cos(t), 0], [0, 0, 1]])


def standard_frame(x_s_rec: np.ndarray):
    """Rigid motion that puts the source trajectory into standard form
    (run_mgfbp.py:733-760): fit the plane a x + b y + c z = 1 through the
    sources, rotate it onto z = const (about y, then x), fit the circle in that
    plane and shift its centre to the origin, then rotate about z so that the
    first source sits on +x.  x_s_rec: (3, V) float64.  Returns (R, x_s) with
    R the total 3 x 3 rotation and x_s the moved sources (3, V)."""
    V = x_s_rec.shape[1]
    plane = XuLeastSquareSolution(x_s_rec.transpose(), np.ones((V, 1)))[:, 0]
    ay = float(np.arctan(plane[0] / plane[2]))
    ax = float(np.arctan(plane[1] / np.sqrt(plane[0] ** 2 + plane[2] ** 2)) * np.sign(plane[2]))
    Ry, Rx = _rotation_y(ay), _rotation_x(ax)
    rot = np.matmul(Rx, np.matmul(Ry, x_s_rec))
    zc = np.mean(rot[2, :])
    xy = rot[0:2, :]
    B = np.concatenate((-xy.transpose(), np.ones((V, 1))), axis=1)
    circ = XuLeastSquareSolution(B, -np.sum(xy ** 2, axis=0))
    shifted = rot - np.array([[circ[0] / 2], [circ[1] / 2], [zc]])
    az = math.atan2(shifted[1, 0], shifted[0, 0])
    Rz = _rotation_z(az)
    return np.matmul(Rz, np.matmul(Rx, Ry)), np.matmul(Rz, shifted)


def source_trajectory(pm, view_num: int, modify: bool, pm_w: float, pm_h: float, n_u: int,
                      det_w: float, n_v: int, det_h: float) -> dict:
    """ChangePMatrix_SourceTrajectory (run_mgfbp.py:721-816).

    Per view: rescale P so that its pixel pitch matches (pm_w, pm_h), move
    (modify = True) or keep the frame, rebuild P from the moved e_u, e_v,
    detector origin and source, and derive the scan parameters.  Returns a
    dict: pmatrix (float32, 12 * V), positive_u_is_positive_y,
    positive_v_is_positive_z (+-1 from the first view), det_offset_horizontal,
    det_offset_vertical, source_isocenter_dis, source_det_dis (means over the
    views), the per-view arrays (sid_each, sdd_each, det_offset_horizontal_each,
    det_offset_vertical_each, scan_angle_each in degrees) and scan_sum (rad,
    the summed source angle steps)."""
    P = _views(pm, view_num)
    x_s_rec = np.zeros((3, view_num))
    for j in range(view_num):
        a = np.linalg.inv(P[j, :, 0:3])
        x_s_rec[:, j:j + 1] = -np.matmul(a, P[j, :, 3]).reshape(3, 1)
    R, x_s_fin = standard_frame(x_s_rec)
    if not modify:
        R, x_s_fin = np.eye(3), x_s_rec
    out = np.asarray(pm, dtype=np.float32).copy()
    u_c = np.zeros((view_num, 1))
    v_c = np.zeros((view_num, 1))
    x_d_c = np.zeros((3, view_num))
    scan_each = np.zeros((view_num, 1))
    scan_sum = 0.0
    pos_u = pos_v = None
    for j in range(view_num):
        Pj = P[j]
        a0 = np.linalg.inv(Pj[:, 0:3])
        ratio = (pm_w / np.linalg.norm(a0[:, 0]) + pm_h / np.linalg.norm(a0[:, 1])) * 0.5
        Pj = Pj / ratio
        a = np.linalg.inv(Pj[:, 0:3])
        x_s = x_s_fin[:, j:j + 1]
        if j <= view_num - 2:
            nxt = x_s_fin[:, j + 1:j + 2]
            step = math.atan2(nxt[1, 0], nxt[0, 0]) - math.atan2(x_s[1, 0], x_s[0, 0])
            if abs(step) > PI:
                step = step - 2 * PI * np.sign(step)
            scan_sum = scan_sum + step
            scan_each[j + 1, 0] = scan_sum / PI * 180.0
        e_v = np.matmul(R, a[:, 1])
        e_u = np.matmul(R, a[:, 0])
        if j == 0:
            pos_u = np.sign(e_u[1])
            if pos_u == 1:
                print("Attention: +u direction is along +y based on pmatrix! ")
            pos_v = np.sign(e_v[2])
            if pos_v == 1:
                print("Attention: +v direction is along +z based on pmatrix! ")
        x_do = np.matmul(R, a[:, 2])
        a[:, 0], a[:, 1], a[:, 2] = e_u, e_v, x_do
        Pn = _from_inverse(np.linalg.inv(a), x_s)
        u_c[j, 0] = Pn[0, 3] / Pn[2, 3]
        v_c[j, 0] = -np.dot(x_do, e_v) / (pm_h ** 2)
        x_d_c[:, j:j + 1] = x_do.reshape((3, 1)) + u_c[j, 0] * e_u.reshape((3, 1)) + v_c[j, 0] * e_v.reshape((3, 1))
        out[12 * j:12 * j + 12] = Pn.reshape(12)
    off_v_each = pos_v * ((n_v * det_h / pm_h - 1) / 2.0 - v_c) * pm_h
    off_u_each = -pos_u * ((n_u * det_w / pm_w - 1) / 2.0 - u_c) * pm_w
    sid_each = np.sqrt(np.sum(x_s_fin * x_s_fin, axis=0))
    sdd_each = np.sqrt(np.sum(x_d_c * x_d_c, axis=0))
    return dict(pmatrix=out, positive_u_is_positive_y=pos_u, positive_v_is_positive_z=pos_v,
                det_offset_horizontal=np.squeeze(np.mean(off_u_each, axis=0)).tolist(),
                det_offset_vertical=np.squeeze(np.mean(off_v_each, axis=0)).tolist(),
                source_isocenter_dis=np.squeeze(np.mean(sid_each, axis=0)).tolist(),
                source_det_dis=np.squeeze(np.mean(sdd_each, axis=0)).tolist(),
                sid_each=sid_each, sdd_each=sdd_each, det_offset_horizontal_each=off_u_each,
                det_offset_vertical_each=off_v_each, scan_angle_each=scan_each, scan_sum=scan_sum)


def pixel_size(pm, view_num: int, det_w: float, det_h: float, pm_w: float, pm_h: float) -> np.ndarray:
    """ChangePMatrix_PMatrixPixelSize (run_mgfbp.py:818-835): scale e_u and e_v
    from the P-matrix pitch to the scan pitch and shift the detector origin as
    the reference does.  Returns the new float32 array."""
    P = _views(pm, view_num)
    out = np.asarray(pm, dtype=np.float32).copy()
    for j in range(view_num):
        a = np.linalg.inv(P[j, :, 0:3])
        x_s = -np.matmul(a, P[j, :, 3]).reshape([3, 1])
        e_v0, e_u0 = a[:, 1].copy(), a[:, 0].copy()
        e_v = e_v0 * det_h / pm_h
        e_u = e_u0 * det_w / pm_w
        x_do = a[:, 2] + np.multiply(0.5, e_v - e_v0) + np.multiply(0.5, e_u - e_v0)
        a[:, 0], a[:, 1], a[:, 2] = e_u, e_v, x_do
        out[12 * j:12 * j + 12] = _from_inverse(np.linalg.inv(a), x_s).reshape(12)
    return out


def vertical_recon_range(pm, view_num: int, begin: int) -> np.ndarray:
    """ChangePMatrix_VerticalReconRange (run_mgfbp.py:837-852): move the
    detector origin to row `begin`, so v index 0 is the first reconstructed
    row.  Returns the new float32 array."""
    P = _views(pm, view_num)
    out = np.asarray(pm, dtype=np.float32).copy()
    for j in range(view_num):
        a = np.linalg.inv(P[j, :, 0:3])
        x_s = -np.matmul(a, P[j, :, 3]).reshape([3, 1])
        a[:, 2] = a[:, 2] + np.multiply(begin, a[:, 1])
        out[12 * j:12 * j + 12] = _from_inverse(np.linalg.inv(a), x_s).reshape(12)
    return out
