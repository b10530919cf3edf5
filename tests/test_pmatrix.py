"""P-matrix geometry on the host (reconstruction_program_taichi_amd/pmatrix.py
and the PMatrixFile path of Mgfbp), pinned by first-principles P-matrices of
a known circular scan: the normalisation must recover the scan's distances,
offsets and angles, undo a rigid motion of the whole scanner when asked to,
and move the detector origin exactly as the reference's steps do
(run_mgfbp.py:721-852)."""
import json
import math

import numpy as np
import pytest

from reconstruction_program_taichi_amd import geometry as geo
from reconstruction_program_taichi_amd import pmatrix as pmx
from tests._cases import PI, pmatrix_for

SID, SDD, NU, NV, W, H, OFF_U, OFF_V, VIEWS = 500.0, 1000.0, 100, 40, 0.8, 0.8, 1.2, -0.8, 90


def _scan(scan_deg=360.0, views=VIEWS):
    """Circular cone-beam scan in the reference's FBP conventions: u runs
    along -y (u_i = -(i - (N-1)/2) w - off_u), v along +z."""
    scan = scan_deg / 180.0 * PI
    ang = geo.angle_array(views, scan, 0.0)
    u = geo.det_pos_array(NU, NU, -W, -OFF_U, 0)
    v = geo.det_pos_array(NV, NV, H, OFF_V, 0)
    return geo.GeometrySpec(img_dim=16, img_dim_z=8, pix=1.0, vox_h=1.0, n_views=views, sid=SID,
                            sdd=SDD, total_scan_angle=scan, angles=ang, u=u, v=v, cone=True,
                            n_v_full=NV, det_h=H)


def _project(pm, j, X):
    P = np.asarray(pm, np.float64)[12 * j:12 * j + 12].reshape(3, 4)
    q = P @ np.append(X, 1.0)
    return q[:2] / q[2]


def test_matrix_a_and_source_invert_the_projection():
    spec = _scan()
    pm = pmatrix_for(spec).astype(np.float32)
    A, xs = pmx.matrix_a_and_source(pm, VIEWS)
    assert A.dtype == np.float32 and A.shape == (9 * VIEWS,) and xs.shape == (3 * VIEWS,)
    for j in (0, 17, 63):
        Aj, sj = A[9 * j:9 * j + 9].reshape(3, 3).astype(np.float64), xs[3 * j:3 * j + 3].astype(np.float64)
        b = float(spec.angles[j])
        assert np.allclose(sj, [SID * math.cos(b), SID * math.sin(b), 0.0], atol=2e-2)
        for uv in ((3.0, 5.0), (71.5, 20.25)):
            X = sj + 0.7 * (Aj @ np.array([uv[0], uv[1], 1.0]))
            assert np.allclose(_project(pm, j, X), uv, atol=2e-3)


def test_source_trajectory_recovers_the_scan():
    spec = _scan()
    pm = pmatrix_for(spec).astype(np.float32)
    r = pmx.source_trajectory(pm, VIEWS, False, W, H, NU, W, NV, H)
    assert r["positive_u_is_positive_y"] == -1 and r["positive_v_is_positive_z"] == 1
    assert r["source_isocenter_dis"] == pytest.approx(SID, rel=1e-4)
    assert r["source_det_dis"] == pytest.approx(SDD, rel=1e-4)
    assert r["det_offset_horizontal"] == pytest.approx(OFF_U, abs=1e-2)
    assert r["det_offset_vertical"] == pytest.approx(OFF_V, abs=1e-2)
    step = 360.0 / VIEWS
    assert np.allclose(r["scan_angle_each"][:, 0], np.arange(VIEWS) * step, atol=1e-3)
    assert r["scan_sum"] == pytest.approx((VIEWS - 1) * step / 180 * PI, rel=1e-5)
    # already in standard form with the scan's pitch: the matrices keep
    # mapping every point where the original ones do
    X = np.array([12.0, -7.0, 3.0])
    for j in (0, 44):
        assert np.allclose(_project(r["pmatrix"], j, X), _project(pm, j, X), atol=2e-3)


def test_source_trajectory_standard_form_undoes_a_rigid_motion():
    """ModifyPMatrixToStandardForm: the P-matrices of the same scan seen
    through a tilted, shifted world frame normalise back to the original."""
    spec = _scan()
    pm = pmatrix_for(spec)
    ax, ay, az = 0.05, -0.03, 0.7
    Rx = np.array([[1, 0, 0], [0, math.cos(ax), -math.sin(ax)], [0, math.sin(ax), math.cos(ax)]])
    Ry = np.array([[math.cos(ay), 0, math.sin(ay)], [0, 1, 0], [-math.sin(ay), 0, math.cos(ay)]])
    Rz = np.array([[math.cos(az), -math.sin(az), 0], [math.sin(az), math.cos(az), 0], [0, 0, 1]])
    R0 = Rz @ Ry @ Rx
    t = np.array([5.0, -3.0, 40.0])
    moved = []
    for j in range(VIEWS):
        P = pm[12 * j:12 * j + 12].reshape(3, 4)
        M = np.hstack([P[:, :3] @ R0.T, (P[:, 3] - P[:, :3] @ R0.T @ t)[:, None]])
        moved.append(M.reshape(12))
    moved = np.concatenate(moved).astype(np.float32)
    X = np.array([10.0, 6.0, -4.0])
    assert not np.allclose(_project(moved, 5, X), _project(pm, 5, X), atol=1.0)
    r = pmx.source_trajectory(moved, VIEWS, True, W, H, NU, W, NV, H)
    for j in (0, 5, 60):
        assert np.allclose(_project(r["pmatrix"], j, X), _project(pm, j, X), atol=2e-2)
    assert r["source_isocenter_dis"] == pytest.approx(SID, rel=1e-4)
    assert r["det_offset_horizontal"] == pytest.approx(OFF_U, abs=2e-2)
    # without the modification the frame stays moved
    keep = pmx.source_trajectory(moved, VIEWS, False, W, H, NU, W, NV, H)
    assert np.allclose(_project(keep["pmatrix"], 5, R0 @ X + t), _project(pm, 5, X), atol=2e-2)


def test_short_scan_angle_from_the_trajectory():
    spec = _scan(scan_deg=200.0, views=100)
    pm = pmatrix_for(spec).astype(np.float32)
    r = pmx.source_trajectory(pm, 100, False, W, H, NU, W, NV, H)
    assert r["scan_sum"] / 99 * 100 == pytest.approx(200.0 / 180 * PI, rel=1e-5)


@pytest.mark.parametrize("pm_w, shift_u", [(W, None), (W / 2, None)])
def test_pixel_size_step_moves_the_detector_origin(pm_w, shift_u):
    """ChangePMatrix_PMatrixPixelSize rescales e_u by det_w / pm_w and moves
    the origin by 0.5 (e_v - e_v0) + 0.5 (e_u - e_v0) (the reference's
    expression).  With the scan's e_u = k e_u0 (k = det_w / pm_w) and e_v
    unchanged, a point at P-matrix index (u, v) lands at scan index
    ((u - 0.5 k) / k, v + 0.5)."""
    spec = _scan()
    pm = pmatrix_for(spec).astype(np.float32)
    out = pmx.pixel_size(pm, VIEWS, W, H, pm_w, H)
    k = W / pm_w
    X = np.array([8.0, -11.0, 2.5])
    for j in (0, 33):
        u, v = _project(pm, j, X)
        un, vn = _project(out, j, X)
        assert un == pytest.approx((u - 0.5 * k) / k, abs=3e-3)
        assert vn == pytest.approx(v + 0.5, abs=3e-3)


def test_vertical_recon_range_shifts_rows():
    spec = _scan()
    pm = pmatrix_for(spec).astype(np.float32)
    out = pmx.vertical_recon_range(pm, VIEWS, 7)
    X = np.array([-3.0, 4.0, 6.0])
    for j in (2, 80):
        u, v = _project(pm, j, X)
        un, vn = _project(out, j, X)
        assert un == pytest.approx(u, abs=2e-3) and vn == pytest.approx(v - 7, abs=2e-3)


def _write_config(tmp_path, pm, extra=None):
    (tmp_path / "in").mkdir(exist_ok=True)
    (tmp_path / "pm.jsonc").write_text(json.dumps({"Value": [float(x) for x in pm]}))
    cfg = {"InputDir": str(tmp_path / "in"), "OutputDir": str(tmp_path / "out"), "InputFiles": "sgm_.*",
           "OutputFilePrefix": "", "OutputFileReplace": ["sgm_", "rec_"],
           "SinogramWidth": NU, "DetectorElementWidth": W, "DetectorOffcenter": 0.0,
           "SliceCount": NV, "Views": VIEWS, "TotalScanAngle": 360.0,
           "SourceIsocenterDistance": 123.0, "SourceDetectorDistance": 456.0,
           "ImageDimension": 16, "PixelSize": 1.0, "HammingFilter": 1.0, "ConeBeam": True,
           "DetectorElementHeight": H, "ImageSliceCount": 8, "VoxelHeight": 1.0,
           "PMatrixFile": str(tmp_path / "pm.jsonc")}
    cfg.update(extra or {})
    return cfg


def test_mgfbp_pmatrix_file_sets_the_geometry(tmp_path, capsys):
    from reconstruction_program_taichi_amd.fbp import Mgfbp

    spec = _scan()
    pm = pmatrix_for(spec).astype(np.float32)
    cfg = _write_config(tmp_path, pm, {"SaveModifiedConfigFolder": str(tmp_path / "saved")})
    fbp = Mgfbp(cfg)
    assert "--PMatrix applied" in capsys.readouterr().out
    assert fbp.bool_apply_pmatrix == 1
    # the config dict carries the derived parameters, as the reference's does
    assert cfg["SourceIsocenterDistance"] == pytest.approx(SID, rel=1e-4)
    assert cfg["SourceDetectorDistance"] == pytest.approx(SDD, rel=1e-4)
    assert cfg["DetectorOffsetHorizontal"] == pytest.approx(OFF_U, abs=1e-2)
    assert cfg["DetectorOffsetVertical"] == pytest.approx(OFF_V, abs=1e-2)
    g = fbp.geometry_spec()
    assert g.pmatrix is not None and g.pmatrix.shape == (12 * VIEWS,)
    assert g.sid == pytest.approx(SID, rel=1e-4)
    # the stored matrices are the normalised ones (pixel-size step included)
    X = np.array([1.0, 2.0, 3.0])
    u, v = _project(pm, 3, X)
    assert np.allclose(_project(g.pmatrix, 3, X), (u - 0.5, v + 0.5), atol=3e-3)
    for name in ("config_mgfbp", "pmatrix_file", "sid_file", "sdd_file", "det_offset_horizontal_file",
                 "det_offset_vertical_file", "scan_angle_file"):
        assert (tmp_path / "saved" / f"{name}.jsonc").exists(), name
    saved = json.loads((tmp_path / "saved" / "pmatrix_file.jsonc").read_text())["Value"]
    assert np.allclose(saved, g.pmatrix)


def test_mgfbp_pmatrix_file_errors(tmp_path):
    from reconstruction_program_taichi_amd.fbp import Mgfbp

    pm = pmatrix_for(_scan()).astype(np.float32)
    with pytest.raises(SystemExit):
        Mgfbp(_write_config(tmp_path, pm[:-12]))
    (tmp_path / "bad.jsonc").write_text(json.dumps({"Values": [1.0]}))
    with pytest.raises(SystemExit):
        Mgfbp(_write_config(tmp_path, pm, {"PMatrixFile": str(tmp_path / "bad.jsonc")}))
    with pytest.raises(SystemExit):
        Mgfbp(_write_config(tmp_path, pm, {"ModifyPMatrixToStandardForm": 1}))
    with pytest.raises(SystemExit):
        Mgfbp(_write_config(tmp_path, pm, {"PMatrixDetectorElementWidth": -1.0}))
