"""run_mgfpj_v2 / run_mgfpj_v3 drop-ins: configuration and the geometry
their forward projections use (host only; GPU parity is in
test_gpu_fpj_variants.py).  Expected arrays are written out from the
reference's formulas: Mgfbp.InitializeArrays (run_mgfbp.py:1244-1259),
GenerateDectPixPosArrayFPJ (run_mgfpj_v2.py:229-233) and the v3 call order
(run_mgfpj_v3.py:82-90)."""
import json

import numpy as np
import pytest

from reconstruction_program_taichi_amd import geometry as geo

N, NV, VIEWS, W, H = 40, 12, 30, 0.6, 0.5


def _cfg(tmp_path, **kw):
    (tmp_path / "in").mkdir(exist_ok=True)
    c = {"InputDir": str(tmp_path / "in"), "OutputDir": str(tmp_path / "out"), "InputFiles": "img_.*",
         "OutputFilePrefix": "", "OutputFileReplace": ["img_", "sgm_"], "SinogramWidth": N,
         "DetectorElementWidth": W, "DetectorOffcenter": 0.3, "SliceCount": NV, "Views": VIEWS,
         "SourceIsocenterDistance": 300.0, "SourceDetectorDistance": 500.0, "ImageDimension": 24,
         "PixelSize": 0.5, "HammingFilter": 1.0, "ConeBeam": True, "DetectorElementHeight": H,
         "ImageSliceCount": 10, "VoxelHeight": 0.5, "SliceOffCenter": 0.2,
         "DetectorElementVerticalReconRange": [2, 9], "OversampleSize": 2}
    c.update(kw)
    return c


def _fbp_u(off=0.3):
    return geo.det_pos_array(N, N, -W, -off, 0)


def test_v2_geometry(tmp_path):
    from reconstruction_program_taichi_amd.fpj_v2 import Mgfpj

    f = Mgfpj(_cfg(tmp_path, StartAngle=30))
    s = f.fp_geometry_spec()
    assert s.n_u == 2 * N and np.array_equal(s.u[:N], _fbp_u()) and not s.u[N:].any()
    assert np.array_equal(s.v, geo.det_pos_array(NV, 8, H, 0.2, 2))
    # StartAngle enters the angle array as given (no degree conversion)
    assert np.array_equal(s.angles, geo.angle_array(VIEWS, 2 * geo.PI, 30))
    assert s.fpj_count_mode == 0 and s.center_z is None and s.matrix_a is None
    assert f.img_sgm.shape == (NV, VIEWS, N)


def test_v3_geometry(tmp_path):
    from reconstruction_program_taichi_amd.fpj_v3 import Mgfpj_v3

    cz = np.linspace(-1, 1, VIEWS).astype(np.float32)
    (tmp_path / "cz.jsonc").write_text(json.dumps({"Value": cz.tolist()}))
    (tmp_path / "ang.jsonc").write_text(json.dumps({"Value": list(range(VIEWS))}))
    f = Mgfpj_v3(_cfg(tmp_path, ImageCenterZFile=str(tmp_path / "cz.jsonc"),
                      ScanAngleFile=str(tmp_path / "ang.jsonc")))
    s = f.fp_geometry_spec()
    fpj_u = geo.det_pos_array(2 * N, 2 * N, -W / 2, -0.3, 0)
    assert np.array_equal(s.u[:N], _fbp_u()) and np.array_equal(s.u[N:], fpj_u[N:])
    assert np.array_equal(s.v, geo.det_pos_array(NV, 8, H, 0.2, 2))
    assert np.array_equal(s.center_z, cz)
    assert s.fpj_count_mode == 1
    # the uneven angles are read, then InitializeArrays regenerates uniform ones
    assert f.bool_uneven_scan_angle == 1
    assert np.array_equal(s.angles, geo.angle_array(VIEWS, 2 * geo.PI, 0.0))


def test_v3_file_errors(tmp_path):
    from reconstruction_program_taichi_amd.fpj_v3 import Mgfpj_v3

    (tmp_path / "short.jsonc").write_text(json.dumps({"Value": [0.0, 1.0]}))
    with pytest.raises(SystemExit):
        Mgfpj_v3(_cfg(tmp_path, ImageCenterZFile=str(tmp_path / "short.jsonc")))
    (tmp_path / "noval.jsonc").write_text(json.dumps({"V": [0.0]}))
    with pytest.raises(SystemExit):
        Mgfpj_v3(_cfg(tmp_path, ScanAngleFile=str(tmp_path / "noval.jsonc")))
    with pytest.raises(SystemExit):
        Mgfpj_v3(_cfg(tmp_path, OutputFileForm="image"))
    with pytest.raises(SystemExit):
        Mgfpj_v3(_cfg(tmp_path, OversampleSize=1.5))


def test_v2_fan_geometry(tmp_path):
    from reconstruction_program_taichi_amd.fpj_v2 import Mgfpj

    f = Mgfpj(_cfg(tmp_path, ConeBeam=False, OversampleSize=1,
                   DetectorElementVerticalReconRange=[0, NV - 1]))
    s = f.fp_geometry_spec()
    assert not s.cone and s.img_dim_z == NV and s.n_v == NV and not s.v.any()
    assert np.array_equal(s.u, _fbp_u())
