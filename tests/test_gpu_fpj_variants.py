"""GPU parity of the run_mgfpj_v2 / run_mgfpj_v3 drop-ins (config file in,
sinogram file out) against the oracle's forward projection on the geometry
each variant uses; rows outside the vertical range, noise and the P-matrix
branch included.  Tolerance: rel-L2 <= 1e-5 (noise: 1e-6 against the
oracle's Philox stream on the same projection)."""
import json

import numpy as np
import pytest

from oracle import recon
from tests._cases import oracle_dict, pmatrix_for, rel_l2

pytestmark = pytest.mark.gpu

N, NV, VIEWS = 40, 12, 30


def _cfg(tmp_path, **kw):
    (tmp_path / "in").mkdir(exist_ok=True)
    c = {"InputDir": str(tmp_path / "in"), "OutputDir": str(tmp_path / "out"), "InputFiles": "img_.*",
         "OutputFilePrefix": "", "OutputFileReplace": ["img_", "sgm_"], "SinogramWidth": N,
         "DetectorElementWidth": 0.61, "DetectorOffcenter": 0.3, "SliceCount": NV, "Views": VIEWS,
         "SourceIsocenterDistance": 300.0, "SourceDetectorDistance": 500.0, "ImageDimension": 24,
         "PixelSize": 0.5, "HammingFilter": 1.0, "ConeBeam": True, "DetectorElementHeight": 0.52,
         "ImageSliceCount": 10, "VoxelHeight": 0.5, "SliceOffCenter": 0.2,
         "DetectorElementVerticalReconRange": [2, 9], "ForwardProjectionStepSize": 0.3}
    c.update(kw)
    return c


def _run(tmp_path, cfg, entry):
    img = np.random.default_rng(31).uniform(0, 0.02, (cfg.get("ImageSliceCount", NV), 24, 24))
    img = img.astype(np.float32)
    img.tofile(tmp_path / "in" / "img_a.raw")
    (tmp_path / "c.jsonc").write_text(json.dumps(cfg))
    out = entry(str(tmp_path / "c.jsonc"))
    written = np.fromfile(tmp_path / "out" / "sgm_a.raw", dtype=np.float32).reshape(out.shape)
    assert np.array_equal(written, out)
    return img, out


def _expected(fpj, img, os_=1):
    spec = fpj.fp_geometry_spec()
    g = oracle_dict(spec)
    if spec.matrix_a is not None:
        g.update(matrix_a=spec.matrix_a, x_s=spec.x_s)
    part = recon.forward_project(img, g, oversample=os_, count_mode=spec.fpj_count_mode, dtype=np.float32)
    full = np.zeros((NV, VIEWS, N), np.float32)
    b = fpj.det_elem_vertical_recon_range_begin
    full[b:b + part.shape[0]] = part
    return full, b, part.shape[0]


@pytest.mark.parametrize("os_", [1, 2])
def test_run_mgfpj_v2_cone(tmp_path, os_):
    from reconstruction_program_taichi_amd.fpj_v2 import Mgfpj, run_mgfpj_v2

    cfg = _cfg(tmp_path, OversampleSize=os_)
    img, out = _run(tmp_path, cfg, run_mgfpj_v2)
    ref, b, rows = _expected(Mgfpj(_cfg(tmp_path, OversampleSize=os_)), img, os_)
    got = out[::-1]  # cone-beam 'sinogram' output flips the rows
    assert not got[:b].any() and not got[b + rows:].any()
    assert rel_l2(got, ref) <= 1e-5


def test_run_mgfpj_v2_fan(tmp_path):
    from reconstruction_program_taichi_amd.fpj_v2 import Mgfpj, run_mgfpj_v2

    kw = dict(ConeBeam=False, DetectorElementVerticalReconRange=[0, NV - 1], ImageSliceCount=NV)
    img, out = _run(tmp_path, _cfg(tmp_path, **kw), run_mgfpj_v2)
    ref, _, _ = _expected(Mgfpj(_cfg(tmp_path, **kw)), img)
    assert rel_l2(out, ref) <= 1e-5


def test_run_mgfpj_v3_center_z_pmatrix_noise(tmp_path):
    from reconstruction_program_taichi_amd.fpj_v3 import Mgfpj_v3, run_mgfpj_v3

    cz = np.linspace(-0.6, 0.6, VIEWS).astype(np.float32)
    (tmp_path / "cz.jsonc").write_text(json.dumps({"Value": cz.tolist()}))
    base = Mgfpj_v3(_cfg(tmp_path, DetectorElementVerticalReconRange=[0, NV - 1]))
    pm = pmatrix_for(base.geometry_spec()).astype(np.float32)
    (tmp_path / "pm.jsonc").write_text(json.dumps({"Value": [float(x) for x in pm]}))
    kw = dict(ImageCenterZFile=str(tmp_path / "cz.jsonc"), PMatrixFile=str(tmp_path / "pm.jsonc"))
    img, out = _run(tmp_path, _cfg(tmp_path, **kw), run_mgfpj_v3)
    fpj = Mgfpj_v3(_cfg(tmp_path, **kw))
    assert fpj.bool_apply_pmatrix
    ref, _, _ = _expected(fpj, img)
    got = out[::-1]
    split = np.abs(got - ref) > 1e-5 * np.abs(ref).max()  # samples on the volume's edge
    assert split.mean() <= 1e-3 and rel_l2(got[~split], ref[~split]) <= 1e-5
    # noise over every row, rows outside the range included (run_mgfpj_v3.py:112-115)
    kw.update(PhotonNumber=1e5, NoiseSeed=7)
    img, out = _run(tmp_path, _cfg(tmp_path, **kw), run_mgfpj_v3)
    got = out[::-1]
    want = recon.add_noise(ref, 1e5, 7)
    assert rel_l2(got[~split], want[~split]) <= 1e-5
    assert got[:2].any() and np.isfinite(got).all()
