"""GPU tests of the drop-in entry points (config file in, files out) and of the
device IRN-CGLS / PICCS solver against the oracle's restatement of the loop.

Tolerances: IR iterates rel-L2 <= 1e-4 against the FP64 oracle loop (FP32
projectors, FP64 CG scalars; CG amplifies the FP32 rounding of P^T P);
drop-in outputs against the oracle chain at the kernels' own tolerances.
"""
import json

import numpy as np
import pytest
import torch

from oracle import recon
from reconstruction_program_taichi_amd import geometry as geo
from tests._cases import PI, oracle_dict, rel_l2

pytestmark = pytest.mark.gpu


def _write_cfg(path, cfg):
    path.write_text("// test config\n" + json.dumps(cfg, indent=1))
    return str(path)


def _fan_cfg(tmp_path, views=60, **kw):
    c = {"InputDir": str(tmp_path / "in"), "OutputDir": str(tmp_path / "out"),
         "InputFiles": "sgm_.*\\.raw", "OutputFilePrefix": "p_", "OutputFileReplace": ["sgm", "rec"],
         "DetectorElementCountHorizontal": 72, "DetectorElementWidth": 0.8,
         "DetectorOffsetHorizontal": 0.0, "DetectorElementCountVertical": 2, "Views": views,
         "SourceIsocenterDistance": 250.0, "SourceDetectorDistance": 500.0, "ImageDimension": 40,
         "PixelSize": 0.6, "HammingFilter": 1.0}
    c.update(kw)
    return c


def test_run_mgfbp_file_roundtrip(tmp_path):
    from reconstruction_program_taichi_amd.fbp import Mgfbp, run_mgfbp

    (tmp_path / "in").mkdir()
    cfg = _fan_cfg(tmp_path, WaterMu=0.02)
    rng = np.random.default_rng(0)
    sgm = rng.uniform(0, 1, (2, 60, 72)).astype(np.float32)
    sgm.tofile(tmp_path / "in" / "sgm_a.raw")
    out = run_mgfbp(_write_cfg(tmp_path / "c.jsonc", cfg))
    written = np.fromfile(tmp_path / "out" / "p_rec_a.raw", dtype=np.float32).reshape(out.shape)
    assert np.array_equal(written, out)
    spec = Mgfbp(_fan_cfg(tmp_path)).geometry_spec()
    w = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle, False)
    taps = geo.hamming_taps(72, 0.8, 1.0, 500.0)
    fr = recon.filter_rows(w, taps, 0.8, np.float32)
    ref = recon.backproject_fbp(fr, oracle_dict(spec), np.float32)
    ref_hu = (ref / 0.02 - 1) * 1000
    assert rel_l2(out, ref_hu) <= 1e-5


def test_run_mgfbp_post_log_images_input(tmp_path):
    """InputFileForm post_log_images (run_mgfbp.py:1206-1222): the file holds
    one [rows][n_u] image per view; the recon-range rows of the first Views
    images are read (offset / gap per image) and back-projected.  The result
    equals the same data given as a sinogram file whose first slice is the
    detector's top row (post-log images are upright, run_mgfbp.py:177-181)
    bit for bit, and the oracle chain on the rows it should have read."""
    from reconstruction_program_taichi_amd.fbp import Mgfbp, run_mgfbp

    (tmp_path / "in").mkdir()
    (tmp_path / "in2").mkdir()
    cone = dict(ConeBeam=True, DetectorElementHeight=0.6, ImageDimensionZ=8, VoxelHeight=0.5,
                DetectorElementCountVertical=20, DetectorElementVerticalReconRange=[4, 15],
                DetectorOffsetVertical=0.0, SinogramHeight=26, WaterMu=0.02)
    rng = np.random.default_rng(7)
    images = rng.uniform(0, 1, (26, 20, 72)).astype(np.float32)  # [view][row][col]
    images.tofile(tmp_path / "in" / "sgm_p.raw")
    cfg = _fan_cfg(tmp_path, views=24, InputFileForm="post_log_images", **cone)
    out = run_mgfbp(_write_cfg(tmp_path / "c.jsonc", cfg))
    # the same rows as a sinogram file [row][view][col]
    np.ascontiguousarray(images.transpose(1, 0, 2)).tofile(tmp_path / "in2" / "sgm_p.raw")
    cfg2 = _fan_cfg(tmp_path, views=24, FirstSinogramSliceIsDetectorTopRow=True, **cone)
    cfg2.update(InputDir=str(tmp_path / "in2"), OutputDir=str(tmp_path / "out2"))
    out2 = run_mgfbp(_write_cfg(tmp_path / "c2.jsonc", cfg2))
    assert out.shape == (8, 40, 40) and np.array_equal(out, out2)
    spec = Mgfbp(cfg).geometry_spec()
    sgm = np.ascontiguousarray(images[:24, 4:16, :].transpose(1, 0, 2))
    w = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle, False)
    fr = recon.filter_rows(w, geo.hamming_taps(72, 0.8, 1.0, 500.0), 0.8, np.float32)
    ref = (recon.backproject_fbp(fr, oracle_dict(spec), np.float32) / 0.02 - 1) * 1000
    assert rel_l2(out, ref) <= 1e-5


def test_run_mgfpj_cone_file_roundtrip(tmp_path):
    from reconstruction_program_taichi_amd.fpj import Mgfpj, run_mgfpj

    (tmp_path / "in").mkdir()
    cfg = {"InputDir": str(tmp_path / "in"), "OutputDir": str(tmp_path / "out"),
           "InputFiles": "img_.*", "OutputFilePrefix": "", "OutputFileReplace": ["img", "sgm"],
           "ImageDimension": 32, "PixelSize": 0.5, "ImageDimensionZ": 12, "ConeBeam": True,
           "VoxelHeight": 0.5, "DetectorElementCountHorizontal": 48, "DetectorElementWidth": 0.6,
           "DetectorElementCountVertical": 20, "DetectorElementHeight": 0.6, "Views": 16,
           "SourceIsocenterDistance": 300.0, "SourceDetectorDistance": 500.0}
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 0.02, (12, 32, 32)).astype(np.float32)
    img.tofile(tmp_path / "in" / "img_b.raw")
    out = run_mgfpj(_write_cfg(tmp_path / "c.jsonc", cfg))
    written = np.fromfile(tmp_path / "out" / "sgm_b.raw", dtype=np.float32).reshape(20, 16, 48)
    assert np.array_equal(written, out)
    spec = Mgfpj(cfg).geometry_spec()
    ref = recon.forward_project(img, oracle_dict(spec), dtype=np.float32)
    # cone-beam 'sinogram' output flips the detector rows (run_mgfpj.py:657-661)
    assert rel_l2(out, ref[::-1]) <= 1e-5


def test_mgfpj_noise_is_seeded(tmp_path):
    from reconstruction_program_taichi_amd.fpj import Mgfpj

    cfg = {"InputDir": str(tmp_path), "OutputDir": str(tmp_path / "o"), "InputFiles": "x",
           "OutputFileReplace": ["a", "b"], "ImageDimension": 24, "PixelSize": 0.5, "SliceCount": 2,
           "DetectorElementCountHorizontal": 40, "DetectorElementWidth": 0.6, "Views": 12,
           "SourceIsocenterDistance": 200.0, "SourceDetectorDistance": 400.0, "PhotonNumber": 1e4,
           "NoiseSeed": 5}
    vol = torch.rand((2, 24, 24), device="cuda") * 0.02
    a = Mgfpj(cfg).project(vol).cpu().numpy()
    b = Mgfpj(cfg).project(vol).cpu().numpy()
    assert np.array_equal(a, b)
    clean = Mgfpj({k: v for k, v in cfg.items() if k != "PhotonNumber"}).project(vol).cpu().numpy()
    assert rel_l2(a, recon.add_noise(clean, 1e4, 5)) <= 1e-5


def _ir_cfg(tmp_path, **kw):
    c = _fan_cfg(tmp_path, views=30, DetectorElementCountVertical=3, Lambda=1e-3, BetaTV=1e-6,
                 NumberOfIRNIterations=1, NumberOfIterations=4)
    c.update(kw)
    return c


@pytest.mark.parametrize("couple", [True, False])
def test_ir_solver_matches_oracle_loop(tmp_path, couple):
    from reconstruction_program_taichi_amd.ir import Mgfbp_ir

    ir = Mgfbp_ir(_ir_cfg(tmp_path, TVCoupleSlices=couple))
    spec = ir.ir_geometry_spec()
    rng = np.random.default_rng(2)
    sgm = rng.uniform(0, 1, (3, 30, 72)).astype(np.float32)
    x0 = rng.uniform(0, 0.01, (3, 40, 40)).astype(np.float32)
    got = ir.solve(torch.from_numpy(sgm).cuda(), torch.from_numpy(x0).cuda()).cpu().numpy()
    g = oracle_dict(spec)
    ref = recon.ir_solve(sgm, g, x0, ir.coef_lambda, ir.beta_tv, 1, 4, couple_z=couple)
    err = rel_l2(got, ref)
    print(f"IRN-CGLS couple={couple}: rel-L2 vs FP64 oracle loop {err:.2e}")
    assert err <= 1e-4
    assert len(ir.loss) == 2 and np.all(np.isfinite(ir.loss))


def test_piccs_solver_matches_oracle_loop(tmp_path):
    from reconstruction_program_taichi_amd.ir import Mgfbp_ir_piccs

    rng = np.random.default_rng(3)
    prior = rng.uniform(0, 0.01, (3, 40, 40)).astype(np.float32)
    prior.tofile(tmp_path / "prior.raw")
    ir = Mgfbp_ir_piccs(_ir_cfg(tmp_path, Alpha=0.5, PriorImageFile=str(tmp_path / "prior.raw")))
    spec = ir.ir_geometry_spec()
    sgm = rng.uniform(0, 1, (3, 30, 72)).astype(np.float32)
    got = ir.solve(torch.from_numpy(sgm).cuda(), torch.from_numpy(prior).cuda()).cpu().numpy()
    ref = recon.ir_solve(sgm, oracle_dict(spec), prior, ir.coef_lambda, ir.beta_tv, 1, 4,
                         prior=prior, alpha=0.5)
    err = rel_l2(got, ref)
    print(f"PICCS: rel-L2 vs FP64 oracle loop {err:.2e}")
    assert err <= 1e-4


def test_ir_cone_solver_matches_oracle_loop(tmp_path):
    from reconstruction_program_taichi_amd.ir import Mgfbp_ir

    cfg = _ir_cfg(tmp_path, ConeBeam=True, DetectorElementHeight=0.6, ImageDimensionZ=10,
                  VoxelHeight=0.5, DetectorElementCountVertical=16, DetectorOffsetVertical=0.0,
                  NumberOfIterations=3)
    ir = Mgfbp_ir(cfg)
    spec = ir.ir_geometry_spec()
    rng = np.random.default_rng(4)
    sgm = rng.uniform(0, 1, (16, 30, 72)).astype(np.float32)
    x0 = rng.uniform(0, 0.01, (10, 40, 40)).astype(np.float32)
    got = ir.solve(torch.from_numpy(sgm).cuda(), torch.from_numpy(x0).cuda()).cpu().numpy()
    ref = recon.ir_solve(sgm, oracle_dict(spec), x0, ir.coef_lambda, ir.beta_tv, 1, 3)
    err = rel_l2(got, ref)
    print(f"IRN-CGLS cone: rel-L2 vs FP64 oracle loop {err:.2e}")
    assert err <= 1e-4


def test_run_mgfbp_ir_end_to_end(tmp_path):
    """run_mgfbp_ir (run_mgfbp_iterative_recon_IRN_CGLS.py:14-49, 138-176)
    against the oracle chain: FBP seed in HU, HU -> mu, IRN-CGLS-TV from that
    seed on the same sinogram, mu -> HU; compared in mu units."""
    from reconstruction_program_taichi_amd.fbp import Mgfbp
    from reconstruction_program_taichi_amd.ir import Mgfbp_ir, run_mgfbp_ir

    (tmp_path / "in").mkdir()
    cfg = _ir_cfg(tmp_path, WaterMu=0.02)
    rng = np.random.default_rng(5)
    sgm = rng.uniform(0, 1, (3, 30, 72)).astype(np.float32)
    sgm.tofile(tmp_path / "in" / "sgm_c.raw")
    out = run_mgfbp_ir(_write_cfg(tmp_path / "c.jsonc", cfg))
    assert out.shape == (3, 40, 40) and np.all(np.isfinite(out))
    written = np.fromfile(tmp_path / "out" / "p_rec_c.raw", dtype=np.float32).reshape(out.shape)
    assert np.array_equal(written, out)
    # oracle chain
    spec = Mgfbp(cfg).geometry_spec()
    w = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle, False)
    fr = recon.filter_rows(w, geo.hamming_taps(72, 0.8, 1.0, 500.0), 0.8, np.float32)
    seed_hu = ((recon.backproject_fbp(fr, oracle_dict(spec), np.float32) / np.float32(0.02) - 1)
               * 1000).astype(np.float32)
    seed_mu = ((seed_hu / np.float32(1000) + 1) * np.float32(0.02)).astype(np.float32)
    ir = Mgfbp_ir(cfg)
    x = recon.ir_solve(sgm, oracle_dict(ir.ir_geometry_spec()), seed_mu, ir.coef_lambda, ir.beta_tv,
                       1, 4)
    got_mu = (out.astype(np.float64) / 1000 + 1) * 0.02
    err = rel_l2(got_mu, x)
    print(f"run_mgfbp_ir vs oracle chain (FBP seed, HU round trip, IRN-CGLS): rel-L2 {err:.2e} (mu)")
    assert err <= 1e-4
