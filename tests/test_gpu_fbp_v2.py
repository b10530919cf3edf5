"""GPU tests of the view-block streaming FBP (run_mgfbp_v2.py drop-in) through
the C ABI (mct_copy_views, mct_weight_views, mct_backproject_views).

Parity bars:
  * streamed == resident run_mgfbp path, bit for bit (same kernels, same
    per-voxel summation order), for ragged block sizes, every back-projector
    (LDS-staged, direct-gather, fan, generic) and the weighting options that
    read the view index (short-scan Parker weight, BH polynomial);
  * against the oracle's restatement of the reference's view loop
    (recon.backproject_fbp_v2): rel-L2 <= 1e-5 with identical zero masks;
  * the -10000 / -5000 truncation marks: sums below -5000 cleared as the
    oracle's view loop clears them; voxels that leave the detector 0.
"""
import json

import numpy as np
import pytest
import torch

from oracle import recon
from tests._cases import oracle_dict, rel_l2

pytestmark = pytest.mark.gpu


def _cone_cfg(tmp_path, **kw):
    c = {"InputDir": str(tmp_path / "in"), "OutputDir": str(tmp_path / "out"),
         "InputFiles": "sgm_.*\\.raw", "OutputFilePrefix": "", "OutputFileReplace": ["sgm", "rec"],
         "ConeBeam": True, "DetectorElementCountHorizontal": 96, "DetectorElementWidth": 0.6,
         "DetectorOffsetHorizontal": 0.0,
         "DetectorElementCountVertical": 40, "DetectorElementHeight": 0.6, "Views": 90,
         "SourceIsocenterDistance": 300.0, "SourceDetectorDistance": 500.0, "ImageDimension": 48,
         "PixelSize": 0.5, "ImageDimensionZ": 24, "VoxelHeight": 0.5, "HammingFilter": 1.0}
    c.update(kw)
    return c


def _fan_cfg(tmp_path, **kw):
    c = {"InputDir": str(tmp_path / "in"), "OutputDir": str(tmp_path / "out"),
         "InputFiles": "sgm_.*\\.raw", "OutputFilePrefix": "", "OutputFileReplace": ["sgm", "rec"],
         "DetectorElementCountHorizontal": 100, "DetectorElementWidth": 0.8,
         "DetectorOffsetHorizontal": 0.0, "DetectorElementCountVertical": 3, "Views": 120, "SourceIsocenterDistance": 250.0,
         "SourceDetectorDistance": 500.0, "ImageDimension": 64, "PixelSize": 0.6,
         "HammingFilter": 0.54}
    c.update(kw)
    return c


def _sinogram(f, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 2, (f.det_elem_count_vertical_actual, f.view_num,
                              f.det_elem_count_horizontal)).astype(np.float32)


def _pair(cfg, block, kernel="auto"):
    from reconstruction_program_taichi_amd.fbp import Mgfbp
    from reconstruction_program_taichi_amd.fbp_v2 import Mgfbp_v2

    one = Mgfbp(dict(cfg))
    streamed = Mgfbp_v2(dict(cfg, ViewsPerBlock=block))
    one.ctx().set_bp_kernel(kernel)
    streamed.ctx().set_bp_kernel(kernel)
    return one, streamed


def _resident_and_streamed(one, streamed, sgm):
    res = one.reconstruct(torch.from_numpy(sgm.copy()).cuda()).cpu().numpy()
    host = torch.from_numpy(sgm.copy()).pin_memory()
    got = streamed.reconstruct_streamed(host).cpu().numpy()
    return res, got, one.ctx().bp_last_kernel(), streamed.ctx().bp_last_kernel()


CASES = {
    "cone-lds": (_cone_cfg, {}, "lds"),
    "cone-direct": (_cone_cfg, {}, "direct"),
    "cone-shortscan-bh": (_cone_cfg, {"TotalScanAngle": 230.0,
                                      "BeamHardeningCorrectionCoefficients": [1.0, 0.05]}, "auto"),
    "cone-vertical-gauss": (_cone_cfg, {"DetectorElementVerticalGaussFilterSize": 1.2}, "auto"),
    "cone-coronal-generic": (_cone_cfg, {"ReconViewMode": "coronal", "ImageDimensionZ": 16}, "auto"),
    "fan-shortscan": (_fan_cfg, {"TotalScanAngle": 200.0}, "auto"),
    "fan-f16": (_fan_cfg, {"SinogramStorage": "f16", "RampFilterMethod": "bf16x3"}, "auto"),
    # weighting, MFMA filter and ROWS taps in one call per block (mct_weight_filter_pack)
    "cone-bf16x3-fused": (_cone_cfg, {"RampFilterMethod": "bf16x3",
                                      "BeamHardeningCorrectionCoefficients": [1.0, 0.05]}, "lds"),
}
EXPECT = {"cone-lds": "lds", "cone-direct": "direct", "cone-coronal-generic": "generic",
          "fan-shortscan": "fan", "fan-f16": "fan", "cone-bf16x3-fused": "lds"}


@pytest.mark.parametrize("block", [7, 32])
@pytest.mark.parametrize("case", sorted(CASES))
def test_streamed_equals_resident(tmp_path, case, block):
    make, extra, kernel = CASES[case]
    one, streamed = _pair(make(tmp_path, **extra), block, kernel)
    sgm = _sinogram(one, 11)
    res, got, k1, k2 = _resident_and_streamed(one, streamed, sgm)
    assert k1 == k2
    assert one.fused_filter_pack() == streamed.fused_filter_pack() == (case == "cone-bf16x3-fused")
    if case in EXPECT:
        assert k2 == EXPECT[case]
    assert np.array_equal(got, res), f"{case}: max |diff| {np.abs(got - res).max():.3g}"
    assert np.count_nonzero(res) > res.size // 4


def _oracle_chain(f, sgm, v2=True):
    spec = f.geometry_spec()
    bh = list(f.array_bh_coefficients) if f.bool_bh_correction else None
    w = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle,
                              spec.short_scan, bh=bh, dtype=np.float32)
    fr = recon.filter_rows(w, f.recon_kernel(), f.det_elem_width, np.float32)
    g = oracle_dict(spec)
    if v2:
        return recon.backproject_fbp_v2(fr, g, dtype=np.float32)
    return recon.backproject_fbp(fr, g, np.float32)


@pytest.mark.parametrize("case", ["cone", "fan-shortscan"])
def test_streamed_matches_oracle_view_loop(tmp_path, case):
    from reconstruction_program_taichi_amd.fbp_v2 import Mgfbp_v2

    cfg = (_cone_cfg(tmp_path, RampFilterMethod="f32") if case == "cone"
           else _fan_cfg(tmp_path, TotalScanAngle=200.0, RampFilterMethod="f32"))
    f = Mgfbp_v2(dict(cfg, ViewsPerBlock=13))
    sgm = _sinogram(f, 12)
    got = f.reconstruct_streamed(torch.from_numpy(sgm.copy()).pin_memory()).cpu().numpy()
    ref = _oracle_chain(f, sgm)
    assert np.array_equal(got == 0, ref == 0)
    err = rel_l2(got, ref)
    print(f"{case}: streamed FBP vs oracle view loop rel-L2 {err:.2e}")
    assert err <= 1e-5


def test_truncation_marks(tmp_path):
    """Unfiltered values so large and negative that every running sum passes
    -5000: SetTruncatedRegionToZero (run_mgfbp_v2.py:232-236) clears every
    voxel, exactly as the oracle's view loop does.  With the same values
    positive, exactly the voxels that leave the detector in some view are 0
    (as in run_mgfbp; the reference's view loop would leave -10000 plus the
    later views there once those add up to more than 5000, which no
    attenuation sinogram reaches)."""
    from reconstruction_program_taichi_amd.fbp_v2 import Mgfbp_v2

    cfg = _cone_cfg(tmp_path, ImageDimension=64)  # the corners leave the detector
    del cfg["HammingFilter"]
    f = Mgfbp_v2(dict(cfg, ViewsPerBlock=17, **{"None": 0}))
    assert f.recon_kernel() is None
    sgm = np.full((40, 90, 96), -1e5, np.float32)
    got = f.reconstruct_streamed(torch.from_numpy(sgm.copy()).pin_memory()).cpu().numpy()
    ref = _oracle_chain_unfiltered(f, sgm)
    assert not ref.any() and np.array_equal(got, ref)
    pos = -sgm
    got = f.reconstruct_streamed(torch.from_numpy(pos.copy()).pin_memory()).cpu().numpy()
    ref, alive = _oracle_chain_unfiltered(f, pos, mask=True)
    assert 0 < alive.sum() < alive.size
    assert np.array_equal(got == 0, ~alive)
    assert rel_l2(got, ref) <= 1e-5


def _oracle_chain_unfiltered(f, sgm, mask=False):
    spec = f.geometry_spec()
    w = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle,
                              spec.short_scan, dtype=np.float32)
    g = oracle_dict(spec)
    if mask:
        return recon.backproject_fbp(w, g, np.float32, return_mask=True)
    return recon.backproject_fbp_v2(w, g, dtype=np.float32)


def test_backproject_views_rejects_ranges_outside_the_scan(tmp_path):
    from reconstruction_program_taichi_amd.fbp_v2 import Mgfbp_v2

    f = Mgfbp_v2(_cone_cfg(tmp_path))
    plan, ctx = f.plan(), f.ctx()
    vol = torch.zeros((24, 48, 48), device="cuda")
    blk = torch.zeros((40, 10, 96), device="cuda")
    packed = ctx.pack(blk, "f32", "rows")
    # the binding refuses the range before any call ...
    with pytest.raises(ValueError, match="outside the plan"):
        plan.backproject_views(packed, "f32", "rows", 85, 10, vol)
    with pytest.raises(ValueError, match="outside the plan"):
        plan.weight_views(blk, -1)
    # ... and so does the C ABI itself
    import ctypes as C
    from reconstruction_program_taichi_amd import _abi
    from reconstruction_program_taichi_amd.device import DTYPES, LAYOUTS
    with pytest.raises(RuntimeError, match="outside the scan"):
        _abi.call("mct_backproject_views", plan.h, packed.data_ptr(), DTYPES["f32"], LAYOUTS["rows"],
                  85, 10, 0, vol.data_ptr())
    with pytest.raises(RuntimeError, match="outside the scan"):
        _abi.call("mct_weight_views", plan.h, blk.data_ptr(), -1, 10, None, 0)
    host = torch.zeros((40, 90, 96))
    with pytest.raises(RuntimeError, match="outside the sinogram"):
        ctx.copy_views(host, 81, 10, blk)


def test_run_mgfbp_v2_file_equals_run_mgfbp(tmp_path):
    from reconstruction_program_taichi_amd.fbp import run_mgfbp
    from reconstruction_program_taichi_amd.fbp_v2 import Mgfbp_v2, run_mgfbp_v2

    (tmp_path / "in").mkdir()
    cfg = _cone_cfg(tmp_path, Views=80, SinogramHeight=84, DetectorElementVerticalReconRange=[3, 36],
                    SaveFilteredSinogram=True, WaterMu=0.02, OutputFileFormat="tif")
    rng = np.random.default_rng(4)
    full = rng.uniform(0, 2, (40, 84, 96)).astype(np.float32)
    full.tofile(tmp_path / "in" / "sgm_1.raw")
    (tmp_path / "a.jsonc").write_text(json.dumps(dict(cfg, OutputDir=str(tmp_path / "o1"))))
    (tmp_path / "b.jsonc").write_text(json.dumps(dict(cfg, OutputDir=str(tmp_path / "o2"),
                                                      ViewsPerBlock=9)))
    a = run_mgfbp(str(tmp_path / "a.jsonc"))
    b = run_mgfbp_v2(str(tmp_path / "b.jsonc"))
    assert np.array_equal(a, b)
    assert a.shape == (24, 48, 48) and np.count_nonzero(a) > a.size // 4
    assert (tmp_path / "o2" / "rec_1.tif").exists()
    # the filtered-sinogram file holds every view in run_mgfbp and the last
    # view only in the view-by-view drop-in (its device buffer holds one view)
    s1 = (tmp_path / "o1" / "sgm_filtered_rec_1.tif").stat().st_size
    s2 = (tmp_path / "o2" / "sgm_filtered_rec_1.tif").stat().st_size
    plane = 34 * 96 * 4
    assert s1 >= 80 * plane and plane <= s2 < 2 * plane
    assert Mgfbp_v2(dict(cfg, ViewsPerBlock=9)).views_per_block == 9
