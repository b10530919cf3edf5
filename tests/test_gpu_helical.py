"""GPU parity of the helical rebinning FBP (run_mgfbp_helical.py) through the C
ABI: rebinning against the oracle's FP32 restatement, and the drop-in's
reconstruction (rebin, per-slice weighting, ramp filter, batched fan
back-projection) against the oracle chain on the same slice geometry.
Tolerances: rebinning rel-L2 <= 1e-6; reconstruction rel-L2 <= 1e-5."""
import json

import numpy as np
import pytest
import torch

from oracle import recon
from tests._cases import rel_l2
from tests.test_helical import CYL, GAMMA, N, PIX, cylinder_sinogram, helical_cfg, oracle_helical

pytestmark = pytest.mark.gpu


def test_helical_rebin_matches_oracle(tmp_path):
    from reconstruction_program_taichi_amd.fbp_helical import Mgfbp_helical

    f = Mgfbp_helical(helical_cfg(tmp_path, SliceOffCenter=0.37))
    sgm = np.random.default_rng(41).uniform(0, 2, (f.det_elem_count_vertical, f.view_num,
                                                   f.det_elem_count_horizontal)).astype(np.float32)
    got = f.rebin(torch.from_numpy(sgm).cuda()).cpu().numpy()
    ref = recon.helical_rebin(sgm, f.array_z_pos, f.view_count, f.det_offset_vertical, f.det_elem_height,
                              f.source_det_dis, f.mag_ratio_isocenter, f.dis_per_view, np.float32)
    assert got.shape == ref.shape
    assert np.array_equal(got == 0, ref == 0)
    assert rel_l2(got, ref) <= 1e-6


def test_run_mgfbp_helical_matches_oracle_chain(tmp_path):
    from reconstruction_program_taichi_amd.fbp_helical import Mgfbp_helical, run_mgfbp_helical

    cfg = helical_cfg(tmp_path)
    f = Mgfbp_helical(dict(cfg))
    sgm = cylinder_sinogram(f).astype(np.float32)
    sgm.tofile(tmp_path / "in" / "sgm_h.raw")
    (tmp_path / "c.jsonc").write_text(json.dumps(cfg))
    out = run_mgfbp_helical(str(tmp_path / "c.jsonc"))
    written = np.fromfile(tmp_path / "out" / "rec_h.raw", dtype=np.float32).reshape(out.shape)
    assert np.array_equal(written, out)
    reb_ref, ref = oracle_helical(f, sgm, np.float32)
    reb = np.fromfile(tmp_path / "in" / "temp" / "sgm_rebin.raw", dtype=np.float32).reshape(reb_ref.shape)
    assert rel_l2(reb, reb_ref) <= 1e-6
    assert rel_l2(out, ref) <= 1e-5
    # and it is the cylinder: mu (1 + GAMMA z) inside, centred where it is
    cx, cy, R, mu = CYL
    xs = (np.arange(N) - (N - 1) / 2) * PIX
    X, Y = np.meshgrid(xs, -xs)
    r = np.hypot(X - cx, Y - cy)
    for z in range(out.shape[0]):
        want = mu * (1 + GAMMA * float(f.array_z_pos[z]))
        assert abs(np.median(out[z][r < R - 2]) - want) < 0.02 * want
