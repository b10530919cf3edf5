"""The projector autotuner: candidates and cache on the host, tuning on the GPU."""
import json

import numpy as np
import pytest
import torch

from reconstruction_program_taichi_amd import autotune
from tests._cases import cone_case, fan_case, oracle_dict, rel_l2


def test_candidates_cover_both_kernels_and_valid_tiles():
    cone = autotune.bp_candidates(True)
    assert ("lds", 0, 0) in cone
    assert all(b <= zc for k, zc, b in cone if k == "direct")
    assert {zc for k, zc, _ in cone if k == "direct"} == {8, 16, 32}
    fan = autotune.bp_candidates(False)
    assert all(k == "direct" for k, _, _ in fan) and len(fan) == 4


class _FakePlan:
    def __init__(self, spec):
        self.spec = spec


class _FakeCtx:
    def __init__(self):
        self.calls = []

    def set_bp_kernel(self, k):
        self.calls.append(("kernel", k))

    def set_bp_tile(self, zc, b):
        self.calls.append(("tile", zc, b))


def test_cached_choice_is_applied_without_timing(tmp_path):
    spec = cone_case()
    key = autotune.geometry_key(_FakePlan(spec), "bp", "f16")
    assert key != autotune.geometry_key(_FakePlan(spec), "bp", "f32")
    assert key != autotune.geometry_key(_FakePlan(fan_case()), "bp", "f16")
    path = tmp_path / "tune.json"
    path.write_text(json.dumps({key: {"kernel": "direct", "zc": 32, "batch": 8, "ms": 1.0}}))
    ctx = _FakeCtx()
    tuner = autotune.Autotuner(ctx, cache_path=str(path))
    got = tuner.tune_backprojector(_FakePlan(spec), None, "f16", "quad", None)
    assert got["zc"] == 32 and ctx.calls == [("kernel", "direct"), ("tile", 32, 8)]


@pytest.mark.gpu
def test_autotuner_picks_and_applies_a_candidate(ctx, tmp_path):
    from oracle import recon
    from reconstruction_program_taichi_amd.device import Plan

    spec = cone_case(nz=37)
    rng = np.random.default_rng(2)
    f = rng.standard_normal((spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    plan = Plan(ctx, spec)
    packed = ctx.pack(torch.from_numpy(f).cuda(), "f32", "quad")
    vol = torch.empty((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32, device="cuda")
    tuner = autotune.Autotuner(ctx, cache_path=str(tmp_path / "t.json"), reps=1)
    try:
        best = tuner.tune_backprojector(plan, packed, "f32", "quad", vol)
        assert (best["kernel"], best["zc"], best["batch"]) in autotune.bp_candidates(True)
        assert best["candidates"] == len(autotune.bp_candidates(True))
        plan.backproject_fbp(packed, "f32", "quad", vol)
        assert ctx.bp_last_kernel() == best["kernel"]
        ref = recon.backproject_fbp(f, oracle_dict(spec), np.float32)
        assert rel_l2(vol.cpu().numpy(), ref) <= 1e-5
        # forward projector
        out = torch.empty((spec.n_v, spec.n_views, spec.n_u), dtype=torch.float32, device="cuda")
        fp = tuner.tune_forward_projector(plan, vol, out)
        assert fp["kernel"] in autotune.FP_CANDIDATES and fp["candidates"] == 2
        assert json.loads((tmp_path / "t.json").read_text())
    finally:
        ctx.set_bp_kernel("auto")
        ctx.set_bp_tile(0, 0)
        ctx.set_fp_kernel("auto")
