"""The smallest geometries through every operator, against the oracle:
FBP/FDK chains (FP32 and split-BF16 ramp filter; LDS-staged and direct
cone back-projectors), forward projection, IR back-projection and the ray
mask.  Two views and three detector elements (and three rows for cone
beam) with images of 1-3 pixels; with the two elements and two rows the
plan accepts at least, no voxel passes the reference's u range test and
no ray meets the image, so every result is 0 - checked too.

Tolerances as in the other parity tests: rel-L2 <= 1e-5 with identical
masks for FP32 paths, <= 1e-3 for the split-BF16 filter.
"""
import numpy as np
import pytest
import torch

from oracle import recon
from reconstruction_program_taichi_amd.device import Plan
from tests._cases import cone_case, fan_case, oracle_dict, rel_l2

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()


def tiny_fan(nx=3, nu=3):
    return fan_case(nx=nx, nv=2, nu=nu, pix=0.1, w=0.8, rows=1)


def tiny_cone(nx=3, nz=2, nu=3):
    return cone_case(nx=nx, nz=nz, nv=2, nu=nu, nrows=nu, pix=0.1, vh=0.1, w=0.8, h=0.8)


@pytest.mark.parametrize("nu", [2, 3])
@pytest.mark.parametrize("kind,method,bp", [("fan", "f32", "auto"), ("fan", "bf16x3", "auto"),
                                            ("cone", "f32", "lds"), ("cone", "f32", "direct"),
                                            ("cone", "bf16x3", "auto")])
def test_smallest_fbp_chain(ctx, kind, method, bp, nu):
    spec = tiny_fan(nu=nu) if kind == "fan" else tiny_cone(nu=nu)
    du = float(abs(spec.u[1] - spec.u[0]))
    rng = np.random.default_rng(3)
    sgm = rng.uniform(0.5, 1.5, (spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    taps = recon.ramlak_taps(spec.n_u, du, np.float32)
    plan = Plan(ctx, spec)
    w, filt = dev(sgm), torch.empty((spec.n_v, spec.n_views, spec.n_u), device="cuda")
    if method == "f32":
        plan.weight(w)
        ctx.filter_rows(w, filt, taps, du, "f32")
    else:
        plan.weight_filter(w, filt, taps, du, method)
    layout = "rows" if spec.cone else "pair"
    vol = torch.full((spec.img_dim_z, spec.img_dim, spec.img_dim), 7.0, device="cuda")
    ctx.set_bp_kernel(bp)
    try:
        plan.backproject_fbp(ctx.pack(filt, "f32", layout), "f32", layout, vol)
        if kind == "cone" and bp != "auto":
            assert ctx.bp_last_kernel() == bp
    finally:
        ctx.set_bp_kernel("auto")
    got = vol.cpu().numpy()
    wo = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle, False)
    ref = recon.backproject_fbp(recon.filter_rows(wo, taps, du, np.float32), oracle_dict(spec), np.float32)
    assert np.array_equal(got == 0, ref == 0)
    if nu == 2:
        assert not ref.any()
        return
    assert np.abs(ref).max() > 0 and rel_l2(got, ref) <= (1e-5 if method == "f32" else 1e-3)


@pytest.mark.parametrize("kind,nx,nu", [("fan", 1, 3), ("fan", 2, 3), ("fan", 2, 2), ("cone", 1, 3),
                                         ("cone", 2, 3), ("cone", 2, 2)])
@pytest.mark.parametrize("kernel", ["auto", "direct"])
def test_smallest_forward_projection(ctx, kind, nx, nu, kernel):
    """A one-pixel image has no cell, so every line sum is 0 (no sample
    passes 0 <= index, index + 1 < dim); two pixels make one cell, which
    the centre ray of three detector elements crosses."""
    spec = tiny_fan(nx, nu) if kind == "fan" else tiny_cone(nx, nu=nu)
    spec.fpj_step = 0.2
    vol = np.random.default_rng(4).uniform(0.5, 1.5, (spec.img_dim_z, nx, nx)).astype(np.float32)
    out = torch.full((spec.n_v, spec.n_views, spec.n_u), 7.0, device="cuda")
    ctx.set_fp_kernel(kernel)
    try:
        Plan(ctx, spec).forward_project(dev(vol), out)
    finally:
        ctx.set_fp_kernel("auto")
    got = out.cpu().numpy()
    ref = recon.forward_project(vol, oracle_dict(spec), dtype=np.float32)
    if nx == 1 or nu == 2:
        assert not ref.any() and not got.any()
    else:
        assert ref.max() > 0 and rel_l2(got, ref) <= 1e-5 and np.array_equal(got == 0, ref == 0)


@pytest.mark.parametrize("kind", ["fan", "cone"])
def test_smallest_ir_backprojection_and_ray_mask(ctx, kind):
    spec = tiny_fan() if kind == "fan" else tiny_cone()
    plan = Plan(ctx, spec)
    mask = torch.empty((spec.n_v, spec.n_views, spec.n_u), device="cuda")
    plan.ray_mask(mask)
    mref = recon.ray_mask(oracle_dict(spec))
    assert np.array_equal(mask.cpu().numpy(), mref)
    sgm = np.random.default_rng(5).uniform(0.5, 1.5, (spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    vol = torch.full((spec.img_dim_z, spec.img_dim, spec.img_dim), 7.0, device="cuda")
    plan.backproject_ir(dev(sgm), mask, vol)
    got = vol.cpu().numpy()
    ref = recon.backproject_ir(sgm, oracle_dict(spec), ray_mask=mref, dtype=np.float32)
    assert np.abs(ref).max() > 0 and np.array_equal(got == 0, ref == 0)
    assert rel_l2(got, ref) <= 1e-5
