"""Volumes past 4 GiB: a 1024^3 cone geometry (0.15 mm voxels, 1080 views
of a 1024 x 1024 detector at 0.3 mm, SID 750 / SDD 1250 mm).

- Back-projection: the LDS-staged kernel's window table is 2.3 GB here
  (64 x 64 tiles x 64 slice chunks x 1080 views x 8 B) and the packed
  sinogram (ROWS layout) 9 GB; the kernel runs (no fallback) and two slices (the first
  and a middle one) match the oracle's C restatement, masks identical.
  The input is a smooth field, as filtered projections are: the kernel's
  magnification comes from the hardware reciprocal (about 1 ulp), which
  moves the detector coordinate by up to ~5e-5 of a column at 1024
  columns - on white noise that alone is 5e-5 rel-L2, on smooth data far
  below the 1e-5 bound.
- Forward projection: the volume's records would need 17 GB, past 32-bit
  offsets, so the direct kernel runs and the plan says so with a
  RuntimeWarning; the sums match the oracle on a subset of rays.

About 30 GB of device memory and 10 GB of host memory.
"""
import numpy as np
import pytest
import torch

from oracle import cbp
from reconstruction_program_taichi_amd import geometry as geo
from reconstruction_program_taichi_amd.device import Plan
from tests._cases import oracle_dict, rel_l2

pytestmark = pytest.mark.gpu


def cube_spec(n_views=1080):
    ang = geo.angle_array(n_views, 2 * geo.PI, 0.0)
    u = geo.det_pos_array(1024, 1024, -0.3, 0.0)
    v = geo.det_pos_array(1024, 1024, 0.3, 0.0)
    return geo.GeometrySpec(img_dim=1024, img_dim_z=1024, pix=0.15, vox_h=0.15, n_views=n_views,
                            sid=750.0, sdd=1250.0, total_scan_angle=2 * geo.PI, angles=ang, u=u, v=v,
                            cone=True, n_v_full=1024, det_h=0.3)


def test_backprojection_of_a_1024_cube_with_a_2gb_window_table(ctx):
    import warnings

    spec = cube_spec()
    gen = torch.Generator(device="cuda").manual_seed(21)
    ph = torch.rand((4, 1, 1080, 1), generator=gen, device="cuda") * 2 * np.pi
    c = torch.arange(1024, device="cuda", dtype=torch.float32).view(1, 1, 1024)
    r = torch.arange(1024, device="cuda", dtype=torch.float32).view(1024, 1, 1)
    filt = torch.sin(c * (2 * np.pi / 173) + ph[0]) * torch.cos(r * (2 * np.pi / 311) + ph[1])
    filt += 0.5 * torch.sin(c * (2 * np.pi / 251) + ph[2]) * torch.cos(r * (2 * np.pi / 197) + ph[3])
    plan = Plan(ctx, spec)
    packed = ctx.pack(filt, "f32", "rows")
    vol = torch.empty((1024, 1024, 1024), dtype=torch.float32, device="cuda")
    with warnings.catch_warnings():
        warnings.simplefilter("error", RuntimeWarning)  # no fallback
        plan.backproject_fbp(packed, "f32", "rows", vol)
    assert ctx.bp_last_kernel() == "lds"
    del packed
    fh = filt.cpu().numpy()
    del filt
    g = dict(oracle_dict(spec), short_scan=False)
    for z in (0, 600):
        ref = cbp.backproject_fbp(fh, g, z, z + 1)
        got = vol[z:z + 1].cpu().numpy()
        assert np.array_equal(got == 0, ref == 0)
        assert (ref != 0).any() and rel_l2(got, ref) <= 1e-5, z
    plan.close()


def test_forward_projection_of_a_1024_cube_falls_back_with_a_warning(ctx):
    spec = cube_spec(n_views=2)
    spec.fpj_step = 0.2
    gen = torch.Generator(device="cuda").manual_seed(22)
    vol = torch.rand((1024, 1024, 1024), generator=gen, device="cuda") * 0.02
    out = torch.empty((1024, 2, 1024), dtype=torch.float32, device="cuda")
    plan = Plan(ctx, spec)
    with pytest.warns(RuntimeWarning, match="direct kernel"):
        plan.forward_project(vol, out)
    assert ctx.fp_last_kernel() == "direct"
    rows = list(range(0, 1024, 64))
    got = out[rows].cpu().numpy()
    g = dict(oracle_dict(spec), v=spec.v[rows])
    ref = cbp.forward_project(vol.cpu().numpy(), g, dtype=np.float32)
    assert ref.max() > 1.0 and rel_l2(got, ref) <= 1e-5
    plan.close()
