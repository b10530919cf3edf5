"""GPU parity of the P-matrix branches against the oracle, through the C ABI:
FBP back-projection (run_mgfbp.py:1133-1141, 1165-1168), IR back-projection
(run_mgfbp_iterative_recon_IRN_CGLS.py:423-458), forward projection and ray
mask (run_mgfpj_v2.py:334-355), and the PMatrixFile drop-in path.

The P-matrices are those of a circular scan, built from first principles and
rounded to FP32 (as the reference stores them).  Tolerances: rel-L2 <= 1e-5
against the FP32 reference-order oracle with at most 0.1% of range-test
decisions on an FP32 boundary.
"""
import json

import numpy as np
import pytest
import torch

from oracle import recon
from reconstruction_program_taichi_amd import geometry as geo
from reconstruction_program_taichi_amd import pmatrix as pmx
from reconstruction_program_taichi_amd.device import Plan
from tests._cases import cone_case, oracle_dict, pmatrix_for, rel_l2

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _spec(rot=0.0, tilt=False):
    # generic pitches and distances: no sample or voxel sits exactly on a cell
    # boundary, where FP32 and the oracle's order could split a tie
    spec = cone_case(nx=32, nz=14, nv=30, nu=64, nrows=36, pix=0.49, vh=0.51, w=0.53, h=0.47,
                     sid=201.3, sdd=361.7, off_v=0.77)
    if rot:
        spec.img_rot = rot
        spec.angles = geo.angle_array(spec.n_views, spec.total_scan_angle, rot)
    pm = pmatrix_for(spec)
    if tilt:  # a detector slightly rotated in its plane: no standard geometry
        for j in range(spec.n_views):  # mixes u and v: P <- T P with T a small in-plane rotation
            P = pm[12 * j:12 * j + 12].reshape(3, 4)
            T = np.array([[np.cos(0.02), -np.sin(0.02), 0], [np.sin(0.02), np.cos(0.02), 0], [0, 0, 1]])
            pm[12 * j:12 * j + 12] = (T @ P).reshape(12)
    spec.pmatrix = pm.astype(np.float32)
    return spec


@pytest.mark.parametrize("case", ["plain", "rot", "tilt"])
@pytest.mark.parametrize("layout", ["quad", "rows", "plain"])
def test_bp_pmatrix_matches_oracle(ctx, case, layout):
    spec = _spec(rot=0.375 if case == "rot" else 0.0, tilt=case == "tilt")
    rng = np.random.default_rng(21)
    f = rng.standard_normal((spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    plan = Plan(ctx, spec)
    packed = ctx.pack(dev(f), "f32", layout)
    vol = torch.empty((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32, device=DEV)
    plan.backproject_fbp(packed, "f32", layout, vol)
    assert ctx.bp_last_kernel() == "generic"
    got = host(vol)
    g = dict(oracle_dict(spec), pmatrix=spec.pmatrix)
    ref, alive = recon.backproject_fbp(f, g, np.float32, return_mask=True)
    assert 0.2 < alive.mean() < 0.95
    assert (alive != (got != 0)).mean() <= 1e-3
    both = alive & (got != 0)
    assert rel_l2(got[both], ref[both]) <= 1e-5


@pytest.mark.parametrize("case", ["plain", "tilt"])
def test_bp_ir_pmatrix_matches_oracle(ctx, case):
    spec = _spec(tilt=case == "tilt")
    rng = np.random.default_rng(22)
    sgm = rng.standard_normal((spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    mask = (rng.uniform(0, 1, sgm.shape) > 0.2).astype(np.float32)
    plan = Plan(ctx, spec)
    vol = torch.empty((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32, device=DEV)
    plan.backproject_ir(dev(sgm), dev(mask), vol)
    got = host(vol)
    ref = recon.backproject_ir(sgm, dict(oracle_dict(spec), pmatrix=spec.pmatrix), ray_mask=mask,
                               dtype=np.float32)
    assert (ref != 0).mean() > 0.2
    assert ((got != 0) != (ref != 0)).mean() <= 1e-3
    assert rel_l2(got, ref) <= 1e-5


def _fp_spec(tilt=False, rot=0.0):
    spec = _spec(tilt=tilt, rot=rot)
    spec.fpj_step = 0.3
    A, xs = pmx.matrix_a_and_source(spec.pmatrix, spec.n_views)
    spec.matrix_a, spec.x_s = A, xs
    return spec


@pytest.mark.parametrize("kernel", ["cells", "direct"])
@pytest.mark.parametrize("case", ["plain", "tilt", "rot"])
def test_fp_pmatrix_matches_oracle(ctx, kernel, case):
    spec = _fp_spec(tilt=case == "tilt", rot=0.375 if case == "rot" else 0.0)
    vol = np.random.default_rng(23).uniform(0, 0.02, (spec.img_dim_z, spec.img_dim, spec.img_dim))
    plan = Plan(ctx, spec)
    out = torch.zeros((spec.n_v, spec.n_views, spec.n_u), dtype=torch.float32, device=DEV)
    ctx.set_fp_kernel(kernel)
    try:
        plan.forward_project(dev(vol), out, None, 1)
        assert ctx.fp_last_kernel() == kernel
    finally:
        ctx.set_fp_kernel("auto")
    got = host(out)
    g = dict(oracle_dict(spec), matrix_a=spec.matrix_a, x_s=spec.x_s)
    ref = recon.forward_project(vol.astype(np.float32), g, dtype=np.float32)
    assert (ref > 0).mean() > 0.3
    # A sample within FP32 rounding of the volume's edge is counted by one
    # order of operations and not by the other (the FP32 and FP64 oracles
    # disagree on 5 of these 69,120 rays the same way): such rays are
    # counted, the rest must agree to 1e-5.
    split = np.abs(got - ref) > 1e-5 * np.abs(ref).max()
    assert split.mean() <= 1e-3
    assert rel_l2(got[~split], ref[~split]) <= 1e-5


def test_ray_mask_pmatrix_matches_oracle(ctx):
    spec = _fp_spec(tilt=True)
    plan = Plan(ctx, spec)
    mask = torch.empty((spec.n_v, spec.n_views, spec.n_u), dtype=torch.float32, device=DEV)
    plan.ray_mask(mask)
    got = host(mask)
    ref = recon.ray_mask(dict(oracle_dict(spec), matrix_a=spec.matrix_a, x_s=spec.x_s))
    assert 0.05 < (ref == 0).mean() < 0.95
    assert (got != ref).mean() <= 1e-3


def test_fp_pmatrix_needs_matrix_a(ctx):
    spec = _spec()
    plan = Plan(ctx, spec)
    out = torch.zeros((spec.n_v, spec.n_views, spec.n_u), dtype=torch.float32, device=DEV)
    vol = torch.zeros((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32, device=DEV)
    with pytest.raises(RuntimeError, match="matrix_a"):
        plan.forward_project(vol, out, None, 1)


def test_run_mgfbp_pmatrix_file(tmp_path):
    """PMatrixFile through the drop-in: the reconstruction equals the oracle
    chain (weighting, ramp filter, P-matrix back-projection) on the
    normalised matrices and derived scan parameters."""
    from reconstruction_program_taichi_amd.fbp import Mgfbp, run_mgfbp

    nu, nv, views, w, h = 64, 24, 40, 0.5, 0.5
    ang = geo.angle_array(views, 2 * np.pi, 0.0)
    u = geo.det_pos_array(nu, nu, -w, -0.5, 0)
    v = geo.det_pos_array(nv, nv, h, 0.25, 0)
    scan = geo.GeometrySpec(img_dim=24, img_dim_z=10, pix=0.5, vox_h=0.5, n_views=views, sid=200.0,
                            sdd=360.0, total_scan_angle=2 * np.pi, angles=ang, u=u, v=v, cone=True,
                            n_v_full=nv, det_h=h)
    pm = pmatrix_for(scan).astype(np.float32)
    (tmp_path / "in").mkdir()
    (tmp_path / "pm.jsonc").write_text(json.dumps({"Value": [float(x) for x in pm]}))
    cfg = {"InputDir": str(tmp_path / "in"), "OutputDir": str(tmp_path / "out"), "InputFiles": "sgm_.*",
           "OutputFilePrefix": "", "OutputFileReplace": ["sgm_", "rec_"], "SinogramWidth": nu,
           "DetectorElementWidth": w, "DetectorOffcenter": 0.0, "SliceCount": nv, "Views": views,
           "ImageDimension": 24, "PixelSize": 0.5, "HammingFilter": 1.0, "ConeBeam": True,
           "DetectorElementHeight": h, "ImageSliceCount": 10, "VoxelHeight": 0.5, "ImageCenterZ": 0.0,
           "PMatrixFile": str(tmp_path / "pm.jsonc")}
    sgm = np.random.default_rng(24).uniform(0, 1, (nv, views, nu)).astype(np.float32)
    sgm.tofile(tmp_path / "in" / "sgm_x.raw")
    (tmp_path / "c.jsonc").write_text(json.dumps(cfg))
    out = run_mgfbp(str(tmp_path / "c.jsonc"))
    spec = Mgfbp(dict(cfg)).geometry_spec()
    assert spec.pmatrix is not None
    wgt = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle, False)
    taps = geo.hamming_taps(nu, w, 1.0, spec.sdd)
    fr = recon.filter_rows(wgt, taps, w, np.float32)
    ref, alive = recon.backproject_fbp(fr, dict(oracle_dict(spec), pmatrix=spec.pmatrix), np.float32,
                                       return_mask=True)
    assert alive.mean() > 0.2
    assert (alive != (out != 0)).mean() <= 1e-3
    both = alive & (out != 0)
    assert rel_l2(out[both], ref[both]) <= 1e-5
