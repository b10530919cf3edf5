"""Parity at the benchmark workloads' real sizes, through the C ABI, against
the oracle (the C restatements of oracle/cbp.py run the large cases on the
host cores in seconds; they equal the NumPy oracle, tests/test_oracle.py).

* C0: fan FBP 512 x 512 at 0.3 mm, 720 views, 800 elements at 0.4 mm, SID 500
  / SDD 1000, Ram-Lak; input = the oracle's FP64 bilinear forward projection
  (step 0.2) of 10 seeded ellipses (4 x 4 supersampled), plus the reference's
  noise model (AddPossionNoise: Gaussian-approximated Poisson, clamp at
  1e-6) at I0 = 1e5.  Bounds: rel-L2 <= 1e-5 against the FP32 reference-order
  chain with identical masks; <= 1e-3 against the FP64 chain.
* C1: cone FDK 512 x 512 x 256 at 0.3 mm from 1080 views x 256 x 1024 at 0.4 mm,
  SID 750 / SDD 1250, Shepp-Logan taps; 30 seeded ellipsoids.  The weighting
  and FP32 ramp filter on three detector rows, the back-projection on slabs
  of 38 slices around the top and bottom mask transitions, and the forward
  projection on 4 views x 64 rows: each <= 1e-5 rel-L2 against the FP32
  reference-order restatement, masks identical.
* C3: IRN-CGLS-TV on 3 coupled fan slices at C0's geometry with 120 views,
  Lambda 1e-3, BetaTV 1e-6, 1 IRN pass x 3 CG iterations; PICCS alpha = 0.5
  with a perturbed prior: <= 1e-4 against the FP64 oracle loop.
* C4: the full-size 36-job served mix (FP16 storage): every served result
  equals its plain one-job-at-a-time result bit for bit, and one job of each
  kind matches the oracle (FBP/FDK <= 1e-3 against the FP32 chain, the
  precision bound of FP16 storage with the split-BF16 filter, measured
  2e-5; CGLS <= 1e-4).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import cbp, recon
from reconstruction_program_taichi_amd import geometry as geo
from reconstruction_program_taichi_amd import phantom
from tests._cases import oracle_dict, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _fan_cfg(views, rows=1, **kw):
    c = {"InputDir": "/nonexistent", "OutputDir": "/tmp/mct_workload_tests", "InputFiles": "x",
         "OutputFilePrefix": "", "OutputFileReplace": ["a", "b"],
         "DetectorElementCountHorizontal": 800, "DetectorElementWidth": 0.4,
         "DetectorOffsetHorizontal": 0.0, "DetectorElementCountVertical": rows, "Views": views,
         "SourceIsocenterDistance": 500.0, "SourceDetectorDistance": 1000.0, "ImageDimension": 512,
         "PixelSize": 0.3, "HammingFilter": 1.0}
    c.update(kw)
    return c


R_FOV_C0 = min(256 * 0.3, 400 * 0.4 * 500 / 1000) - 2 * 0.3


# ----------------------------------------------------------------- C0
def test_c0_fan_fbp_of_noisy_fp64_projection():
    from reconstruction_program_taichi_amd.fbp import Mgfbp

    _say("C0: fan FBP 512^2 / 720 views / 800 elements from a noisy FP64 projection")
    fbp = Mgfbp(_fan_cfg(720))
    spec = fbp.geometry_spec()
    g = oracle_dict(spec)
    ell = phantom.random_ellipsoids(10, seed=0, r_fov=R_FOV_C0)
    img = phantom.voxelize(ell, 512, 1, 0.3, 0.0, clip_r=R_FOV_C0, supersample=4).cpu().numpy()
    g64 = dict(g, angles=spec.angles.astype(np.float64), u=spec.u.astype(np.float64),
               v=spec.v.astype(np.float64))
    clean = cbp.forward_project(img.astype(np.float64), g64, dtype=np.float64)
    assert clean.max() > 1.0
    sgm = recon.add_noise(clean.astype(np.float32), 1e5, seed=1)
    got = fbp.reconstruct(torch.from_numpy(sgm).cuda()).cpu().numpy()
    # FP32 reference-order chain
    w = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle, False)
    fl = recon.filter_rows(w, fbp.recon_kernel(), 0.4, np.float32)
    ref32 = cbp.backproject_fbp(fl, g)
    e32 = rel_l2(got, ref32)
    masks = int(((got == 0) != (ref32 == 0)).sum())
    # FP64 chain
    w64 = recon.weight_sinogram(sgm.astype(np.float64), g64["u"], g64["v"], g64["angles"], spec.sdd,
                                spec.total_scan_angle, False, dtype=np.float64)
    fl64 = recon.filter_rows(w64, recon.ramlak_taps(800, 0.4), 0.4, np.float64)
    ref64, alive = recon.backproject_fbp(fl64, g64, np.float64, return_mask=True)
    keep = alive & (got != 0)
    e64 = rel_l2(got[keep], ref64[keep])
    _say(f"C0: rel-L2 {e32:.2e} vs FP32 chain (mask disagreements {masks}), {e64:.2e} vs FP64 chain, "
         f"{int((alive != (got != 0)).sum())} FP32/FP64 range-test boundary voxels")
    assert masks == 0 and e32 <= 1e-5
    assert e64 <= 1e-3 and int((alive != (got != 0)).sum()) <= 1e-3 * got.size


# ----------------------------------------------------------------- C1
C1_CFG = {"InputDir": "/nonexistent", "OutputDir": "/tmp/mct_workload_tests", "InputFiles": "x",
          "OutputFilePrefix": "", "OutputFileReplace": ["a", "b"],
          "DetectorElementCountHorizontal": 1024, "DetectorElementWidth": 0.4,
          "DetectorOffsetHorizontal": 0.0, "DetectorElementCountVertical": 256, "Views": 1080,
          "TotalScanAngle": 360.0, "SourceIsocenterDistance": 750.0, "SourceDetectorDistance": 1250.0,
          "ImageDimension": 512, "PixelSize": 0.3, "SheppLoganFilter": 1, "ConeBeam": True,
          "DetectorElementHeight": 0.4, "ImageDimensionZ": 256, "VoxelHeight": 0.3,
          "DetectorOffsetVertical": 0.0}


@pytest.fixture(scope="module")
def c1():
    from reconstruction_program_taichi_amd.fbp import Mgfbp

    _say("C1: analytic projections of 30 ellipsoids, 1080 x 256 x 1024")
    fbp = Mgfbp(dict(C1_CFG))
    spec = fbp.geometry_spec()
    ell = phantom.random_ellipsoids(30, seed=2, r_fov=70.0, z_half=30.0)
    sgm = phantom.project(ell, spec.angles, spec.u, spec.v, spec.sid, spec.sdd, device="cuda")
    return fbp, spec, ell, sgm


def test_c1_weight_filter_and_backprojection_match_oracle(c1):
    fbp, spec, _, sgm = c1
    ctx, plan = fbp.ctx(), fbp.plan()
    raw_rows = sgm[[0, 100, 255]].cpu().numpy()
    w = sgm.clone()
    plan.weight(w)
    filt = torch.empty_like(w)
    taps = fbp.recon_kernel()
    assert fbp.kernel_name == "SheppLoganFilter"
    ctx.filter_rows(w, filt, taps, 0.4, "f32")
    packed = ctx.pack(filt, "f32", "rows")
    # poisoned: slices no column keeps (the first and last 16-slice chunks
    # here) are written by a fill, not by the back-projection launch
    vol = torch.full((256, 512, 512), float("nan"), dtype=torch.float32, device="cuda")
    plan.backproject_fbp(packed, "f32", "rows", vol)
    assert ctx.bp_last_kernel() == "lds"
    got = vol.cpu().numpy()
    f_host = filt.cpu().numpy()
    # weighting + filter, three detector rows, all views (FP32 reference order)
    rows = [0, 100, 255]
    wr = recon.weight_sinogram(raw_rows, spec.u, spec.v[rows], spec.angles, spec.sdd,
                               spec.total_scan_angle, False)
    fr = recon.filter_rows(wr, taps, 0.4, np.float32)
    ef = rel_l2(f_host[rows], fr)
    _say(f"C1 weighting + FP32 Shepp-Logan filter (3 rows x 1080 views): rel-L2 {ef:.2e}")
    assert ef <= 1e-5
    # back-projection on slabs around the mask transition: the detector
    # covers |z| < 30.72 mm at the centre and < 26.3 mm at the volume's
    # corners, so slices 25-40 and 215-230 are partly masked (a voxel is 0
    # once some view projects it off the detector rows)
    g = oracle_dict(spec)
    errs, partial = [], 0
    for z0, z1 in ((24, 40), (126, 130), (214, 232)):
        ref = cbp.backproject_fbp(f_host, g, z0, z1)
        sl = got[z0:z1]
        errs.append(rel_l2(sl, ref))
        assert np.array_equal(sl == 0, ref == 0), (z0, int(((sl == 0) != (ref == 0)).sum()))
        partial += sum(1 for k in range(z1 - z0) if 0 < (sl[k] == 0).sum() < sl[k].size)
    _say(f"C1 back-projection slabs: rel-L2 {[f'{e:.2e}' for e in errs]}, partly masked slices {partial}")
    assert max(errs) <= 1e-5 and partial >= 10
    # |z| >= 30.72 mm (slices <= 25 and >= 230) leaves the detector for every column
    assert (got[:24] == 0).all() and (got[232:] == 0).all()
    assert np.isfinite(got).all()


def test_c1_forward_projection_matches_oracle(c1):
    from reconstruction_program_taichi_amd.fpj import Mgfpj

    _, _, ell, _ = c1
    cfg = {"InputDir": "/nonexistent", "OutputDir": "/tmp/mct_workload_tests", "InputFiles": "x",
           "OutputFileReplace": ["a", "b"], "ImageDimension": 512, "PixelSize": 0.3,
           "ImageDimensionZ": 256, "ConeBeam": True, "VoxelHeight": 0.3,
           "DetectorElementCountHorizontal": 1024, "DetectorElementWidth": 0.4,
           "DetectorElementCountVertical": 256, "DetectorElementHeight": 0.4, "Views": 1080,
           "SourceIsocenterDistance": 750.0, "SourceDetectorDistance": 1250.0}
    fpj = Mgfpj(cfg)
    spec = fpj.geometry_spec()
    vol = phantom.voxelize(ell, 512, 256, 0.3, 0.3, clip_r=76.0, supersample=1)
    out = torch.empty((256, 1080, 1024), dtype=torch.float32, device="cuda")
    fpj.plan().forward_project(vol, out)
    views, rows = [0, 271, 540, 1001], list(range(0, 256, 4))
    got = out[rows][:, views].cpu().numpy()
    g = dict(oracle_dict(spec), angles=spec.angles[views], v=spec.v[rows])
    ref = cbp.forward_project(vol.cpu().numpy(), g, dtype=np.float32)
    e = rel_l2(got, ref)
    _say(f"C1 forward projection ({fpj.ctx().fp_last_kernel()}), 4 views x 64 rows x 1024: rel-L2 {e:.2e}")
    assert ref.max() > 1.0 and e <= 1e-5
    assert np.array_equal(got == 0, ref == 0)


# ----------------------------------------------------------------- C3
def _c3_inputs(n_slices=3):
    from reconstruction_program_taichi_amd.ir import Mgfbp_ir

    cfg = _fan_cfg(120, n_slices, Lambda=1e-3, BetaTV=1e-6, NumberOfIRNIterations=1,
                   NumberOfIterations=3)
    ir = Mgfbp_ir(cfg)
    spec = ir.ir_geometry_spec()
    sg = torch.empty((n_slices, 120, 800), dtype=torch.float32, device="cuda")
    for s in range(n_slices):
        ell = phantom.random_ellipsoids(10, seed=100 + s, r_fov=R_FOV_C0)
        sg[s:s + 1] = phantom.project(ell, spec.angles, spec.u, spec.v[:1], spec.sid, spec.sdd)
    return ir, spec, sg


def test_c3_irn_cgls_tv_matches_oracle_loop():
    _say("C3: IRN-CGLS-TV, 3 coupled slices of 512^2, 120 views x 800, 1 IRN x 3 CG")
    ir, spec, sg = _c3_inputs()
    x0 = torch.zeros((3, 512, 512), dtype=torch.float32, device="cuda")
    got = ir.solve(sg, x0).cpu().numpy()
    ref = recon.ir_solve(sg.cpu().numpy(), oracle_dict(spec), np.zeros((3, 512, 512)), ir.coef_lambda,
                         ir.beta_tv, 1, 3, couple_z=True,
                         projectors=(cbp.forward_project, cbp.backproject_ir))
    e = rel_l2(got, ref)
    _say(f"C3 IRN-CGLS-TV: rel-L2 vs FP64 oracle loop {e:.2e}")
    assert np.abs(ref).max() > 1e-3 and e <= 1e-4


def test_c3_full_size_projectors_match_oracle():
    """The projector pair of every CG iteration at C3's full size (64 slices
    of 512^2, 120 views x 800): the forward projection (64 slices: one ray
    per wavefront over the slices, each cell's steps summed as one term) on a
    subset of views, and the IR back-projection (slice-chunk records) of all
    views, against the C restatement of the oracle."""
    ir, spec, _ = _c3_inputs(n_slices=64)
    plan = ir.ir_plan()
    g = oracle_dict(spec)
    views = [0, 37, 71, 119]
    gv = dict(g, angles=spec.angles[views])
    fp = torch.empty((64, 120, 800), dtype=torch.float32, device="cuda")
    # a phantom (30 seeded ellipsoids, voxelised): against the FP32 oracle
    ell = phantom.random_ellipsoids(30, seed=12, r_fov=R_FOV_C0, z_half=8.0)
    x = phantom.voxelize(ell, 512, 64, 0.3, 0.3, clip_r=R_FOV_C0, supersample=1)
    plan.forward_project(x, fp)
    got = fp[:, views].cpu().numpy()
    ref = cbp.forward_project(x.cpu().numpy(), gv, dtype=np.float32)
    e = rel_l2(got, ref)
    _say(f"C3 forward projection ({ir.ctx().fp_last_kernel()}), 64 slices x 4 views x 800: rel-L2 {e:.2e}")
    assert ref.max() > 1.0 and e <= 1e-5 and np.array_equal(got == 0, ref == 0)
    # white noise: FP32 formulations differ by ~2e-5 here, so against FP64:
    # the kernel is no further from it than the FP32 oracle is
    gen = torch.Generator(device="cuda").manual_seed(11)
    x = torch.rand((64, 512, 512), generator=gen, device="cuda") * 0.02
    plan.forward_project(x, fp)
    got = fp[:, views].cpu().numpy()
    r32 = cbp.forward_project(x.cpu().numpy(), gv, dtype=np.float32)
    r64 = cbp.forward_project(x.cpu().numpy().astype(np.float64), gv, dtype=np.float64)
    e_got, e_o32 = rel_l2(got, r64), rel_l2(r32, r64)
    _say(f"C3 forward projection of white noise: rel-L2 vs FP64 {e_got:.2e} (FP32 oracle {e_o32:.2e})")
    assert e_got <= 1.1 * e_o32
    sgm = torch.rand((64, 120, 800), generator=gen, device="cuda")
    bp = torch.empty((64, 512, 512), dtype=torch.float32, device="cuda")
    plan.backproject_ir(sgm, None, bp)
    got = bp.cpu().numpy()
    ref = cbp.backproject_ir(sgm.cpu().numpy(), g, dtype=np.float32)
    e = rel_l2(got, ref)
    _say(f"C3 IR back-projection, 64 slices x 120 views: rel-L2 {e:.2e}, truncated voxels "
         f"{int((ref == 0).sum())}")
    assert e <= 1e-5 and np.array_equal(got == 0, ref == 0) and (ref == 0).any()


def test_c3_piccs_matches_oracle_loop(tmp_path):
    from reconstruction_program_taichi_amd.ir import Mgfbp_ir_piccs

    _say("C3: PICCS alpha 0.5, 3 slices of 512^2, 120 views x 800, 1 IRN x 3 CG")
    _, spec, sg = _c3_inputs()
    rng = np.random.default_rng(7)
    prior = np.zeros((3, 512, 512), np.float32)
    ys, xs = np.mgrid[0:512, 0:512]
    for k in range(3):  # two seeded ellipses per slice as the perturbed prior
        for _ in range(2):
            cx, cy, a, b = rng.uniform(150, 350), rng.uniform(150, 350), rng.uniform(20, 60), rng.uniform(20, 60)
            prior[k][((xs - cx) / a) ** 2 + ((ys - cy) / b) ** 2 <= 1] += rng.uniform(0.005, 0.02)
    prior.tofile(tmp_path / "prior.raw")
    cfg = _fan_cfg(120, 3, Lambda=1e-3, BetaTV=1e-6, NumberOfIRNIterations=1, NumberOfIterations=3,
                   Alpha=0.5, PriorImageFile=str(tmp_path / "prior.raw"))
    ir = Mgfbp_ir_piccs(cfg)
    got = ir.solve(sg, torch.from_numpy(prior).cuda()).cpu().numpy()
    ref = recon.ir_solve(sg.cpu().numpy(), oracle_dict(spec), prior, ir.coef_lambda, ir.beta_tv, 1, 3,
                         prior=prior, alpha=0.5, projectors=(cbp.forward_project, cbp.backproject_ir))
    e = rel_l2(got, ref)
    _say(f"C3 PICCS: rel-L2 vs FP64 oracle loop {e:.2e}")
    assert e <= 1e-4


# ----------------------------------------------------------------- C4
def _fbp_oracle_fp32(job):
    """FP32 reference-order chain of one FBP/FDK job (weight, FP64 ramp
    filter rounded to FP32, FP32 back-projection of the C restatement)."""
    s = job.spec
    sgm = job.sgm_host.numpy()
    w = recon.weight_sinogram(sgm, s.u, s.v, s.angles, s.sdd, s.total_scan_angle, s.short_scan)
    fl = recon.filter_rows(w.astype(np.float64), job.taps.astype(np.float64), job.det_w,
                           np.float64).astype(np.float32)
    g = oracle_dict(s)
    if s.cone:  # 8 central slices (the end slices are masked throughout)
        z = s.img_dim_z // 2
        return cbp.backproject_fbp(fl, g, z - 4, z + 4), slice(z - 4, z + 4)
    return cbp.backproject_fbp(fl, g), slice(None)


def test_c4_full_mix_served_equals_plain_and_oracle():
    from reconstruction_program_taichi_amd import serve

    _say("C4: full-size 36-job mix (fan FBP 256-1024, cone FDK 256/512, CGLS-TV 512), FP16 storage")
    jobs = serve.make_mix(36, qps=200.0, seed=3, storage="f16")
    plain = [j.fresh() for j in jobs]
    srv = serve.Server()
    try:
        st = srv.run(jobs)
    finally:
        srv.close()
    sp = serve.run_plain(plain)
    _say(f"C4 served {st['jobs_per_s']:.1f} jobs/s, plain {sp['jobs_per_s']:.1f} jobs/s")
    assert st["rejected"] == 0
    for a, b in zip(jobs, plain):
        assert torch.equal(a.result, b.result), f"job {a.jid} ({a.kind}) differs from its plain run"
    for kind in ("fbp", "fdk"):
        job = min((j for j in jobs if j.kind == kind), key=lambda j: j.nbytes())
        ref, sl = _fbp_oracle_fp32(job)
        got = job.result.numpy()[sl]
        keep = (got != 0) & (ref != 0)
        assert keep.sum() > 0.25 * got.size
        e = rel_l2(got[keep], ref[keep])
        _say(f"C4 {kind} job {job.jid} ({job.spec.img_dim}^2 x {job.spec.img_dim_z}, {job.spec.n_views} views): "
             f"rel-L2 {e:.2e} vs FP32 chain, mask disagreements {int(((got == 0) != (ref == 0)).sum())}")
        assert e <= 1e-3 and int(((got == 0) != (ref == 0)).sum()) <= 1e-3 * got.size
    job = next(j for j in jobs if j.kind == "cgls")
    s = job.spec
    ref = recon.ir_solve(job.sgm_host.numpy(), oracle_dict(s), np.zeros(job.vol_shape()), job.lam,
                         job.beta_tv, 0, job.iters, projectors=(cbp.forward_project, cbp.backproject_ir))
    e = rel_l2(job.result.numpy(), ref)
    _say(f"C4 cgls job {job.jid} ({s.img_dim}^2, {s.n_views} views, {job.iters} iterations): rel-L2 {e:.2e}")
    assert e <= 1e-4
