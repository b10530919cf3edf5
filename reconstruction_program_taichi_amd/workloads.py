"""Timed synthetic workloads besides the cone FDK bench line.

  fbp2d: fan FBP 512 x 512 at 0.3 mm, 720 views over 360 deg, 800 elements at
         0.4 mm, SID 500 / SDD 1000 mm, Ram-Lak; 10 seeded random ellipses.
  fp3d:  cone forward projection of a 512 x 512 x 256 volume (0.3 mm voxels,
         30 seeded random ellipsoids) into 1080 views x 256 rows x 1024
         columns at 0.4 mm, SID 750 / SDD 1250 mm, step 0.2 of the voxel
         diagonal (optionally a leading subset of the views).
  fpfan: the IR loop's fan forward projection alone: 64 slices of 512 x 512
         into 120 views x 800 columns (fbp2d's geometry), ray mask applied.
  ir:    IRN-CGLS-TV on 64 fan slices of 512 x 512 (geometry of fbp2d with
         120 views), 5 reweighting passes x 20 CG iterations, Lambda 1e-3,
         BetaTV 1e-6, slices coupled through TV as the reference does.
  filter: the ramp filter alone on the rows of the cone FDK sinogram
         (256 x 1080 rows of 1024 samples, Hamming t = 1 taps at 0.4 mm),
         seeded uniform input, with --filter's method.
  pmatrix: the cone back-projection of the FDK benchmark's geometry (512 x 512
         x 256 at 0.3 mm, 1024 x 256 detector at 0.4 mm, SID 750 / SDD 1250 mm)
         given as per-view P-matrices (the P-matrix branch, generic kernel,
         FP32 QUAD taps) for the first --fp-views views; seeded uniform taps.
  helical: helical rebinning FBP (run_mgfbp_helical's path): 64 detector
         rows x 736 columns at 1.0 mm, SID 570 / SDD 1040 mm, 1152 views per
         turn over 3 turns, pitch 0.8, slices of 512 x 512 at 0.7 mm every
         0.5 mm; seeded uniform sinogram; rebinning, then weighting, ramp
         filter and the batched fan back-projection of every slice.
Each returns a dict (printed as one JSON line by bench.py).
"""
from __future__ import annotations

import time
from dataclasses import replace

import numpy as np
import torch

from . import phantom
from .device import Plan


def _fan_cfg(views: int, slices: int = 1) -> dict:
    return {"InputDir": "/nonexistent", "OutputDir": "/tmp/mct_workloads", "InputFiles": "x",
            "OutputFilePrefix": "", "OutputFileReplace": ["a", "b"],
            "DetectorElementCountHorizontal": 800, "DetectorElementWidth": 0.4,
            "DetectorOffsetHorizontal": 0.0, "DetectorElementCountVertical": slices, "Views": views,
            "TotalScanAngle": 360.0, "SourceIsocenterDistance": 500.0,
            "SourceDetectorDistance": 1000.0, "ImageDimension": 512, "PixelSize": 0.3,
            "HammingFilter": 1.0}


def _sync_time(fn, reps: int) -> float:
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def fbp2d(args) -> dict:
    from .fbp import Mgfbp

    dev = torch.device("cuda", 0)
    fbp = Mgfbp(_fan_cfg(720))
    spec = fbp.geometry_spec()
    r_fov = min(256 * 0.3, 400 * 0.4 * 500 / 1000) - 2 * 0.3
    ell = phantom.random_ellipsoids(10, seed=0, r_fov=r_fov)
    raw = phantom.project(ell, spec.angles, spec.u, spec.v, spec.sid, spec.sdd, device=dev)
    work = torch.empty_like(raw)
    vol = torch.empty((1, 512, 512), dtype=torch.float32, device=dev)

    def step():
        work.copy_(raw)
        fbp.reconstruct(work, vol)

    for _ in range(3):
        step()
    t = _sync_time(step, max(args.steps, 20))
    U = 512 * 512 * 720
    return {"workload": "fan FBP 512x512, 720 views, 800 elements (FP32, Ram-Lak)",
            "ms_per_step": round(t * 1e3, 4), "gups": round(U / t / 1e9, 2),
            "ramp_filter": fbp.filter_method}


def fp3d(args) -> dict:
    from .fpj import Mgfpj

    dev = torch.device("cuda", 0)
    nv = int(args.fp_views)
    cfg = {"InputDir": "/nonexistent", "OutputDir": "/tmp/mct_workloads", "InputFiles": "x",
           "OutputFileReplace": ["a", "b"], "ImageDimension": 512, "PixelSize": 0.3,
           "ImageDimensionZ": 256, "ConeBeam": True, "VoxelHeight": 0.3,
           "DetectorElementCountHorizontal": 1024, "DetectorElementWidth": 0.4,
           "DetectorElementCountVertical": 256, "DetectorElementHeight": 0.4, "Views": 1080,
           "SourceIsocenterDistance": 750.0, "SourceDetectorDistance": 1250.0}
    fpj = Mgfpj(cfg)
    spec = fpj.geometry_spec()
    if nv < spec.n_views:  # leading subset of the views (same per-view work)
        spec = replace(spec, n_views=nv, angles=spec.angles[:nv])
    ell = phantom.random_ellipsoids(30, seed=2, r_fov=70.0, z_half=30.0)
    vol = phantom.voxelize(ell, 512, 256, 0.3, 0.3, clip_r=76.0, supersample=1, device=dev)
    out = torch.empty((256, nv, 1024), dtype=torch.float32, device=dev)
    plan = Plan(fpj.ctx(), spec)
    fpj.ctx().set_fp_kernel(getattr(args, "fp_kernel", "auto"))
    plan.forward_project(vol, out)  # warm-up
    t = _sync_time(lambda: plan.forward_project(vol, out), max(1, min(args.steps, 3)))
    diag = float(np.sqrt(np.float32(2) * np.float32(0.3) ** 2 + np.float32(0.3) ** 2))
    dim = np.float32(512 * 0.3)
    D = float(np.sqrt(np.float32(2) * dim * dim + np.float32(256 * 0.3) ** 2))
    steps = int(np.floor(np.float32(D) / (np.float32(0.2) * np.float32(diag))))
    rays = 256 * nv * 1024
    S = rays * steps
    return {"workload": f"cone forward projection 512x512x256 -> {nv} views x 256 x 1024",
            "kernel": fpj.ctx().fp_last_kernel(),
            "ms": round(t * 1e3, 3), "rays": rays, "steps_per_ray": steps,
            "samples_per_s": round(S / t / 1e9, 2), "unit": "G samples/s (reference step count)"}


def fpfan(args) -> dict:
    """The IR loop's forward projection alone: 64 fan slices of 512 x 512
    (0.3 mm) -> 120 views x 800 columns (0.4 mm), SID 500 / SDD 1000, the
    ray mask applied, step 0.2 voxel."""
    from .ir import Mgfbp_ir

    dev = torch.device("cuda", 0)
    irr = Mgfbp_ir({**_fan_cfg(120, 64), "Lambda": 1e-3})
    spec = irr.ir_geometry_spec()
    plan = Plan(irr.ctx(), spec)
    x = torch.rand((64, 512, 512), generator=torch.Generator(device=dev).manual_seed(1), device=dev)
    out = torch.empty((64, 120, 800), dtype=torch.float32, device=dev)
    mask = torch.empty_like(out)
    plan.ray_mask(mask)
    irr.ctx().set_fp_kernel(getattr(args, "fp_kernel", "auto"))
    plan.forward_project(x, out, mask, 1)  # warm-up
    t = _sync_time(lambda: plan.forward_project(x, out, mask, 1), max(1, args.steps))
    return {"workload": "fan forward projection 64 x 512x512 -> 120 views x 800 (IR loop geometry)",
            "kernel": irr.ctx().fp_last_kernel(), "ms": round(t * 1e3, 3),
            "rays": 64 * 120 * 800, "checksum": float(out.double().sum())}


def ir(args) -> dict:
    from .ir import Mgfbp_ir

    dev = torch.device("cuda", 0)
    cfg = {**_fan_cfg(120, 64), "Lambda": 1e-3, "BetaTV": 1e-6, "NumberOfIRNIterations": 5,
           "NumberOfIterations": 20}
    irr = Mgfbp_ir(cfg)
    spec = irr.ir_geometry_spec()
    r_fov = min(256 * 0.3, 400 * 0.4 * 500 / 1000) - 2 * 0.3
    sg = torch.empty((64, 120, 800), dtype=torch.float32, device=dev)
    for s in range(64):  # an independent random phantom per slice
        ell = phantom.random_ellipsoids(10, seed=100 + s, r_fov=r_fov)
        sg[s:s + 1] = phantom.project(ell, spec.angles, spec.u, spec.v[:1], spec.sid, spec.sdd, device=dev)
    x0 = torch.zeros((64, 512, 512), dtype=torch.float32, device=dev)
    n_irn, n_it = irr.num_irn_iter, irr.num_iter
    irr.num_irn_iter, irr.num_iter = 0, 1
    irr.solve(sg, x0)  # warm-up (allocation, plan)
    irr.num_irn_iter, irr.num_iter = n_irn, n_it
    irr.num_iter_runned = 0
    irr.loss = np.empty(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    irr.solve(sg, x0)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    n_cg = (1 + irr.num_irn_iter) * irr.num_iter
    return {"workload": "IRN-CGLS-TV, 64 fan slices 512x512, 120 views, (1+5) passes x 20 CG iterations",
            "total_s": round(t, 3), "cg_iterations": n_cg, "ms_per_cg_iteration": round(t / n_cg * 1e3, 3),
            "fp_bp_pairs": irr.num_iter_runned, "loss_first_last": [float(irr.loss[0]), float(irr.loss[-1])]}


def filter_rows(args) -> dict:
    from . import geometry as geo
    from .device import Context

    dev = torch.device("cuda", 0)
    ctx = Context(0)
    rows, n = 256 * 1080, 1024
    gen = torch.Generator(device=dev).manual_seed(3)
    x = torch.rand((rows, n), generator=gen, device=dev)
    out = torch.empty_like(x)
    taps = geo.hamming_taps(n, 0.4, 1.0, 1250.0)
    method = args.filter

    def step():
        ctx.filter_rows(x, out, taps, 0.4, method)

    step()
    t = _sync_time(step, max(args.steps, 1))
    products = {"f32": 1, "bf16": 1, "bf16x3": 3}[method]
    flops = 2.0 * rows * n * n * products
    line = {"workload": f"ramp filter, {rows} rows x {n} samples", "method": method,
            "ms": round(t * 1e3, 4)}
    if method != "f32":
        line.update({"tflops": round(flops / t / 1e12, 1), "mfma_peak_frac": round(flops / t / 2.5e15, 4)})
    return line


def helical(args) -> dict:
    from .fbp_helical import Mgfbp_helical

    turns, per_turn = 3, 1152
    cfg = {"InputDir": "/nonexistent", "OutputDir": "/tmp/mct_workloads", "InputFiles": "x",
           "OutputFilePrefix": "", "OutputFileReplace": ["a", "b"], "SinogramWidth": 736,
           "DetectorElementWidth": 1.0, "DetectorOffcenter": 0.0, "SliceCount": 64,
           "Views": per_turn * turns, "TotalScanAngle": 360.0 * turns,
           "SourceIsocenterDistance": 570.0, "SourceDetectorDistance": 1040.0, "ImageDimension": 512,
           "PixelSize": 0.7, "HammingFilter": 1.0, "ConeBeam": True, "DetectorElementHeight": 1.0,
           "ImageSliceCount": 1, "VoxelHeight": 0.5, "SliceOffCenter": 0.0, "HelicalPitch": 0.8}
    f = Mgfbp_helical(cfg)
    rng = torch.Generator(device="cuda").manual_seed(5)
    sgm = torch.rand((f.det_elem_count_vertical, f.view_num, f.det_elem_count_horizontal),
                     generator=rng, device="cuda")
    reb = f.rebin(sgm)
    nz, vc, n = reb.shape
    out = torch.empty((nz, 512, 512), dtype=torch.float32, device="cuda")
    work = torch.empty_like(reb)

    def step():
        r = f.rebin(sgm)
        f.reconstruct_rebinned(r, out)

    def rebin_only():
        work.copy_(f.rebin(sgm))

    step()
    t = _sync_time(step, max(args.steps, 3))
    t_rebin = _sync_time(rebin_only, max(args.steps, 3))
    U = float(nz) * 512 * 512 * vc
    return {"workload": f"helical FBP: 64 rows x 736 columns, {per_turn} views/turn x {turns} turns, "
                        f"pitch 0.8 -> {nz} slices of 512x512 from {vc} rebinned views each",
            "ms_per_step": round(t * 1e3, 3), "rebin_ms": round(t_rebin * 1e3, 3),
            "gups": round(U / t / 1e9, 2), "ramp_filter": f.filter_method}


def circular_pmatrices(spec) -> np.ndarray:
    """FP64 P-matrices (12 per view) of spec's circular flat-detector scan:
    view j has its source at sid (cos b, sin b, 0), b = angle_j - img_rot,
    detector columns along (-sin b, cos b, 0) from u[0] in steps of du, rows
    along z from v[0] in steps of dv, at distance sdd; P = [A^-1 | -A^-1 S]
    with A = [du e, dv e_z, sdd n + u[0] e + v[0] e_z] maps a point to
    (u index, v index) after division by its third component."""
    u, v = np.asarray(spec.u, np.float64), np.asarray(spec.v, np.float64)
    du, dv = u[1] - u[0], v[1] - v[0]
    out = np.empty((spec.n_views, 12))
    for j, b in enumerate(np.asarray(spec.angles, np.float64) - spec.img_rot):
        c, s = np.cos(b), np.sin(b)
        e, n, ez = np.array([-s, c, 0.0]), -np.array([c, s, 0.0]), np.array([0.0, 0.0, 1.0])
        A = np.column_stack([du * e, dv * ez, spec.sdd * n + u[0] * e + v[0] * ez])
        Ai = np.linalg.inv(A)
        out[j] = np.hstack([Ai, -Ai @ (spec.sid * np.array([c, s, 0.0]))[:, None]]).reshape(12)
    return out.reshape(-1)


def pmatrix(args) -> dict:
    from .fbp import Mgfbp

    nv = int(args.fp_views)
    cfg = {"InputDir": "/nonexistent", "OutputDir": "/tmp/mct_workloads", "InputFiles": "x",
           "OutputFilePrefix": "", "OutputFileReplace": ["a", "b"],
           "DetectorElementCountHorizontal": 1024, "DetectorElementWidth": 0.4,
           "DetectorOffsetHorizontal": 0.0, "DetectorElementCountVertical": 256, "Views": 1080,
           "TotalScanAngle": 360.0, "SourceIsocenterDistance": 750.0, "SourceDetectorDistance": 1250.0,
           "ImageDimension": 512, "PixelSize": 0.3, "HammingFilter": 1.0, "ConeBeam": True,
           "DetectorElementHeight": 0.4, "ImageDimensionZ": 256, "VoxelHeight": 0.3,
           "DetectorOffsetVertical": 0.0, "ImageCenterZ": 0.0}
    fbp = Mgfbp(cfg)
    spec = fbp.geometry_spec()
    if nv < spec.n_views:  # leading subset of the views (same per-view work)
        spec = replace(spec, n_views=nv, angles=spec.angles[:nv])
    spec = replace(spec, pmatrix=circular_pmatrices(spec).astype(np.float32))
    ctx = fbp.ctx()
    plan = Plan(ctx, spec)
    g = torch.Generator(device="cuda").manual_seed(9)
    filt = torch.rand((spec.n_v, spec.n_views, spec.n_u), generator=g, device="cuda")
    packed = ctx.pack(filt, "f32", "quad")  # the layout the drop-in packs for the P-matrix path
    vol = torch.empty((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32, device="cuda")
    plan.backproject_fbp(packed, "f32", "quad", vol)
    kernel = ctx.bp_last_kernel()
    t = _sync_time(lambda: plan.backproject_fbp(packed, "f32", "quad", vol), max(1, min(args.steps, 3)))
    U = 512.0 * 512 * 256 * spec.n_views
    return {"workload": f"P-matrix cone back-projection 512x512x256 from {spec.n_views} views x 256 x 1024",
            "kernel": kernel, "ms": round(t * 1e3, 3), "gups": round(U / t / 1e9, 2)}


def run(name: str, args) -> dict:
    return {"fbp2d": fbp2d, "fp3d": fp3d, "fpfan": fpfan, "ir": ir, "filter": filter_rows,
            "helical": helical, "pmatrix": pmatrix}[name](args)
