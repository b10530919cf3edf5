"""Benchmark: cone-beam FDK, 512x512x256 volume from 1080 views of a
1024 x 256 flat detector (FP32 compute), on one MI355X.

One step = one pass of the FDK hot path over the resident sinogram:
cosine weighting, ramp filter (split-BF16 MFMA by default; the FP32 filter's
step is timed beside it), repack into the back-projector's tap layout,
pixel-driven back-projection.  value = GUPS = voxel-view updates of the step
(512*512*256*1080) / step time.

roofline: the back-projection kernel's, timed with HIP events on the stream
it runs on.  Its algorithmic bytes per launch are the compulsory ones: the
filtered sinogram read once (4 B per detector sample, 2 B with FP16/BF16
storage) and the volume written once (4 B per voxel), against HBM's 8 TB/s.
The kernel is far from HBM's bound; VALU issue is its largest single cost.
The line carries, from the counter record of the same kernel sources (a JSON
block in a profiles/*.md summary whose source hash matches; null otherwise),
the measured HBM traffic and the VALU issue rate against two ceilings: two
wave64 instructions per CU per clock (gfx950 issues two VALU instructions in
one SIMD cycle when both qualify), and the rate the issue-rate probe
(issue_rate_probe) measures for the back-projector's own slice-loop
instruction mix, of which 3 in 8 (convert, fract, address multiply-add)
issue one at a time.  The per-update gather figure is context.

cpu_baseline: the oracle's C/OpenMP restatement of the same back-projection
on a slab of slices (all views, all columns) on all host cores, plus a
1-core run and the weighting, filter, forward-projection and IR
back-projection restatements on samples of their workloads.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--storage f32]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0
# Device memory the timed steps' private input copies may take.
RESIDENT_INPUT_BYTES = 32 << 30
# Per-CU fill rates: vector L1 64 B/clk; LDS 256 B/clk for 128-bit reads on
# CDNA4 (twice CDNA3's 128 B/clk: a read-only stream of ds_read_b128 measured
# 131 TB/s on MI355X, above the 78.6 TB/s that 128 B/clk allows).  VALU: each
# SIMD issues one wave64 instruction per 4-clock cycle, or two when both are
# FP32 add/mul/FMA or 32-bit add/and with VGPR, inline-constant or literal
# operands (issue_rate_probe: 1.6 per cycle for such streams, 0.95 for
# converts, fract, 24-bit multiply-adds, min and any SGPR operand); that is 1
# (single) or 2 (dual) wave64 instructions per CU per clock.  MI355X peak
# engine clock 2.4 GHz.
GATHER_BYTES_PER_CLK_PER_CU = {"direct": 64, "lds": 256}
GATHER_PATH = {"direct": "vector L1", "lds": "LDS"}
KERNEL_NAME = {"direct": "bp_cone_axial_kernel", "lds": "bp_cone_lds_kernel"}
CLOCK_GHZ = 2.4
VALU_DUAL_ISSUE = 2  # wave64 VALU instructions per CU per clock, both issue slots used
PROBE_EXE = ROOT / "reconstruction_program_taichi_amd" / "issue_rate_probe"


def gather_roofline(kernel: str, gather_bytes: float, ms: float) -> dict:
    """Every update reads its four taps (16 B FP32, 8 B FP16/BF16) through the
    vector L1 (direct kernel) or from LDS (staged kernel): that path's
    ceiling (context, not the bound)."""
    import torch

    if kernel not in GATHER_PATH:
        return None
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    peak = cus * GATHER_BYTES_PER_CLK_PER_CU[kernel] * CLOCK_GHZ  # GB/s
    ach = gather_bytes / (ms * 1e-3) / 1e9
    return {"path": GATHER_PATH[kernel], "achieved": round(ach, 1), "peak": round(peak, 1),
            "unit": "GB/s", "frac": round(ach / peak, 4), "cus": cus}


def counter_record(kernel: str, workload: str, source: str):
    """The counter record of `kernel` on `workload` (a ```json block of a
    profiles/*.md summary) whose source hash equals that of the sources at
    hand; (record, where) or (None, why not)."""
    from reconstruction_program_taichi_amd import _build

    want = _build.kernel_source_hash(source)
    why = f"no counter record of {kernel} on this workload"
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for blk in re.findall(r"```json\n(.*?)\n```", f.read_text(), re.S):
            try:
                rec = json.loads(blk)
            except ValueError:
                continue
            if rec.get("counters_of") != kernel or rec.get("workload") != workload:
                continue
            if rec.get("source_hash") == want:
                return rec, f"profiles/{f.name}"
            why = f"profiles/{f.name} holds counters of other sources ({rec.get('source_hash')} != {want})"
    return None, why


def issue_probe_rates() -> dict | None:
    """Wave64 VALU instructions per CU per ns that issue_rate_probe measures
    for each instruction form and mix (None when the probe is not built)."""
    import subprocess

    if not PROBE_EXE.exists():
        return None
    r = subprocess.run([str(PROBE_EXE)], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        return None
    rows = [json.loads(x) for x in r.stdout.splitlines() if x.startswith("{")]
    return {d["probe"]: d["wave_insts_per_cu_per_ns"] for d in rows}


def measured_bounds(rec: dict, where: str, ms: float, cus: int, probe: dict | None) -> dict:
    """HBM traffic (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a plain
    copy's readings) and VALU issue of one launch from the counter record,
    over this run's launch time (ms) on cus compute units; probe: the
    issue-rate probe's rates, measured on this GPU."""
    c, cal = rec["per_dispatch"], rec["calibration"]
    traffic = (c["FETCH_SIZE_KiB"] * 1024.0 * cal["fetch_factor"]
               + c["WRITE_SIZE_KiB"] * 1024.0 * cal["write_factor"])
    t = ms * 1e-3
    insts = c["SQ_INSTS_VALU"]
    rate = insts / cus / (t * 1e9)  # wave64 VALU instructions per CU per ns
    out = {"profile": where, "source_hash": rec["source_hash"], "hbm_bytes": traffic,
           "hbm_frac": round(traffic / t / (HBM_PEAK_GBS * 1e9), 4),
           "valu_insts": insts, "valu_insts_per_cu_per_ns": round(rate, 4),
           # against two instructions per CU per clock at the peak clock
           "valu_issue_frac": round(rate / (VALU_DUAL_ISSUE * CLOCK_GHZ), 4),
           # against one per CU per clock (every instruction issued alone)
           "valu_single_issue_frac": round(rate / CLOCK_GHZ, 4)}
    if c.get("GRBM_GUI_ACTIVE"):
        # over the cycles the GPU was busy with the dispatch (clock as run)
        out["valu_issue_frac_busy_cycles"] = round(insts / (VALU_DUAL_ISSUE * cus * c["GRBM_GUI_ACTIVE"]), 4)
        if c.get("SQ_ACTIVE_INST_VALU2"):
            # SIMD cycles (4 clocks) that issued two VALU instructions
            out["valu_dual_issue_cycles_frac"] = round(c["SQ_ACTIVE_INST_VALU2"] / (cus * c["GRBM_GUI_ACTIVE"]), 4)
    if probe and probe.get("bp_slice_mix"):
        out["slice_mix_rate_per_cu_per_ns"] = probe["bp_slice_mix"]
        out["valu_frac_of_slice_mix_rate"] = round(rate / probe["bp_slice_mix"], 4)
    return out


def c1_config(storage: str, filter_method: str) -> dict:
    return {
        "InputDir": "/nonexistent", "OutputDir": os.environ.get("TMPDIR", "/tmp") + "/mct_bench",
        "InputFiles": "x", "OutputFilePrefix": "", "OutputFileReplace": ["a", "b"],
        "DetectorElementCountHorizontal": 1024, "DetectorElementWidth": 0.4,
        "DetectorOffsetHorizontal": 0.0, "DetectorElementCountVertical": 256,
        "Views": 1080, "TotalScanAngle": 360.0,
        "SourceIsocenterDistance": 750.0, "SourceDetectorDistance": 1250.0,
        "ImageDimension": 512, "PixelSize": 0.3, "SheppLoganFilter": True,
        "ConeBeam": True, "DetectorElementHeight": 0.4, "ImageDimensionZ": 256,
        "VoxelHeight": 0.3, "DetectorOffsetVertical": 0.0, "ImageCenterZ": 0.0,
        "SinogramStorage": storage, "RampFilterMethod": filter_method,
    }


def counter_workload(args) -> str:
    """The key of this run's back-projection in the counter records."""
    return (f"cone FDK 512x512x256, 1080 views, 1024x256; {args.storage} taps; "
            f"tile {args.bp_tile}, batch {args.bp_batch}")


def fp32_filter_step(args, raw, vol, steps: int) -> dict:
    """The same FDK step with the FP32 direct ramp filter (C1's arithmetic),
    timed beside the headline's split-BF16 filter on the same inputs."""
    import torch

    from reconstruction_program_taichi_amd.fbp import Mgfbp

    f = Mgfbp(c1_config(args.storage, "f32"))
    f.ctx().set_bp_kernel(args.bp_kernel)
    if args.bp_tile or args.bp_batch:
        f.ctx().set_bp_tile(args.bp_tile, args.bp_batch)
    out = torch.empty_like(vol)
    work = raw.clone()
    f.reconstruct(work, out)  # plan tables, taps
    inputs = [raw.clone() for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        f.reconstruct(inputs[i], out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    del inputs, work, out
    nx, nz, nv = vol.shape[2], vol.shape[0], raw.shape[1]
    f.plan().close()
    f.ctx().close()
    return {"ms_per_step": round(ms, 3), "gups": round(float(nx) * nx * nz * nv / (ms * 1e-3) / 1e9, 2),
            "steps": steps}


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(fbp, filt_host: np.ndarray, raw_rows: np.ndarray, vol_host: np.ndarray,
                 nslices: int) -> dict:
    """Reference-algorithm CPU restatements (the oracle's C/OpenMP code and
    numpy), timed on samples of the workloads: the headline back-projection
    on nslices central slices with every host thread and on 2 slices with
    one, and the weighting, FP32 ramp filter, cone forward projection and fan
    IR back-projection legs, each with its full-workload extrapolation."""
    sys.path.insert(0, str(ROOT))
    from oracle import cbp, recon
    from reconstruction_program_taichi_amd import geometry as geo

    spec = fbp.geometry_spec()
    g = dict(img_dim=spec.img_dim, img_dim_z=spec.img_dim_z, pix=spec.pix, vox_h=spec.vox_h,
             sid=spec.sid, sdd=spec.sdd, angles=spec.angles, u=spec.u, v=spec.v, det_h=spec.det_h,
             cone=True, cz=spec.cz, total_scan_angle=spec.total_scan_angle, short_scan=False)
    nx, nz, nv, nrows, nu = spec.img_dim, spec.img_dim_z, spec.n_views, spec.n_v, spec.n_u
    cbp.lib()  # build/load outside the timed regions
    threads = cbp.max_threads()

    def timed(fn):
        t0 = time.perf_counter()
        r = fn()
        return r, time.perf_counter() - t0

    # back-projection, all threads (the headline CPU figure)
    z0 = nz // 2 - nslices // 2
    out, dt = timed(lambda: cbp.backproject_fbp(filt_host, g, z0, z0 + nslices))
    upd = float(nx) ** 2 * nslices * nv
    res = {"value": upd / dt / 1e9, "unit": "GUPS", "cores": threads, "kind": "port",
           "sample": f"FP32 back-projection of {nslices} central slices x 512x512 columns x "
                     f"1080 views ({upd:.3g} updates) in {dt:.2f} s; oracle/bp_oracle.c, OpenMP "
                     f"({threads} threads)",
           "cpu_model": cpu_model(), "label": "reference-algorithm CPU restatement",
           "_check": out}
    legs = {}
    # the same on one core
    cbp.set_threads(1)
    try:
        _, dt1 = timed(lambda: cbp.backproject_fbp(filt_host, g, nz // 2, nz // 2 + 1))
    finally:
        cbp.set_threads(threads)
    legs["backprojection_1core"] = {"gups": float(nx) ** 2 * nv / dt1 / 1e9, "cores": 1,
                                    "sample": f"the central slice in {dt1:.2f} s"}
    # weighting (numpy, one core), 16 detector rows
    rows = raw_rows.shape[0]
    _, dtw = timed(lambda: recon.weight_sinogram(raw_rows, spec.u, spec.v[:rows], spec.angles, spec.sdd,
                                                 spec.total_scan_angle, False))
    legs["weighting"] = {"samples_per_s": raw_rows.size / dtw, "cores": 1,
                         "sample": f"{rows} detector rows x {nv} views x {nu} in {dtw:.2f} s (numpy)",
                         "full_step_s": round(dtw * nrows / rows, 2)}
    # FP32 ramp filter (O(n^2) per row, all threads), the same 16 detector rows
    taps = fbp.recon_kernel()
    du = float(abs(spec.u[1] - spec.u[0]))
    _, dtf = timed(lambda: cbp.filter_rows(raw_rows, taps, du))
    legs["ramp_filter"] = {"rows_per_s": rows * nv / dtf, "cores": threads,
                           "sample": f"{rows * nv} rows of {nu} in {dtf:.2f} s (oracle C, OpenMP)",
                           "full_step_s": round(dtf * nrows / rows, 2)}
    legs["fdk_step_estimate_s"] = round(legs["weighting"]["full_step_s"] + legs["ramp_filter"]["full_step_s"]
                                        + dt * nz / nslices, 1)
    # cone forward projection of the FDK volume, 2 of the 1080 views
    k = 2
    gfp = dict(g, angles=spec.angles[:k], fpj_step=0.2)
    _, dtp = timed(lambda: cbp.forward_project(vol_host, gfp))
    legs["forward_projection"] = {"rays_per_s": k * nrows * nu / dtp, "cores": threads,
                                  "sample": f"{k} views x {nrows} x {nu} rays, step 0.2 voxel, in {dtp:.2f} s "
                                            "(oracle C, OpenMP)",
                                  "full_1080_views_s": round(dtp * nv / k, 1)}
    # fan IR back-projection, C3 geometry (512x512, 120 views x 800), 16 of the 64 slices
    ang = geo.angle_array(120, 2 * geo.PI, 0.0)
    gir = dict(img_dim=512, img_dim_z=16, pix=0.3, vox_h=0.0, sid=500.0, sdd=1000.0, angles=ang,
               u=geo.det_pos_array(800, 800, -0.4, 0.0), v=geo.det_pos_array(16, 16, 0, 0),
               cone=False, det_h=1.0)
    sg = np.ones((16, 120, 800), np.float32)
    _, dti = timed(lambda: cbp.backproject_ir(sg, gir))
    legs["ir_backprojection"] = {"gups": 512.0 * 512 * 16 * 120 / dti / 1e9, "cores": threads,
                                 "sample": f"16 slices x 512x512 x 120 views in {dti:.2f} s (oracle C, OpenMP)",
                                 "full_64_slices_s": round(dti * 4, 2)}
    res["legs"] = legs
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"])
    # bf16x3: split-BF16 Toeplitz GEMM on MFMA, 2.5e-6 rel-L2 vs FP64 (FP32-grade)
    ap.add_argument("--filter", default="bf16x3", choices=["f32", "bf16", "bf16x3"])
    ap.add_argument("--cpu-slices", type=int, default=32)
    ap.add_argument("--fp32-filter-steps", type=int, default=3,
                    help="timed FDK steps with the FP32 ramp filter beside the headline (0: skip)")
    ap.add_argument("--serve-jobs", type=int, default=36,
                    help="jobs in the served-stream comparison (0: skip it)")
    ap.add_argument("--serve-qps", type=float, default=200.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the timed steps (no FP32-filter, PCIe-inclusive or streamed legs): "
                         "every back-projection launch is then a whole-scan one (profiler runs)")
    ap.add_argument("--bp-tile", type=int, default=0, help="slices per thread (0: library default)")
    ap.add_argument("--bp-batch", type=int, default=0, help="loads per batch (0: library default)")
    ap.add_argument("--bp-kernel", default="auto", choices=["auto", "direct", "lds"],
                    help="cone back-projector: LDS-staged windows or per-update gathers")
    ap.add_argument("--autotune", action="store_true",
                    help="time the back-projector candidates during warm-up and keep the fastest")
    ap.add_argument("--sinogram", default="phantom", choices=["phantom", "random"],
                    help="analytic ellipsoid projections, or seeded uniform noise (profiler runs)")
    ap.add_argument("--workload", default="fdk", choices=["fdk", "fbp2d", "fp3d", "fpfan", "ir", "filter", "helical", "pmatrix"],
                    help="fdk: the bench line; the others print their own measurement")
    ap.add_argument("--fp-views", type=int, default=1080, help="fp3d: views projected")
    ap.add_argument("--fp-kernel", default="auto", choices=["auto", "direct", "cells"],
                    help="fp3d: forward projector")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only copy a buffer of this many MiB 5 times (profiler counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration_mib:
        import torch

        n = args.copy_calibration_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"copied {args.copy_calibration_mib} MiB x 5")
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the reconstruction path runs on a single device)", file=sys.stderr)
        return 2
    if args.workload != "fdk":
        from reconstruction_program_taichi_amd import workloads

        print(json.dumps(workloads.run(args.workload, args)), flush=True)
        return 0

    import torch

    from reconstruction_program_taichi_amd import phantom
    from reconstruction_program_taichi_amd.fbp import Mgfbp

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    fbp = Mgfbp(c1_config(args.storage, args.filter))
    spec = fbp.geometry_spec()
    nx, nz, nv = spec.img_dim, spec.img_dim_z, spec.n_views
    if args.sinogram == "phantom":
        ell = phantom.random_ellipsoids(30, seed=2, r_fov=70.0, z_half=30.0)
        raw = phantom.project(ell, spec.angles, spec.u, spec.v, spec.sid, spec.sdd, device=dev)
    else:  # same work, no setup kernels: for profiler runs
        gen = torch.Generator(device=dev).manual_seed(2)
        raw = torch.rand((spec.n_v, nv, spec.n_u), generator=gen, device=dev)
    work = torch.empty_like(raw)
    vol = torch.empty((nz, nx, nx), dtype=torch.float32, device=dev)
    plan = fbp.plan()
    if args.bp_tile or args.bp_batch:
        fbp.ctx().set_bp_tile(args.bp_tile, args.bp_batch)
    fbp.ctx().set_bp_kernel(args.bp_kernel)
    fbp.autotune = args.autotune  # tunes on the first (warm-up) step, outside the timed region
    stream = torch.cuda.current_stream(dev)

    bp_times = []
    orig_bp = plan.backproject_fbp

    def timed_bp(filt, dtype, layout, out):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        orig_bp(filt, dtype, layout, out)
        e1.record(stream)
        bp_times.append((e0, e1))
        return out

    plan.backproject_fbp = timed_bp

    # The MFMA filters weigh inside their input pass and leave the sinogram as
    # it was, so every timed step reads the one resident input.  The FP32
    # filter weighs in place, so each of its steps gets a resident copy of its
    # own (made before the timed region; past RESIDENT_INPUT_BYTES a device
    # copy of the input inside each step).
    untouched = fbp.filter_method != "f32" and not fbp.apply_gauss_vertical
    per_step = untouched or args.steps * raw.numel() * raw.element_size() <= RESIDENT_INPUT_BYTES
    inputs = [raw.clone() for _ in range(args.steps)] if per_step and not untouched else []

    def step(i):
        if untouched:
            fbp.reconstruct(raw, vol)
        elif per_step:
            fbp.reconstruct(inputs[i], vol)
        else:
            work.copy_(raw)
            fbp.reconstruct(work, vol)

    for _ in range(args.warmup):
        work.copy_(raw)
        fbp.reconstruct(work, vol)
    torch.cuda.synchronize()
    bp_times.clear()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    del inputs
    ms_step = dt / args.steps * 1e3
    bp_ms = float(np.mean([a.elapsed_time(b) for a, b in bp_times]))
    U = float(nx) * nx * nz * nv
    gups = U / (ms_step * 1e-3) / 1e9
    tap_bytes = 4 if args.storage == "f32" else 2
    # compulsory bytes of one back-projection: the filtered sinogram once, the volume once
    bytes_alg = float(spec.n_v) * nv * spec.n_u * tap_bytes + float(nx) * nx * nz * 4
    achieved = bytes_alg / (bp_ms * 1e-3) / 1e9
    kernel = fbp.ctx().bp_last_kernel()
    kname = KERNEL_NAME.get(kernel, kernel)
    rec, where = counter_record(kname, counter_workload(args), "mct_bp.hip")
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    measured = (measured_bounds(rec, where, bp_ms, cus, issue_probe_rates()) if rec
                else {"profile": None, "why": where})
    line = {
        "metric": "FDK/backprojection GUPS", "value": round(gups, 2), "unit": "GUPS",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": ("synthetic: analytic projections of 30 seeded random ellipsoids (seed 2)"
                 if args.sinogram == "phantom" else "synthetic: seeded uniform random sinogram"),
        "config": {"workload": "cone FDK 512x512x256 volume, 1080 views, 1024x256 detector",
                   "inputs": ("one resident sinogram, read unweighted by every step" if untouched
                              else "one resident sinogram per timed step" if per_step
                              else "sinogram restored by a device copy inside each step"),
                   "sinogram_storage": args.storage, "ramp_filter": args.filter,
                   "bp_ms": round(bp_ms, 3), "bp_gups": round(U / (bp_ms * 1e-3) / 1e9, 2),
                   **({"autotune": {"choice": fbp.tuned, "candidates_ms": fbp.tuner.timings}}
                      if fbp.tuned else {})},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured.get("hbm_bytes"), "kernel": kname,
                     "model": (f"compulsory bytes: the filtered sinogram once ({tap_bytes} B per "
                               "detector sample) + the volume once (4 B per voxel)"),
                     "binding": {"resource": "VALU issue", "frac": measured.get("valu_issue_frac"),
                                 "peak": "2 wave64 VALU instructions per CU per clock at 2.4 GHz",
                                 "frac_busy_cycles": measured.get("valu_issue_frac_busy_cycles"),
                                 "frac_single_issue": measured.get("valu_single_issue_frac"),
                                 "frac_of_slice_mix_rate": measured.get("valu_frac_of_slice_mix_rate")},
                     "measured": measured,
                     "gather": gather_roofline(kernel, U * 4 * tap_bytes, bp_ms)},
    }
    if args.filter != "f32" and args.fp32_filter_steps > 0 and not args.no_extras:
        line["config"]["fp32_filter_step"] = fp32_filter_step(args, raw, vol, args.fp32_filter_steps)
    raw_h = torch.empty(raw.shape, dtype=raw.dtype, pin_memory=True)
    raw_h.copy_(raw)
    if not args.no_extras:
        line["config"]["pcie_inclusive"] = pcie_inclusive(raw_h, work, vol, fbp)
        line["config"]["streamed"] = streamed_inclusive(raw_h, vol, args.storage, args.filter)
    if not args.no_cpu:
        # the FP32 filtered sinogram of the step (the headline step writes the
        # back-projector's taps straight from the filter, the same values)
        filt_host = fbp.weight_and_filter(raw.clone()).cpu().numpy()
        cb = cpu_baseline(fbp, filt_host, raw[:16].cpu().numpy(), vol.cpu().numpy(), args.cpu_slices)
        check = cb.pop("_check")
        z0 = nz // 2 - args.cpu_slices // 2
        gpu_slab = vol[z0:z0 + args.cpu_slices].cpu().numpy()
        if args.storage == "f32":
            denom = float(np.linalg.norm(check)) or 1.0
            cb["rel_l2_vs_gpu"] = float(np.linalg.norm(gpu_slab - check) / denom)
        line["cpu_baseline"] = cb
    bp_times.clear()
    plan.backproject_fbp = orig_bp
    fbp.last_filtered = None
    del work, vol, raw
    fbp.plan().close()
    fbp.ctx().close()
    torch.cuda.synchronize()
    if args.serve_jobs > 0:
        torch.cuda.empty_cache()
        line["served"] = served_stream(args.serve_jobs, args.serve_qps)
    print(json.dumps(line), flush=True)
    return 0


def pcie_inclusive(raw_h, work, vol, fbp, reps: int = 3) -> dict:
    """The same FDK step with the sinogram coming from pinned host memory
    (raw_h) and the volume going back to it (the rate a file-driven caller
    sees before disk I/O); reported beside the resident-input value, never
    as it."""
    import torch

    vol_h = torch.empty(vol.shape, dtype=vol.dtype, pin_memory=True)
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        work.copy_(raw_h, non_blocking=True)
        fbp.reconstruct(work, vol)
        vol_h.copy_(vol, non_blocking=True)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ms = float(np.mean(times)) * 1e3
    nx, nz, nv = vol.shape[2], vol.shape[0], raw_h.shape[1]
    gb = (raw_h.numel() + vol.numel()) * 4 / 1e9
    return {"ms_per_step": round(ms, 3), "gups": round(float(nx) * nx * nz * nv / (ms * 1e-3) / 1e9, 2),
            "host_bytes_gb": round(gb, 3)}


def streamed_inclusive(raw_h, vol, storage: str, filter_method: str, reps: int = 3) -> dict:
    """The view-block streaming FBP (run_mgfbp_v2 drop-in, default block
    size) on the same pinned host sinogram: the upload of block b+1 overlaps
    the kernels of block b; the volume is copied back to pinned host memory.
    vol holds the resident path's volume; the streamed one must equal it."""
    import torch

    from reconstruction_program_taichi_amd.fbp_v2 import Mgfbp_v2

    f = Mgfbp_v2(c1_config(storage, filter_method))
    out = torch.empty_like(vol)
    vol_h = torch.empty(vol.shape, dtype=vol.dtype, pin_memory=True)
    f.reconstruct_streamed(raw_h, out)  # plan tables
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        f.reconstruct_streamed(raw_h, out)
        vol_h.copy_(out, non_blocking=True)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ms = float(np.mean(times)) * 1e3
    nx, nz, nv = vol.shape[2], vol.shape[0], raw_h.shape[1]
    return {"views_per_block": f.views_per_block, "ms_per_step": round(ms, 3),
            "gups": round(float(nx) * nx * nz * nv / (ms * 1e-3) / 1e9, 2),
            "equals_resident": bool(torch.equal(out, vol))}


def served_stream(n_jobs: int, qps: float) -> dict:
    """Mixed FBP / FDK / CGLS-TV stream (Poisson arrivals, seed 3, FP16
    sinogram storage): continuous-batching server vs one job at a time.

    served: the stream at the offered QPS (latency under that load; its
    throughput cannot exceed the arrival rate).  served_saturated: the same
    jobs all arriving at once (the server's capacity).  plain: one job at a
    time in arrival order, with its time split into plan setup, host-to-device
    copy, compute and device-to-host copy; compute_bound_jobs_per_s is what a
    server that hid everything but plain's compute would reach."""
    from dataclasses import replace

    from reconstruction_program_taichi_amd import serve

    jobs = serve.make_mix(n_jobs, qps, seed=3, storage="f16")
    plain = [j.fresh() for j in jobs]
    warm = [j.fresh() for j in jobs[:6]]
    burst = [replace(j.fresh(), arrival=0.0) for j in jobs]
    serve.run_plain([replace(j.fresh(), arrival=0.0) for j in jobs[:6]])  # warm-up, untimed
    sp = serve.run_plain(plain)

    def same_as_plain(run):
        # bit-identical to the plain path; then the pinned result buffers go
        # back to the host allocator's cache, as a consumer that has written
        # its results out would release them
        ok = all(a.result is not None and b.result is not None and bool((a.result == b.result).all())
                 for a, b in zip(run, plain))
        for j in run:
            j.result = None
        return ok

    srv = serve.Server(hbm_budget=96 << 30)
    try:
        srv.run(warm)  # compile / allocate once
        same_as_plain(warm)
        st = srv.run(jobs)
        same = same_as_plain(jobs)
        sat = srv.run(burst)
        same = same_as_plain(burst) and same
    finally:
        srv.close()
    mix = {}
    for j in jobs:
        mix[j.kind] = mix.get(j.kind, 0) + 1
    for j in jobs + plain + warm + burst:
        j.result = None
        j.sgm_host = None
    return {"qps_offered": qps, "mix": mix, "served": st, "served_saturated": sat, "plain": sp,
            "throughput_ratio": round(st["jobs_per_s"] / sp["jobs_per_s"], 3),
            "capacity_ratio": round(sat["jobs_per_s"] / sp["jobs_per_s"], 3),
            "compute_bound_jobs_per_s": round(len(jobs) / sp["phase_s"]["compute"], 2),
            "results_identical": same}


if __name__ == "__main__":
    sys.exit(main())
