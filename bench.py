"""Benchmark: cone-beam FDK, 512x512x256 volume from 1080 views of a
1024 x 256 flat detector (FP32 compute), on one MI355X.

One step = one pass of the FDK hot path over the resident sinogram:
restore the raw sinogram (device copy), cosine weighting, ramp filter
(FP32 direct or MFMA), repack into the back-projector's tap layout,
pixel-driven back-projection.  value = GUPS = voxel-view updates of the step
(512*512*256*1080) / step time.  The roofline object is the back-projection
kernel's, timed with HIP events on the stream it runs on; its algorithmic
bytes per update are 4 taps x 4 B (FP32 storage; 2 B for FP16/BF16) plus
4 B per voxel written.  cpu_baseline: the oracle's C/OpenMP restatement of
the same back-projection on a slab of slices (all views, all columns),
timed on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--storage f32]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# Device memory the timed steps' private input copies may take.
RESIDENT_INPUT_BYTES = 32 << 30
# HBM bytes per back-projection launch of this workload, measured with
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and scaled by a 1 GiB copy's counts
# (FETCH x2, WRITE x1); profiles/fdk_c1_kernels.md.  Keyed by
# (kernel, storage, --bp-tile, --bp-batch); other settings report null.
MEASURED_TRAFFIC = {("direct", "f32", 0, 0): 94.2e9, ("lds", "f32", 0, 0): 77.6e9}
# Per-CU fill rates on CDNA: vector L1 64 B/clk, LDS 128 B/clk; MI355X peak
# engine clock 2.4 GHz.
GATHER_BYTES_PER_CLK_PER_CU = {"direct": 64, "lds": 128}
GATHER_PATH = {"direct": "vector L1", "lds": "LDS"}
KERNEL_NAME = {"direct": "bp_cone_axial_kernel", "lds": "bp_cone_lds_kernel"}
CLOCK_GHZ = 2.4


def gather_roofline(kernel: str, gather_bytes: float, ms: float) -> dict:
    """Every update reads its four taps (16 B FP32, 8 B FP16/BF16) through the
    vector L1 (direct kernel) or from LDS (staged kernel): that path's
    ceiling."""
    import torch

    if kernel not in GATHER_PATH:
        return None
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    peak = cus * GATHER_BYTES_PER_CLK_PER_CU[kernel] * CLOCK_GHZ  # GB/s
    ach = gather_bytes / (ms * 1e-3) / 1e9
    return {"path": GATHER_PATH[kernel], "achieved": round(ach, 1), "peak": round(peak, 1),
            "unit": "GB/s", "frac": round(ach / peak, 4), "cus": cus}


def c1_config(storage: str, filter_method: str) -> dict:
    return {
        "InputDir": "/nonexistent", "OutputDir": os.environ.get("TMPDIR", "/tmp") + "/mct_bench",
        "InputFiles": "x", "OutputFilePrefix": "", "OutputFileReplace": ["a", "b"],
        "DetectorElementCountHorizontal": 1024, "DetectorElementWidth": 0.4,
        "DetectorOffsetHorizontal": 0.0, "DetectorElementCountVertical": 256,
        "Views": 1080, "TotalScanAngle": 360.0,
        "SourceIsocenterDistance": 750.0, "SourceDetectorDistance": 1250.0,
        "ImageDimension": 512, "PixelSize": 0.3, "HammingFilter": 1.0,
        "ConeBeam": True, "DetectorElementHeight": 0.4, "ImageDimensionZ": 256,
        "VoxelHeight": 0.3, "DetectorOffsetVertical": 0.0, "ImageCenterZ": 0.0,
        "SinogramStorage": storage, "RampFilterMethod": filter_method,
    }


def cpu_baseline(fbp, filt_host: np.ndarray, nslices: int) -> dict:
    """Times the oracle's C back-projection on nslices central slices."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import cbp

    spec = fbp.geometry_spec()
    g = dict(img_dim=spec.img_dim, img_dim_z=spec.img_dim_z, pix=spec.pix, vox_h=spec.vox_h,
             sid=spec.sid, sdd=spec.sdd, angles=spec.angles, u=spec.u, v=spec.v, det_h=spec.det_h,
             cone=True, cz=spec.cz, total_scan_angle=spec.total_scan_angle, short_scan=False)
    z0 = spec.img_dim_z // 2 - nslices // 2
    cbp.lib()  # build/load outside the timed region
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.perf_counter()
    out = cbp.backproject_fbp(filt_host, g, z0, z0 + nslices)
    dt = time.perf_counter() - t0
    upd = float(spec.img_dim) ** 2 * nslices * spec.n_views
    return {"value": upd / dt / 1e9, "unit": "GUPS", "cores": threads, "kind": "port",
            "sample": f"FP32 back-projection of {nslices} central slices x 512x512 columns x "
                      f"1080 views ({upd:.3g} updates) in {dt:.2f} s; oracle/bp_oracle.c, OpenMP",
            "_check": out}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16", "bf16"])
    # bf16x3: split-BF16 Toeplitz GEMM on MFMA, 2.5e-6 rel-L2 vs FP64 (FP32-grade)
    ap.add_argument("--filter", default="bf16x3", choices=["f32", "bf16", "bf16x3"])
    ap.add_argument("--cpu-slices", type=int, default=32)
    ap.add_argument("--serve-jobs", type=int, default=36,
                    help="jobs in the served-stream comparison (0: skip it)")
    ap.add_argument("--serve-qps", type=float, default=200.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--bp-tile", type=int, default=0, help="slices per thread (0: library default)")
    ap.add_argument("--bp-batch", type=int, default=0, help="loads per batch (0: library default)")
    ap.add_argument("--bp-kernel", default="auto", choices=["auto", "direct", "lds"],
                    help="cone back-projector: LDS-staged windows or per-update gathers")
    ap.add_argument("--autotune", action="store_true",
                    help="time the back-projector candidates during warm-up and keep the fastest")
    ap.add_argument("--sinogram", default="phantom", choices=["phantom", "random"],
                    help="analytic ellipsoid projections, or seeded uniform noise (profiler runs)")
    ap.add_argument("--workload", default="fdk", choices=["fdk", "fbp2d", "fp3d", "ir", "filter", "helical", "pmatrix"],
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

    # Weighting may run in place (FP32 filter, vertical Gaussian; the MFMA
    # filters weigh in their input pass), so every timed step gets a resident copy of the
    # sinogram of its own (made before the timed region); past
    # RESIDENT_INPUT_BYTES the copies are replaced by a device copy of the
    # input inside each step.
    per_step = args.steps * raw.numel() * raw.element_size() <= RESIDENT_INPUT_BYTES
    inputs = [raw.clone() for _ in range(args.steps)] if per_step else []

    def step(i):
        if per_step:
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
    bytes_alg = U * 4 * tap_bytes + float(nx) * nx * nz * 4
    achieved = bytes_alg / (bp_ms * 1e-3) / 1e9
    kernel = fbp.ctx().bp_last_kernel()
    line = {
        "metric": "FDK/backprojection GUPS", "value": round(gups, 2), "unit": "GUPS",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": ("synthetic: analytic projections of 30 seeded random ellipsoids (seed 2)"
                 if args.sinogram == "phantom" else "synthetic: seeded uniform random sinogram"),
        "config": {"workload": "cone FDK 512x512x256 volume, 1080 views, 1024x256 detector",
                   "inputs": ("one resident sinogram per timed step" if per_step
                              else "sinogram restored by a device copy inside each step"),
                   "sinogram_storage": args.storage, "ramp_filter": args.filter,
                   "bp_ms": round(bp_ms, 3), "bp_gups": round(U / (bp_ms * 1e-3) / 1e9, 2),
                   **({"autotune": {"choice": fbp.tuned, "candidates_ms": fbp.tuner.timings}}
                      if fbp.tuned else {})},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": MEASURED_TRAFFIC.get((kernel, args.storage, args.bp_tile,
                                                      args.bp_batch)),
                     "kernel": KERNEL_NAME.get(kernel, kernel),
                     "model": f"gather bytes: 4 taps x {tap_bytes} B per update + 4 B per voxel; "
                              "the taps are served by LDS/L1/L2, so frac > 1 is expected",
                     "gather": gather_roofline(kernel, U * 4 * tap_bytes, bp_ms)},
    }
    raw_h = torch.empty(raw.shape, dtype=raw.dtype, pin_memory=True)
    raw_h.copy_(raw)
    line["config"]["pcie_inclusive"] = pcie_inclusive(raw_h, work, vol, fbp)
    line["config"]["streamed"] = streamed_inclusive(raw_h, vol, args.storage, args.filter)
    if not args.no_cpu:
        filt_host = fbp.last_filtered.cpu().numpy()
        cb = cpu_baseline(fbp, filt_host, args.cpu_slices)
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
