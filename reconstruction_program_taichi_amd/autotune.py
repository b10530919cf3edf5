"""Autotuner for the projectors' kernel and tile shapes.

For one geometry (a Plan) and dtype it runs every candidate configuration on
the caller's own buffers, times each with HIP events on the context's stream
(after one untimed run that builds the plan's tables), keeps the fastest and
applies it to the context.  Results are cached per geometry key in memory and,
optionally, in a JSON file, so a serving process tunes each geometry once.

Back-projector candidates (mct_set_bp_kernel / mct_set_bp_tile):
  cone beam: the LDS-staged kernel, and the direct-gather kernel at every
  (slices per thread, loads per batch) in {8, 16, 32} x {4, 8, 16}, batch <=
  slices; fan beam: slices per thread in {4, 8, 16, 32}.
Forward-projector candidates (mct_set_fp_kernel): bilinear cells, direct.

All candidates compute the same result to rounding (the tests check each
kernel against the oracle); the tuner only chooses speed.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

import torch

from .device import Context, Plan


def bp_candidates(cone: bool) -> list[tuple[str, int, int]]:
    if not cone:
        return [("direct", zc, 0) for zc in (4, 8, 16, 32)]
    out = [("lds", 0, 0)]
    out += [("direct", zc, b) for zc in (8, 16, 32) for b in (4, 8, 16) if b <= zc]
    return out


FP_CANDIDATES = ("cells", "direct")


def geometry_key(plan: Plan, what: str, dtype: str = "f32") -> str:
    s = plan.spec
    return (f"{what}:{'cone' if s.cone else 'fan'}:{'curved' if s.curved else 'flat'}:"
            f"{s.img_dim}x{s.img_dim}x{s.img_dim_z}:v{s.n_views}:d{s.n_u}x{s.n_v}:"
            f"pix{s.pix:.6g}:vh{s.vox_h:.6g}:sid{s.sid:.6g}:sdd{s.sdd:.6g}:{dtype}")


def _time(fn, reps: int) -> float:
    stream = torch.cuda.current_stream()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    fn()  # untimed: tables, caches
    start.record(stream)
    for _ in range(reps):
        fn()
    end.record(stream)
    end.synchronize()
    return start.elapsed_time(end) / reps


@dataclass
class Autotuner:
    ctx: Context
    cache_path: str | None = None
    reps: int = 2
    cache: dict = field(default_factory=dict)
    timings: dict = field(default_factory=dict)

    def __post_init__(self):
        if self.cache_path and os.path.exists(self.cache_path):
            with open(self.cache_path) as f:
                self.cache.update(json.load(f))

    def _save(self) -> None:
        if self.cache_path:
            tmp = self.cache_path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.cache, f, indent=1, sort_keys=True)
            os.replace(tmp, self.cache_path)

    # ---------------------------------------------------------- back-projection
    def apply_bp(self, choice: dict) -> None:
        self.ctx.set_bp_kernel(choice["kernel"])
        self.ctx.set_bp_tile(choice["zc"], choice["batch"])

    def tune_backprojector(self, plan: Plan, packed: torch.Tensor, dtype: str, layout: str,
                           vol: torch.Tensor) -> dict:
        """Times every candidate on (packed, vol), applies and returns the
        fastest {kernel, zc, batch, ms}.  vol is overwritten."""
        key = geometry_key(plan, "bp", dtype)
        if key in self.cache:
            self.apply_bp(self.cache[key])
            return self.cache[key]
        times = {}
        for kernel, zc, batch in bp_candidates(plan.spec.cone):
            choice = {"kernel": kernel, "zc": zc, "batch": batch}
            try:
                self.apply_bp(choice)
                ms = _time(lambda: plan.backproject_fbp(packed, dtype, layout, vol), self.reps)
            except RuntimeError:  # e.g. the LDS window does not fit this geometry
                continue
            times[(kernel, zc, batch)] = ms
        if not times:
            raise RuntimeError("no back-projector candidate ran")
        (kernel, zc, batch), ms = min(times.items(), key=lambda kv: kv[1])
        best = {"kernel": kernel, "zc": zc, "batch": batch, "ms": round(ms, 4),
                "candidates": len(times)}
        self.timings[key] = {f"{k}/{z}/{b}": round(t, 4) for (k, z, b), t in times.items()}
        self.cache[key] = best
        self.apply_bp(best)
        self._save()
        return best

    # -------------------------------------------------------- forward projection
    def tune_forward_projector(self, plan: Plan, vol: torch.Tensor, out: torch.Tensor,
                               oversample: int = 1) -> dict:
        key = geometry_key(plan, "fp")
        if key in self.cache:
            self.ctx.set_fp_kernel(self.cache[key]["kernel"])
            return self.cache[key]
        times = {}
        for kernel in FP_CANDIDATES:
            try:
                self.ctx.set_fp_kernel(kernel)
                ms = _time(lambda: plan.forward_project(vol, out, None, oversample), self.reps)
            except RuntimeError:
                continue
            times[kernel] = ms
        if not times:
            raise RuntimeError("no forward-projector candidate ran")
        kernel, ms = min(times.items(), key=lambda kv: kv[1])
        best = {"kernel": kernel, "ms": round(ms, 4), "candidates": len(times)}
        self.timings[key] = {k: round(t, 4) for k, t in times.items()}
        self.cache[key] = best
        self.ctx.set_fp_kernel(kernel)
        self._save()
        return best
