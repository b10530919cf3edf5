"""Helical FBP by rebinning to fan slices: drop-in for the reference's
run_mgfbp_helical entry and Mgfbp_helical class (run_mgfbp_helical.py:25-258).

For every slice position z the reference rebins a short-scan fan sinogram
from the helical cone-beam data (SinogramRebinning), then weights, filters and
back-projects it as a single fan slice with an image rotation that follows
the first rebinned view, one slice at a time (run_mgfbp_helical.py:111-137).

Here the whole stack runs as five device steps: one launch rebins every
slice (mct_helical_rebin), the weighting runs per slice plan (its Parker
weights follow the slice's angles), one ramp-filter call covers all
slices' rows, one pack, and ONE persistent batched back-projection covers all
slices, each slice a job with its own plan (angles and image rotation).  Each
slice's result equals the single-slice back-projection of its rebinned
sinogram.

As in the reference, the rebinned sinograms (before weighting) are written to
<InputDir>/temp/sgm_rebin.raw, and the slice positions are
arange(z_begin, z_end, VoxelHeight * sign(HelicalPitch)) (float32).
"""
from __future__ import annotations

import gc
import math
import os
import re
import sys
import time
import warnings

import numpy as np
import torch

from . import geometry as geo
from .config import PI, ReadConfigFile, fail, imwriteRaw
from .device import Plan
from .fbp import Mgfbp

MAX_BATCH_JOBS = 4096  # mct_backproject_batch's job limit


def run_mgfbp_helical(file_path):
    """run_mgfbp_helical.py:25-55; returns the last reconstruction (float32
    (slices, N, N), HU when WaterMu is given)."""
    start = time.time()
    print("Performing Helical Rebin and FBP from MandoCT (MI355X HIP) ...")
    warnings.filterwarnings("ignore", category=UserWarning)
    if not os.path.exists(file_path):
        print(f"ERROR: Config File {file_path} does not exist!")
        sys.exit()
    config_dict = ReadConfigFile(file_path)
    fbp = Mgfbp_helical(config_dict)
    img_recon = fbp.MainFunction()
    gc.collect()
    elapsed = time.time() - start
    if fbp.file_processed_count > 0:
        print(f"\nA total of {fbp.file_processed_count:d} file(s) are reconstructed!")
        print(f"Time cost: {elapsed:.3} sec\n")
    else:
        print(f"\nWarning: Did not find files like {fbp.input_files_pattern:s} in {fbp.input_dir:s}.")
        print("No images are reconstructed!")
    return img_recon


class Mgfbp_helical(Mgfbp):
    """run_mgfbp_helical.py:58: Mgfbp plus the helical scan's rebinning
    parameters (run_mgfbp_helical.py:59-97)."""

    def __init__(self, config_dict, device: int = 0, ctx=None):
        super().__init__(config_dict, device, ctx)
        cfg = config_dict
        if "HelicalPitch" not in cfg:
            fail("ERROR: Can not find helical pitch in the config file!")
        self.helical_pitch = cfg["HelicalPitch"]
        self.mag_ratio_isocenter = self.source_det_dis / self.source_isocenter_dis
        self.isocenter_coverage = (self.det_elem_count_vertical_actual * self.det_elem_height
                                   / self.mag_ratio_isocenter)
        self.dis_per_round = self.helical_pitch * self.isocenter_coverage
        self.num_rounds = abs(self.total_scan_angle) / (2 * PI)
        self.angle_per_view = self.total_scan_angle / self.view_num
        self.dis_per_view = self.dis_per_round / (self.view_num / self.num_rounds)
        self.gamma_max = 2 * np.arctan(0.5 * self.det_elem_count_horizontal * self.det_elem_width
                                       / self.source_det_dis) / PI
        self.short_scan_range = 2 * PI / abs(self.helical_pitch)
        if self.short_scan_range < PI + self.gamma_max:
            self.short_scan_range = PI + self.gamma_max
        self.det_offset_vertical_at_isocenter = self.det_offset_vertical / self.mag_ratio_isocenter
        sgn = np.sign(self.helical_pitch)
        z_begin = self.isocenter_coverage / 2.0 * sgn + self.det_offset_vertical_at_isocenter
        z_end = (self.dis_per_view * self.view_num - self.isocenter_coverage / 2.0 * sgn
                 + self.det_offset_vertical_at_isocenter)
        self.array_z_pos = np.arange(z_begin, z_end, self.img_voxel_height * sgn, dtype=np.float32)
        print("Reconstruction Z coverage: %.1f mm to %.1f mm" % (self.array_z_pos[0], self.array_z_pos[-1]))
        self.view_count = int(np.ceil(self.short_scan_range / abs(self.angle_per_view)))
        self.short_scan_range_actual = self.view_count * abs(self.angle_per_view)
        print("Size of rebinned sinogram is %d x %d (%.1f degrees scan)"
              % (self.det_elem_count_horizontal, self.view_count, np.rad2deg(self.short_scan_range_actual)))
        print("There are a total of %d slices to be reconstructed" % len(self.array_z_pos))
        self.img_recon = np.zeros((len(self.array_z_pos), self.img_dim, self.img_dim), dtype=np.float32)
        self.img_rot_add = 0.0

    # ------------------------------------------------------------ geometry
    def rebinned_scan_angle(self) -> tuple[float, int]:
        """ChangeReconParameterValues (run_mgfbp_helical.py:200-218): the
        rebinned sinogram's total scan angle and short-scan flag."""
        tsa = self.short_scan_range_actual * np.sign(self.total_scan_angle)
        short = 0 if abs(tsa % (2 * PI)) < (0.01 / 180 * PI) else 1
        return float(tsa), short

    def added_rotation(self, z_idx: int) -> float:
        """CalculateAddedImgRotation (run_mgfbp_helical.py:160-163)."""
        z_pos = self.array_z_pos[z_idx]
        k0 = (z_pos - self.det_offset_vertical_at_isocenter) / self.dis_per_view
        return float(np.floor(k0 - 0.5 * (self.view_count - 1.0)) * self.angle_per_view)

    def slice_spec(self, z_idx: int) -> geo.GeometrySpec:
        """The single fan slice the reference back-projects for slice z_idx:
        view_count views over the rebinned scan angle, angles and image
        rotation shifted by the slice's added rotation."""
        tsa, short = self.rebinned_scan_angle()
        rot = self.img_rot + self.added_rotation(z_idx)
        n = self.det_elem_count_horizontal
        u = geo.det_pos_array(n, n, self.positive_u_is_positive_y * self.det_elem_width,
                              (-1) * self.det_offset_horizontal, 0)
        v = geo.det_pos_array(1, 1, 0, 0, 0)
        return geo.GeometrySpec(
            img_dim=self.img_dim, img_dim_z=1, pix=self.img_pix_size, vox_h=self.img_voxel_height,
            n_views=self.view_count, sid=self.source_isocenter_dis, sdd=self.source_det_dis,
            total_scan_angle=tsa, angles=geo.angle_array(self.view_count, tsa, rot), u=u, v=v,
            cx=self.img_center_x, cy=self.img_center_y, cz=self.img_center_z, img_rot=rot,
            view_mode=self.recon_view_mode, short_scan=bool(short), cone=False, curved=self.curved_dect,
            n_v_full=1, det_h=self.det_elem_height)

    # ------------------------------------------------------------ device
    def rebin(self, sgm: torch.Tensor) -> torch.Tensor:
        """SinogramRebinning for every slice: (slices, view_count, N)."""
        return self.ctx().helical_rebin(sgm, self.array_z_pos, self.view_count, self.det_offset_vertical,
                                        self.det_elem_height, self.source_det_dis, self.mag_ratio_isocenter,
                                        self.dis_per_view)

    def reconstruct_rebinned(self, reb: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Weighting, ramp filter and back-projection of every rebinned slice
        (run_mgfbp_helical.py:127-137); reb is weighted in place."""
        ctx = self.ctx()
        nz, vc, n = reb.shape
        plans = [Plan(ctx, self.slice_spec(z)) for z in range(nz)]
        try:
            bh = self.array_bh_coefficients if self.bool_bh_correction else None
            for z, plan in enumerate(plans):
                plan.weight(reb[z:z + 1], bh)
            taps = self.recon_kernel()
            src = reb
            if self.apply_gauss_vertical and taps is not None:  # one detector row per slice
                tmp = torch.empty_like(reb)
                ctx.filter_vertical(reb.view(1, nz * vc, n), tmp.view(1, nz * vc, n),
                                    geo.gaussian_taps(1, self.det_elem_height or 1.0,
                                                      self.det_elem_vertical_gauss_filter_size),
                                    self.det_elem_height)
                src = tmp
            if taps is None:  # run_mgfbp.py:1283-1285: kernel 'None' filters nothing
                filt = src
            else:
                filt = torch.empty_like(src)
                ctx.filter_rows(src, filt, taps, self.det_elem_width, self.filter_method)
            self.last_filtered = filt[nz - 1:nz]
            packed = ctx.pack(filt, self.storage, "pair")
            step = packed.numel() // nz  # PAIR is [slice][view][column]: slices are contiguous
            if step * nz != packed.numel():
                raise RuntimeError("packed PAIR sinogram is not a whole number of slices")
            if out is None:
                out = torch.empty((nz, self.img_dim, self.img_dim), dtype=torch.float32, device=reb.device)
            jobs = [(plans[z], packed[z * step:(z + 1) * step], self.storage, "pair", out[z:z + 1])
                    for z in range(nz)]
            for i in range(0, nz, MAX_BATCH_JOBS):
                ctx.backproject_batch(jobs[i:i + MAX_BATCH_JOBS])
            ctx.synchronize()
        finally:
            for p in plans:
                p.close()
        return out

    # ------------------------------------------------------------ files
    def MainFunction(self):
        """run_mgfbp_helical.py:100-158."""
        self.file_processed_count = 0
        os.makedirs(self.input_dir + "/temp", exist_ok=True)
        dev = torch.device("cuda", self._device)
        for file in os.listdir(self.input_dir):
            if re.match(self.input_files_pattern, file) and self.ReadSinogram(file):
                if self.det_elem_count_vertical_actual != self.det_elem_count_vertical:
                    fail("ERROR: helical rebinning needs the full detector row range!")
                self.file_processed_count += 1
                print("\nReconstructing %s ..." % self.input_path)
                sgm = torch.from_numpy(self.img_sgm).to(dev)
                reb = self.rebin(sgm)
                imwriteRaw(reb.cpu().numpy(), self.input_dir + "/temp/sgm_rebin.raw")
                vol = self.reconstruct_rebinned(reb)
                self.img_recon = vol.cpu().numpy()
                self.SaveFilteredSinogram()
                print("\nSaving to %s !" % self.output_path)
                self.SaveReconImg()
        return self.img_recon
