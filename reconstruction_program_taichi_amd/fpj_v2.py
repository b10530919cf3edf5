"""Forward projection on Mgfbp's geometry: drop-in for the reference's
run_mgfpj_v2 entry and its Mgfpj class (run_mgfpj_v2.py:35-470), which
inherits the FBP configuration (P-matrix included) and adds the
forward-projection keys.

Differences from run_mgfpj (v1) that this module keeps, as the reference
computes them:
  * detector arrays and angles are Mgfbp's (InitializeArrays,
    run_mgfbp.py:1244-1259): u has the FBP convention and only its first
    DetectorElementCountHorizontal entries are set (the oversampled entries
    beyond them stay 0, run_mgfpj_v2.py:210-211 with :74); v covers the
    vertical reconstruction range;
  * "StartAngle" replaces the image rotation without a degree-to-radian
    conversion (run_mgfpj_v2.py:203-207);
  * rows outside the vertical reconstruction range stay 0 in the output and
    receive no noise (the reference projects row by row inside the range);
  * the P-matrix branch (PMatrixFile) traces s * A_j [u, v, 1] + x_s_j.

The projection itself is one launch over every row and view (mct_forward_
project), the oversampled elements binned in its epilogue, the noise one more
launch over the projected rows.
"""
from __future__ import annotations

import gc
import os
import re
import sys
import time

import numpy as np
import torch

from . import geometry as geo
from . import pmatrix as pmx
from .config import fail, imwriteRaw, is_num
from .fbp import Mgfbp


def run_mgfpj_v2(file_path):
    """run_mgfpj_v2.py:35-67; returns the last sinogram (float32
    (DetectorElementCountVertical, views, n_u), rows flipped for cone-beam
    'sinogram' output)."""
    return _run(file_path, Mgfpj, "Performing FPJ from MandoCT (MI355X HIP, ver 0.2) ...")


def _run(file_path, cls, banner):
    from .config import ReadConfigFile

    print(banner)
    start = time.time()
    if not os.path.exists(file_path):
        print(f"ERROR: Config File {file_path} does not exist!")
        sys.exit()
    fpj = cls(ReadConfigFile(file_path))
    img_sgm = fpj.MainFunction()
    elapsed = time.time() - start
    if fpj.file_processed_count > 0:
        print(f"\nA total of {fpj.file_processed_count:d} files are forward projected!")
        print(f"Time cost is {elapsed:.3} sec\n")
    else:
        print(f"\nWarning: Did not find files like {fpj.input_files_pattern:s} in {fpj.input_dir:s}.")
        print("No images are forward projected!\n")
    gc.collect()
    return img_sgm


class Mgfpj(Mgfbp):
    """run_mgfpj_v2.py:70 (Mgfpj inheriting Mgfbp)."""

    count_mode = 0  # steps = floor((l_max - l_min) / (step * diag)), run_mgfpj_v2.py:299-300
    noise_all_rows = False

    def __init__(self, config_dict, device: int = 0, ctx=None):
        super().__init__(config_dict, device, ctx)
        cfg = config_dict
        self.output_file_form = cfg.get("OutputFileForm", "sinogram")
        if self.output_file_form not in ("sinogram", "post_log_images"):
            fail("ERROR: OutputFileForm can only be sinogram or post_log_images!")
        self.oversample_size = cfg.get("OversampleSize", 1)
        if not is_num(self.oversample_size):
            fail("ERROR: OversampleSize should be a number!")
        if self.oversample_size % 1 != 0 or self.oversample_size < 0:
            fail("ERROR: OversampleSize must be a positive integer!")
        self.oversample_size = int(self.oversample_size)
        self.fpj_step_size = cfg.get("ForwardProjectionStepSize", 0.2)
        if "HelicalPitch" in cfg:
            self.helical_scan, self.helical_pitch = True, cfg["HelicalPitch"]
            print("--Helical scan")
        else:
            self.helical_scan, self.helical_pitch = False, 0.0
        if self.helical_scan and ("ImageCenterZ" not in cfg or getattr(self, "img_center_z_auto_set_from_fbp", False)):
            self.img_center_z = self.img_voxel_height * (self.img_dim_z - 1) / 2.0 * np.sign(self.helical_pitch)
            print("Warning: ImageCenterZ is not in the config file or is wrongly set from run_mgfbp!")
            print("For helical scans, the first view begins with the bottom or the top of the image object;")
            print("ImageCenterZ is re-set accordingly to be %.1f mm!" % self.img_center_z)
        if "PhotonNumber" in cfg:
            self.add_possion_noise = True
            self.photon_number = cfg["PhotonNumber"]
            if not is_num(self.photon_number):
                fail("ERROR: PhotonNumber must be a number!")
        else:
            self.add_possion_noise = False
            self.photon_number = 0
        if "StartAngle" in cfg:
            self.img_rot = cfg["StartAngle"]
            if not is_num(self.img_rot):
                fail("ERROR: StartAngle must be a number!")
        self.noise_seed = int(cfg.get("NoiseSeed", 0))
        self.img_sgm = np.zeros((self.det_elem_count_vertical, self.view_num, self.det_elem_count_horizontal),
                                dtype=np.float32)
        self.img_image = np.zeros((self.img_dim_z, self.img_dim, self.img_dim), dtype=np.float32)
        if self.bool_apply_pmatrix:  # GenerateMatrixAFromPMatrix, run_mgfpj_v2.py:211-214
            self.matrix_A_each_view, self.x_s_each_view = pmx.matrix_a_and_source(self.array_pmatrix,
                                                                                  self.view_num)
        self._fp_plan = None

    # ------------------------------------------------------------ geometry
    def _u_array(self) -> np.ndarray:
        """Mgfbp's u for the first n_u entries, 0 for the rest of the
        oversampled array (run_mgfpj_v2.py:74, :210-211)."""
        n, os_ = self.det_elem_count_horizontal, self.oversample_size
        u = np.zeros(n * os_, dtype=np.float32)
        u[:n] = geo.det_pos_array(n, n, self.positive_u_is_positive_y * self.det_elem_width,
                                  (-1) * self.det_offset_horizontal, 0)
        return u

    def _center_z(self):
        return None

    def fp_geometry_spec(self) -> geo.GeometrySpec:
        """Geometry of ForwardProjectionBilinear (run_mgfpj_v2.py:236-413):
        Mgfbp's v and angles, the u array above, the forward-projection
        step, helical pitch and P-matrix inverse form."""
        s = Mgfbp.geometry_spec(self)
        s.u = self._u_array()
        s.angles = geo.angle_array(self.view_num, self.total_scan_angle, self.img_rot)
        s.img_rot = 0.0
        s.fpj_step, s.fpj_count_mode = self.fpj_step_size, self.count_mode
        s.helical, s.helical_pitch = self.helical_scan, self.helical_pitch
        s.cz, s.center_z = self.img_center_z, self._center_z()
        s.view_mode = 1
        if self.bool_apply_pmatrix:
            s.matrix_a, s.x_s = self.matrix_A_each_view, self.x_s_each_view
        return s

    def fp_plan(self):
        from .device import Plan

        if self._fp_plan is None:
            self._fp_plan = Plan(self.ctx(), self.fp_geometry_spec())
        return self._fp_plan

    def project(self, vol: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Every row of the vertical range and every view of a resident FP32
        volume (Nz, N, N) in mu units; returns (DetectorElementCountVertical,
        views, N) with the rows outside the range 0."""
        plan = self.fp_plan()
        n, rows, b = self.det_elem_count_horizontal, self.det_elem_count_vertical_actual, \
            self.det_elem_vertical_recon_range_begin
        if out is None:
            out = torch.zeros((self.det_elem_count_vertical, self.view_num, n), dtype=torch.float32,
                              device=vol.device)
        else:
            out.zero_()
        part = out[b:b + rows]
        plan.forward_project(vol, part, None, self.oversample_size)
        if self.add_possion_noise:
            self.ctx().add_noise(out if self.noise_all_rows else part, self.photon_number, self.noise_seed)
        return out

    # ------------------------------------------------------------ files
    def ReadImage(self, file):
        """run_mgfpj_v2.py:423-441."""
        self.input_path = os.path.join(self.input_dir, file)
        self.output_file = re.sub(self.output_file_replace[0], self.output_file_replace[1], file)
        if self.output_file == file:
            print(f"ERROR: did not find string '{self.output_file_replace[0]}' to replace in '{self.output_file}'")
            sys.exit()
        self.output_path = os.path.join(self.output_dir, self.output_file_prefix + self.output_file)
        img = np.fromfile(self.input_path, dtype=np.float32).reshape(self.img_dim_z, self.img_dim, self.img_dim)
        if self.convert_to_HU:
            img = (img + 1000.0) / 1000.0 * self.water_mu
        self.img_image = np.ascontiguousarray(img, dtype=np.float32)
        return True

    def SaveSinogram(self):
        """run_mgfpj_v2.py:458-470."""
        if self.cone_beam:
            if self.output_file_form == "sinogram":
                self.img_sgm = np.flip(self.img_sgm, axis=0)
                imwriteRaw(self.img_sgm, self.output_path)
            else:
                imwriteRaw(self.img_sgm.transpose([1, 0, 2]), self.output_path)
        else:
            imwriteRaw(self.img_sgm, self.output_path)

    def MainFunction(self):
        """run_mgfpj_v2.py:71-111."""
        self.file_processed_count = 0
        dev = torch.device("cuda", self._device)
        for file in os.listdir(self.input_dir):
            if re.match(self.input_files_pattern, file) and self.ReadImage(file):
                print("\nForward projecting %s ..." % self.input_path)
                self.file_processed_count += 1
                vol = torch.from_numpy(self.img_image).to(dev)
                self.img_sgm = self.project(vol).cpu().numpy()
                print("\nSaving to %s !" % self.output_path)
                self.SaveSinogram()
        return self.img_sgm
