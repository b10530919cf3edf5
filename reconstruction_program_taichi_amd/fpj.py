"""Forward projection: drop-in for the reference's run_mgfpj entry and Mgfpj
class (run_mgfpj.py:24-668).

The reference launches one Taichi kernel per detector row (run_mgfpj.py:72),
copies each row back to the host and bins / adds noise per row.  Here one HIP
launch traces every ray of every row and view, bins the oversampled elements
in its epilogue, and the noise kernel runs over the whole sinogram.

Additive config key (no reference counterpart):
  "NoiseSeed": <int>   seed of the counter-based noise generator (default 0;
                       the reference's ti.randn stream is not reproducible)
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
from .config import PI, ReadConfigFile, fail, first_key, imwriteRaw, is_num
from .device import Context, Plan


def run_mgfpj(file_path):
    """run_mgfpj.py:24-52: forward project every matching image file; returns
    the last sinogram (float32 (rows, views, n_u), rows flipped for cone-beam
    'sinogram' output as the reference does)."""
    print("Performing FPJ from MandoCT (MI355X HIP) ...")
    start = time.time()
    if not os.path.exists(file_path):
        print(f"ERROR: Config File {file_path} does not exist!")
        sys.exit()
    config_dict = ReadConfigFile(file_path)
    fpj = Mgfpj(config_dict)
    img_sgm = fpj.MainFunction()
    elapsed = time.time() - start
    if fpj.file_processed_count > 0:
        print(f"\nA total of {fpj.file_processed_count:d} files are forward projected!")
        print(f"Time cost: {elapsed:.3} sec\n")
    else:
        print(f"\nWarning: Did not find files like {fpj.input_files_pattern:s} in {fpj.input_dir:s}.")
        print("No images are  forward projected!\n")
    gc.collect()
    return img_sgm


class Mgfpj:
    """run_mgfpj.py:56: standalone forward projector (ver 0.11)."""

    count_mode = 0  # steps = floor((l_max - l_min) / (step * diag)), run_mgfpj.py:526-527

    def __init__(self, config_dict, device: int = 0, ctx: Context | None = None):
        self.config_dict = cfg = config_dict
        for key, attr in (("InputDir", "input_dir"), ("OutputDir", "output_dir"),
                          ("InputFiles", "input_files_pattern")):
            if key not in cfg:
                fail(f"ERROR: Can not find {key}!")
            if not isinstance(cfg[key], str):
                fail(f"ERROR: {key} must be a string!")
            setattr(self, attr, cfg[key])
        os.makedirs(self.output_dir, exist_ok=True)
        if "OutputFilePrefix" in cfg:
            self.output_file_prefix = cfg["OutputFilePrefix"]
            if not isinstance(self.output_file_prefix, str):
                fail("ERROR: OutputFilePrefix must be a string!")
        else:
            print("Warning: Can not find OutputFilePrefix! Set to be an empty string. ")
            self.output_file_prefix = ""
        if "OutputFileReplace" not in cfg:
            fail("ERROR: Can not find OutputFileReplace in the config file!")
        self.output_file_replace = cfg["OutputFileReplace"]
        self.output_file_form = cfg.get("OutputFileForm", "sinogram")
        if self.output_file_form not in ("sinogram", "post_log_images"):
            fail("ERROR: OutputFileForm can only be sinogram or post_log_images!")
        if "ImageDimension" not in cfg:
            fail("ERROR: can not find ImageDimension!")
        self.img_dim = cfg["ImageDimension"]
        if not isinstance(self.img_dim, int) or self.img_dim < 0:
            fail("ERROR: ImageDimension must be a positive integer!")
        if "PixelSize" not in cfg:
            fail("ERROR: can not find PixelSize!")
        self.img_pix_size = cfg["PixelSize"]
        if not is_num(self.img_pix_size):
            fail("ERROR: PixelSize must be a number!")
        if self.img_pix_size < 0:
            fail("ERROR: PixelSize must be positive!")
        k, nz = first_key(cfg, "SliceCount", "ImageDimensionZ")
        if k is None:
            fail("ERROR: Can not find image dimension along Z direction (ImageDimensionZ or SliceCount)!")
        if not isinstance(nz, int) or nz <= 0:
            fail("ERROR: ImageDimensionZ (ImageSliceCount) must be a positive integer!")
        self.img_dim_z = nz
        if "ImageCenter" in cfg:
            c = cfg["ImageCenter"]
            if not isinstance(c, list) or len(c) != 2:
                fail("ERROR: ImageCenter must be an array with two numbers!")
            self.img_center_x, self.img_center_y = c
        else:
            self.img_center_x = self.img_center_y = 0.0
        self.img_center_z = cfg.get("ImageCenterZ", 0.0)
        if not is_num(self.img_center_z):
            fail("ERROR: ImageCenterZ must be a number!")
        self.cone_beam = cfg.get("ConeBeam", False)
        if not isinstance(self.cone_beam, bool):
            fail("ERROR: ConeBeam must be true or false!")
        self.curved_dect = cfg.get("CurvedDetector", False)
        if not isinstance(self.curved_dect, bool):
            fail("ERROR: CurvedDetector can only be true or false!")
        k, n = first_key(cfg, "DetectorElementCountHorizontal", "SinogramWidth", "DetectorElementCount")
        if k is None:
            fail("ERROR: Can not find DetectorElementCountHorizontal (SinogramWidth or DetectorElementCount)!")
        if n <= 0:
            fail("ERROR: DetectorElementCountHorizontal (SinogramWidth) should be larger than 0!")
        if n % 1 != 0:
            fail("ERROR: DetectorElementCountHorizontal (SinogramWidth) should be an integer!")
        self.det_elem_count_horizontal = int(n)
        k, w = first_key(cfg, "DetectorElementWidth", "DetectorElementSize")
        if k is None:
            fail("ERROR: Can not find DetectorElementWidth (DetectorElementSize)!")
        if w <= 0:
            fail("ERROR: DetectorElementWidth (DetectorElementSize) should be larger than 0!")
        self.det_elem_width = w
        k, off = first_key(cfg, "DetectorOffcenter", "DetectorOffsetHorizontal")
        if k is None:
            off = 0.0
            print("Warning: Can not find DetectorOffsetHorizontal (DetectorOffcenter); Using default value 0")
        if not is_num(off):
            fail("ERROR: DetectorOffsetHorizontal (DetectorOffcenter) should be a number!")
        self.det_offset_horizontal = off
        for key, attr in (("SourceIsocenterDistance", "source_isocenter_dis"),
                          ("SourceDetectorDistance", "source_det_dis")):
            if key not in cfg:
                fail(f"ERROR: Can not find {key}!")
            if not is_num(cfg[key]):
                fail(f"ERROR: {key} must be a number!")
            setattr(self, attr, cfg[key])
        k, sa = first_key(cfg, "StartAngle", "ImageRotation")
        self.start_angle = 0.0 if k is None else sa / 180.0 * PI
        if not is_num(self.start_angle):
            fail("ERROR: ImageRotation must be a number!")
        self.scan_angle = cfg["TotalScanAngle"] / 180.0 * PI if "TotalScanAngle" in cfg else 2 * PI
        if not is_num(self.scan_angle):
            fail("ERROR: TotalScanAngle should be a number!")
        if "Views" not in cfg:
            fail("ERROR: Can not find number of views!")
        self.view_num = cfg["Views"]
        if not is_num(self.view_num):
            fail("ERROR: Views should be a number!")
        if self.view_num % 1 != 0 or self.view_num < 0:
            fail("ERROR: Views must be must be a positive integer!")
        self.view_num = int(self.view_num)
        self.oversample_size = cfg.get("OversampleSize", 1)
        if not is_num(self.oversample_size):
            fail("ERROR: OversampleSize should be a number!")
        if self.oversample_size % 1 != 0 or self.oversample_size < 0:
            fail("ERROR: OversampleSize must be a positive integer!")
        self.oversample_size = int(self.oversample_size)
        self.fpj_step_size = cfg.get("ForwardProjectionStepSize", 0.2)
        if self.cone_beam:
            k, vh = first_key(cfg, "VoxelHeight", "ImageSliceThickness")
            if k is None:
                fail("ERROR: Can not find image voxel height for cone beam recon!")
            self.img_voxel_height = vh
            k, nv = first_key(cfg, "DetectorElementCountVertical", "DetectorZElementCount")
            if k is None:
                fail("ERROR: Can not find detector element count along vertical direction!")
            self.det_elem_count_vertical = nv
            k, h = first_key(cfg, "SliceThickness", "DetectorElementHeight")
            if k is None:
                fail("ERROR: Can not find detector element height for cone beam recon! ")
            self.det_elem_height = h
            k, offv = first_key(cfg, "DetectorZOffcenter", "DetectorOffsetVertical")
            if k is None:
                offv = 0.0
                print("Warning: Can not find vertical detector offset for cone beam recon; Using default value 0")
            self.det_offset_vertical = offv
        else:
            self.det_elem_count_vertical = self.img_dim_z
            self.img_voxel_height = 0.0
            self.det_elem_height = 0.0
            self.det_offset_vertical = 0.0
        if "WaterMu" in cfg:
            self.water_mu = cfg["WaterMu"]
            self.convert_to_HU = True
        else:
            self.convert_to_HU = False
        if "HelicalPitch" in cfg:
            self.helical_scan = True
            self.helical_pitch = cfg["HelicalPitch"]
        else:
            self.helical_scan = False
            self.helical_pitch = 0
        if self.helical_scan and "ImageCenterZ" not in cfg:
            self.img_center_z = self.img_voxel_height * (self.img_dim_z - 1) / 2.0 * np.sign(self.helical_pitch)
            print("ImageCenterZ is re-set accordingly to be %.1f mm!" % self.img_center_z)
        if "PhotonNumber" in cfg:
            self.add_possion_noise = True
            self.photon_number = cfg["PhotonNumber"]
        else:
            self.add_possion_noise = False
            self.photon_number = 0
        self.noise_seed = int(cfg.get("NoiseSeed", 0))
        self.img_sgm = np.zeros((self.det_elem_count_vertical, self.view_num, self.det_elem_count_horizontal),
                                dtype=np.float32)
        self._device = device
        self._ctx = ctx
        self._plan = None

    # ------------------------------------------------------------ geometry
    def geometry_spec(self) -> geo.GeometrySpec:
        """Arrays of MainFunction (run_mgfpj.py:60-65)."""
        os_ = self.oversample_size
        ang = geo.angle_array(self.view_num, self.scan_angle, self.start_angle)
        v = geo.det_pos_array(self.det_elem_count_vertical, self.det_elem_count_vertical,
                              -self.det_elem_height, self.det_offset_vertical)
        u = geo.det_pos_array(self.det_elem_count_horizontal * os_, self.det_elem_count_horizontal * os_,
                              -self.det_elem_width / os_, -self.det_offset_horizontal)
        return geo.GeometrySpec(
            img_dim=self.img_dim, img_dim_z=self.img_dim_z, pix=self.img_pix_size,
            vox_h=self.img_voxel_height, n_views=self.view_num, sid=self.source_isocenter_dis,
            sdd=self.source_det_dis, total_scan_angle=self.scan_angle, angles=ang, u=u, v=v,
            cx=self.img_center_x, cy=self.img_center_y, cz=self.img_center_z, img_rot=0.0,
            cone=self.cone_beam, curved=self.curved_dect, n_v_full=self.det_elem_count_vertical,
            det_h=self.det_elem_height, fpj_step=self.fpj_step_size, fpj_count_mode=self.count_mode,
            helical=self.helical_scan, helical_pitch=self.helical_pitch)

    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = Context(self._device)
        return self._ctx

    def plan(self) -> Plan:
        if self._plan is None:
            self._plan = Plan(self.ctx(), self.geometry_spec())
        return self._plan

    def project(self, vol: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """ForwardProjectionBilinear + BinSinogram (+ AddPossionNoise) for every
        row and view of a resident FP32 volume (Nz, N, N) in mu units."""
        plan = self.plan()
        if out is None:
            out = torch.empty((self.det_elem_count_vertical, self.view_num, self.det_elem_count_horizontal),
                              dtype=torch.float32, device=vol.device)
        plan.forward_project(vol, out, None, self.oversample_size)
        if self.add_possion_noise:
            self.ctx().add_noise(out, self.photon_number, self.noise_seed)
        return out

    # ------------------------------------------------------------ files
    def ReadImage(self, file):
        """run_mgfpj.py:618-638."""
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
        """run_mgfpj.py:656-668."""
        if self.cone_beam:
            if self.output_file_form == "sinogram":
                self.img_sgm = np.flip(self.img_sgm, axis=0)
                imwriteRaw(self.img_sgm, self.output_path)
            else:
                imwriteRaw(self.img_sgm.transpose([1, 0, 2]), self.output_path)
        else:
            imwriteRaw(self.img_sgm, self.output_path)

    def MainFunction(self):
        """run_mgfpj.py:57-96."""
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
