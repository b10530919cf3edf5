"""FBP / FDK: drop-in for the reference's run_mgfbp entry and Mgfbp class.

Reference: run_mgfbp.py:18-47 (run_mgfbp), :51-1310 (Mgfbp).  Config parsing
keeps the reference's keys, aliases, defaults and messages (errors print and
raise SystemExit); the pipeline per file is weighting -> (vertical Gaussian)
-> ramp filter -> back-projection, all as HIP kernels on one device with the
sinogram and the volume resident in HBM.

Additive config keys (no reference counterpart):
  "SheppLoganFilter": <any>      Shepp-Logan taps instead of Hamming/Gaussian
  "SinogramStorage": "f32"|"f16"|"bf16"
                                 dtype of the filtered sinogram the
                                 back-projector reads (default f32)
  "RampFilterMethod": "f32"|"bf16"|"bf16x3"
                                 row-filter arithmetic (default f32)
  "Autotune": true|false         time the back-projector's candidate kernels
                                 and tiles on the first sinogram of a geometry
                                 and keep the fastest (default false)
"""
from __future__ import annotations

import gc
import os
import re
import sys
import time
import warnings

import numpy as np
import torch

from . import geometry as geo
from . import pmatrix as pmx
from .config import (PI, ReadConfigFile, fail, first_key, imreadRaw, imwriteRaw, imwriteTiff,
                     is_num, save_jsonc)
from .device import Context, Plan, _packed_bytes


def run_mgfbp(file_path):
    """run_mgfbp.py:18-47: reconstruct every matching file; returns the last
    reconstruction (float32 (Nz, N, N), HU when WaterMu is given)."""
    print("Performing FBP from MandoCT (MI355X HIP) ...")
    start = time.time()
    warnings.filterwarnings("ignore", category=UserWarning)
    if not os.path.exists(file_path):
        print(f"ERROR: Config File {file_path} does not exist!")
        sys.exit()
    config_dict = ReadConfigFile(file_path)
    fbp = Mgfbp(config_dict)
    img_recon = fbp.MainFunction()
    elapsed = time.time() - start
    if fbp.file_processed_count > 0:
        print(f"\nA total of {fbp.file_processed_count:d} file(s) are reconstructed!")
        print(f"Time cost: {elapsed:.3} sec\n")
    else:
        print(f"\nWarning: Did not find files like {fbp.input_files_pattern:s} in {fbp.input_dir:s}.")
        print("No images are reconstructed!\n")
    gc.collect()
    return img_recon


class Mgfbp:
    """run_mgfbp.py:51: geometry from the config, then MainFunction()."""

    def __init__(self, config_dict, device: int = 0, ctx: Context | None = None):
        self.config_dict = cfg = config_dict
        self._parse_io(cfg)
        self._parse_detector(cfg)
        self._parse_scan(cfg)
        self._parse_image(cfg)
        self._parse_kernel(cfg)
        self._parse_cone(cfg)
        self.storage = cfg.get("SinogramStorage", "f32")
        if self.storage not in ("f32", "f16", "bf16"):
            fail("ERROR: SinogramStorage can only be f32, f16 or bf16!")
        self.filter_method = cfg.get("RampFilterMethod", "f32")
        if self.filter_method not in ("f32", "bf16", "bf16x3"):
            fail("ERROR: RampFilterMethod can only be f32, bf16 or bf16x3!")
        self.autotune = cfg.get("Autotune", False)
        if not isinstance(self.autotune, bool):
            fail("ERROR: Autotune must be true or false!")
        self.tuner = None
        self.tuned = None
        self.img_recon = np.zeros((self.img_dim_z, self.img_dim, self.img_dim), dtype=np.float32)
        self._device = device
        self._ctx = ctx
        self._plan = None
        if "SaveModifiedConfigFolder" in cfg:
            folder = cfg["SaveModifiedConfigFolder"]
            if not isinstance(folder, str):
                fail("ERROR: SaveModifiedConfigFolder must be a string!")
            del cfg["SaveModifiedConfigFolder"]
            os.makedirs(folder, exist_ok=True)
            save_jsonc(folder + "/config_mgfbp.jsonc", cfg)
            print("Modified config files are saved to %s folder." % folder)
            if self.bool_apply_pmatrix:
                # run_mgfbp.py:698-718
                for name, arr in (("pmatrix_file", self.array_pmatrix),
                                  ("sid_file", self.source_isocenter_dis_each_view),
                                  ("sdd_file", self.source_det_dis_each_view),
                                  ("det_offset_horizontal_file", np.squeeze(self.det_offset_horizontal_each_view)),
                                  ("det_offset_vertical_file", np.squeeze(self.det_offset_vertical_each_view)),
                                  ("scan_angle_file", np.squeeze(self.scan_angle_each_view))):
                    save_jsonc(folder + f"/{name}.jsonc", {"Value": arr.tolist()})

    # ------------------------------------------------------------ config
    def _parse_io(self, cfg):
        """run_mgfbp.py:89-207."""
        for key, attr in (("InputDir", "input_dir"), ("OutputDir", "output_dir")):
            if key not in cfg:
                fail(f"ERROR: Can not find {key} in the config file!")
            if type(cfg[key]) != str:
                fail(f"ERROR: {key} is not a string!")
            setattr(self, attr, cfg[key])
        os.makedirs(self.output_dir, exist_ok=True)
        if "InputFiles" not in cfg:
            fail("ERROR: Can not find InputFiles in the config file!")
        if type(cfg["InputFiles"]) != str:
            fail("ERROR: InputFiles is not a string!")
        self.input_files_pattern = cfg["InputFiles"]
        if "OutputFilePrefix" not in cfg:
            fail("ERROR: Can not find OutputFilePrefix in the config file!")
        if type(cfg["OutputFilePrefix"]) != str:
            fail("ERROR: OutputFilePrefix is not a string!")
        self.output_file_prefix = cfg["OutputFilePrefix"]
        if "OutputFileReplace" not in cfg:
            fail("ERROR: Can not find OutputFileReplace in the config file!")
        self.output_file_replace = cfg["OutputFileReplace"]
        self.output_file_format = cfg.get("OutputFileFormat", "raw")
        if self.output_file_format not in ("raw", "tif", "tiff"):
            fail("ERROR: Output file format can only be 'raw', 'tif' or 'tiff' !")
        self.input_file_form = cfg.get("InputFileForm", "sinogram")
        if self.input_file_form not in ("sinogram", "post_log_images"):
            fail("ERROR: InputFileForm can only be sinogram or post_log_images!")
        self.output_file_form = cfg.get("OutputFileForm", "sinogram")
        if self.output_file_form not in ("sinogram", "post_log_images"):
            fail("ERROR: OutputFileForm can only be sinogram or post_log_images!")
        if self.input_file_form == "post_log_images" or self.output_file_form == "post_log_images":
            self.first_slice_top_row = True
        elif "FirstSinogramSliceIsDetectorTopRow" in cfg:
            if not isinstance(cfg["FirstSinogramSliceIsDetectorTopRow"], bool):
                fail("ERROR: FirstSinogramSliceIsDetectorTopRow can only be true or false!")
            self.first_slice_top_row = cfg["FirstSinogramSliceIsDetectorTopRow"]
        else:
            self.first_slice_top_row = False
        self.positive_v_is_positive_z = -1 if self.first_slice_top_row else 1
        self.save_filtered_sinogram = cfg.get("SaveFilteredSinogram", False)
        if not isinstance(self.save_filtered_sinogram, bool):
            fail("ERROR: SaveFilteredSinogram can only be true or false!")

    def _parse_detector(self, cfg):
        """run_mgfbp.py:209-307."""
        self.curved_dect = cfg.get("CurvedDetector", False)
        if not isinstance(self.curved_dect, bool):
            fail("ERROR: CurvedDetector can only be true or false!")
        k, n = first_key(cfg, "DetectorElementCountHorizontal", "SinogramWidth")
        if k is None:
            fail("ERROR: Can not find detector element count along horizontal direction!")
        if n <= 0:
            fail("ERROR: DetectorElementCountHorizontal (SinogramWidth) should be larger than 0!")
        if n % 1 != 0:
            fail("ERROR: DetectorElementCountHorizontal (SinogramWidth) should be an integer!")
        self.det_elem_count_horizontal = int(n)
        k, w = first_key(cfg, "DetectorElementWidth", "DetectorElementSize")
        if k is None:
            fail("ERROR: Can not find detector element width!")
        if w <= 0:
            fail("ERROR: DetectorElementWidth (DetectorElementSize) should be larger than 0!")
        self.det_elem_width = w
        self.positive_u_is_positive_y = -1
        k, off = first_key(cfg, "DetectorOffcenter", "DetectorOffsetHorizontal")
        if k is None:
            # run_mgfbp.py:249-257 warns and leaves the attribute unset, so the
            # next line of the reference raises AttributeError.
            print("Warning: Can not find horizontal detector offset; Using default value 0")
            raise AttributeError("'Mgfbp' object has no attribute 'det_offset_horizontal'")
        if not (is_num(off) or isinstance(off, list)):
            fail("ERROR: DetectorOffsetHorizontal (DetectorOffcenter) should be a number!")
        self.det_offset_horizontal = off[0] if isinstance(off, list) else off
        k, nv = first_key(cfg, "DetectorElementCountVertical", "SliceCount")
        if k is None:
            fail("ERROR: Can not find detector element count along vertical direction!")
        if nv <= 0:
            fail("ERROR: DetectorElementCountVertical (SliceCount) should be larger than 0!")
        if nv % 1 != 0:
            fail("ERROR: DetectorElementCountVertical (SliceCount) should be an integer!")
        self.det_elem_count_vertical = int(nv)
        if "DetectorElementVerticalReconRange" in cfg:
            rr = cfg["DetectorElementVerticalReconRange"]
            if not isinstance(rr, list) or len(rr) != 2:
                fail("ERROR: DetectorElementVerticalReconRange should be an array with two numbers!")
            self.det_elem_vertical_recon_range_begin, self.det_elem_vertical_recon_range_end = rr
            if rr[1] > self.det_elem_count_vertical - 1 or rr[0] < 0:
                fail("ERROR: Out of detector row range!")
            if rr[0] % 1 != 0 or rr[1] % 1 != 0:
                fail("ERROR: DetectorElementVerticalReconRange must be integers!")
            print(f"--Reconstructing from detector row #{rr[0]:d} to #{rr[1]:d}")
            self.det_elem_count_vertical_actual = rr[1] - rr[0] + 1
        else:
            self.det_elem_vertical_recon_range_begin = 0
            self.det_elem_vertical_recon_range_end = self.det_elem_count_vertical - 1
            self.det_elem_count_vertical_actual = self.det_elem_count_vertical
        if "DetectorElementVerticalGaussFilterSize" in cfg:
            self.det_elem_vertical_gauss_filter_size = cfg["DetectorElementVerticalGaussFilterSize"]
            if self.det_elem_vertical_gauss_filter_size <= 0:
                fail("ERROR: DetectorElementVerticalGaussFilterSize should be larger than 0!")
            self.apply_gauss_vertical = True
        else:
            self.apply_gauss_vertical = False
            self.det_elem_vertical_gauss_filter_size = 0.0001

    def _parse_scan(self, cfg):
        """run_mgfbp.py:311-391."""
        if "Views" not in cfg:
            fail("ERROR: Can not find number of views!")
        self.view_num = cfg["Views"]
        if self.view_num % 1 != 0 or self.view_num < 0:
            fail("ERROR: Views must be larger than 0 and must be an integer!")
        self.view_num = int(self.view_num)
        self.sgm_height = cfg.get("SinogramHeight", self.view_num)
        if self.sgm_height % 1 != 0 or self.sgm_height < 0:
            fail("ERROR: SinogramHeight must be larger than 0 and must be an integer!")
        self.sgm_height = int(self.sgm_height)
        if "TotalScanAngle" in cfg:
            tsa = cfg["TotalScanAngle"]
            if isinstance(tsa, list):
                tsa = tsa[0]
            if not is_num(tsa):
                fail("ERROR: TotalScanAngle should be a number!")
            self.total_scan_angle = tsa / 180.0 * PI
        else:
            self.total_scan_angle = 2 * PI
        self.short_scan = 0 if abs(self.total_scan_angle % (2 * PI)) < (0.01 / 180 * PI) else 1
        print(("--Short scan" if self.short_scan else "--Full scan")
              + ", scan angle = %.1f degrees" % (self.total_scan_angle / PI * 180))
        if "BeamHardeningCorrectionCoefficients" in cfg:
            c = cfg["BeamHardeningCorrectionCoefficients"]
            if not isinstance(c, list):
                fail("ERROR: BeamHardeningCorrectionCoefficients must be an array!")
            self.bool_bh_correction = True
            self.array_bh_coefficients = np.array(c, dtype=np.float32)
            self.bh_corr_order = len(c)
        else:
            self.bool_bh_correction = False
        for key, attr in (("SourceIsocenterDistance", "source_isocenter_dis"),
                          ("SourceDetectorDistance", "source_det_dis")):
            if key in cfg:
                val = cfg[key]
                if not is_num(val):
                    fail(f"ERROR: {key} must be a number!")
                if val <= 0.0:
                    fail(f"ERROR: {key} must be positive!")
            else:
                val = self.det_elem_count_horizontal * self.det_elem_width * 1000.0
                print(f"Warning: Did not find {key}; Set to infinity!")
            setattr(self, attr, val)

    def _parse_image(self, cfg):
        """run_mgfbp.py:394-455."""
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
        rot = cfg.get("ImageRotation", 0.0)
        if not is_num(rot):
            fail("ERROR: ImageRotation must be a number!")
        self.img_rot = rot / 180.0 * PI
        if "ImageCenter" in cfg:
            c = cfg["ImageCenter"]
            if not isinstance(c, list) or len(c) != 2:
                fail("ERROR: ImageCenter must be an array with two numbers!")
            self.img_center_x, self.img_center_y = c
        else:
            self.img_center_x = self.img_center_y = 0.0
        mode = cfg.get("ReconViewMode", "axial")
        modes = {"axial": 1, "coronal": 2, "sagittal": 3}
        if mode not in modes:
            fail("ERROR: ReconViewMode can only be axial, coronal or sagittal!")
        self.recon_view_mode = modes[mode]

    def _parse_kernel(self, cfg):
        """run_mgfbp.py:457-492."""
        if "HammingFilter" in cfg:
            self.kernel_name, self.kernel_param = "HammingFilter", cfg["HammingFilter"]
        elif "GaussianApodizedRamp" in cfg:
            self.kernel_name, self.kernel_param = "GaussianApodizedRamp", cfg["GaussianApodizedRamp"]
        elif "SheppLoganFilter" in cfg:
            self.kernel_name, self.kernel_param = "SheppLoganFilter", 0.0
        elif "None" in cfg:
            self.kernel_name, self.kernel_param = "None", 0.0
        else:
            self.kernel_param = 0.0  # kernel_name stays unset, as in the reference
        if not is_num(self.kernel_param):
            fail("ERROR: Kernel parameter must be a number!")
        self.dbt_or_not = cfg.get("DBTMode", 0)
        if "WaterMu" in cfg:
            self.water_mu = cfg["WaterMu"]
            if not is_num(self.water_mu):
                fail("ERROR: WaterMu must be a number!")
            self.convert_to_HU = True
        else:
            self.convert_to_HU = False

    def _parse_pmatrix(self, cfg):
        """P-matrix keys and normalisation (run_mgfbp.py:551-627): read
        PMatrixFile's 'Value' (12 floats per view), derive the scan
        parameters from the matrices (written back into the config dict, as
        the reference does), then adapt the matrices to the scan's pixel
        pitch and vertical reconstruction range."""
        d = ReadConfigFile(cfg["PMatrixFile"])
        if "Value" not in d:
            fail("ERROR: PMatrixFile has no member named 'Value'!")
        val = d["Value"]
        if not isinstance(val, list):
            fail("ERROR: PMatrixFile.Value is not an array")
        if len(val) != self.view_num * 12:
            fail(f"ERROR: view number is {self.view_num:d} while pmatrix has {len(val):d} elements!")
        self.array_pmatrix = np.array(val, dtype=np.float32)
        self.bool_apply_pmatrix = 1
        print("--PMatrix applied")
        self.modify_pmatrix_to_standard_form = cfg.get("ModifyPMatrixToStandardForm", False)
        if not isinstance(self.modify_pmatrix_to_standard_form, bool):
            fail("ERROR: ModifyPMatrixToStandardForm must be True or False!")
        for key, attr, default, what in (
                ("PMatrixDetectorElementWidth", "pmatrix_elem_width", self.det_elem_width, "width"),
                ("PMatrixDetectorElementHeight", "pmatrix_elem_height", self.det_elem_height, "height")):
            if key in cfg:
                print(f"--PMatrix detector pixel {what} is different from the CT scan")
                v = cfg[key]
                if not is_num(v):
                    fail(f"ERROR: {key} must be a number!")
                if v <= 0:
                    fail(f"ERROR: {key} must be positive!")
                setattr(self, attr, v)
            else:
                setattr(self, attr, default)
        r = pmx.source_trajectory(self.array_pmatrix, self.view_num, self.modify_pmatrix_to_standard_form,
                                  self.pmatrix_elem_width, self.pmatrix_elem_height,
                                  self.det_elem_count_horizontal, self.det_elem_width,
                                  self.det_elem_count_vertical, self.det_elem_height)
        self.array_pmatrix = r["pmatrix"]
        self.positive_u_is_positive_y = r["positive_u_is_positive_y"]
        self.positive_v_is_positive_z = r["positive_v_is_positive_z"]
        self.det_offset_horizontal = r["det_offset_horizontal"]
        self.det_offset_vertical = r["det_offset_vertical"]
        self.source_isocenter_dis = r["source_isocenter_dis"]
        self.source_det_dis = r["source_det_dis"]
        self.source_isocenter_dis_each_view = r["sid_each"]
        self.source_det_dis_each_view = r["sdd_each"]
        self.det_offset_horizontal_each_view = r["det_offset_horizontal_each"]
        self.det_offset_vertical_each_view = r["det_offset_vertical_each"]
        self.scan_angle_each_view = r["scan_angle_each"]
        print("Parameters are updated from PMatrix:")
        print("Mean Offset values are %.2f mm and %.2f mm for horizontal and vertical direction respectively;"
              % (self.det_offset_horizontal, self.det_offset_vertical))
        print("Mean Source to Isocenter Distance is %.2f mm;" % self.source_isocenter_dis)
        print("Mean Source to Detector Distance is %.2f mm;" % self.source_det_dis)
        if self.short_scan:
            self.total_scan_angle = r["scan_sum"] / (self.view_num - 1) * self.view_num
            print("Total Scan Angle is %.2f degrees." % (self.total_scan_angle / PI * 180.0))
        else:
            print("Total Scan Angle is not updated for full scan.")
        cfg["SourceIsocenterDistance"] = self.source_isocenter_dis
        cfg["SourceDetectorDistance"] = self.source_det_dis
        cfg["DetectorOffsetHorizontal"] = self.det_offset_horizontal
        cfg["DetectorOffsetVertical"] = self.det_offset_vertical
        cfg["TotalScanAngle"] = self.total_scan_angle / PI * 180.0
        self.array_pmatrix = pmx.pixel_size(self.array_pmatrix, self.view_num, self.det_elem_width,
                                            self.det_elem_height, self.pmatrix_elem_width,
                                            self.pmatrix_elem_height)
        self.array_pmatrix = pmx.vertical_recon_range(self.array_pmatrix, self.view_num,
                                                      self.det_elem_vertical_recon_range_begin)

    def _parse_cone(self, cfg):
        """run_mgfbp.py:495-672, P-matrix branch included."""
        self.cone_beam = cfg.get("ConeBeam", False)
        if not isinstance(self.cone_beam, bool):
            fail("ERROR: ConeBeam must be true or false!")
        self.bool_apply_pmatrix = False
        if not self.cone_beam:
            self.det_elem_height = 0.0
            self.det_offset_vertical = 0.0
            self.img_dim_z = self.det_elem_count_vertical
            self.img_voxel_height = 0.0
            self.img_center_z = 0.0
            return
        k, h = first_key(cfg, "SliceThickness", "DetectorElementHeight")
        if k is None:
            fail("ERROR: Can not find detector element height for cone beam recon! ")
        if not is_num(h):
            fail("ERROR: DetectorElementHeight (SliceThickness) must be a number!")
        if h <= 0:
            fail("ERROR: DetectorElementHeight (SliceThickness) must be positive!")
        self.det_elem_height = h
        k, nz = first_key(cfg, "ImageSliceCount", "ImageDimensionZ")
        if k is None:
            fail("ERROR: Can not find image dimension along Z direction for cone beam recon!")
        if not isinstance(nz, int) or nz <= 0:
            fail("ERROR: ImageDimensionZ (ImageSliceCount) must be a positive integer!")
        self.img_dim_z = nz
        k, vh = first_key(cfg, "VoxelHeight", "ImageSliceThickness")
        if k is None:
            fail("ERROR: Can not find image voxel height for cone beam recon!")
        if not is_num(vh):
            fail("ERROR: VoxelHeight (ImageSliceThickness) must be a number!")
        if vh <= 0:
            fail("ERROR: VoxelHeight (ImageSliceThickness) must be positive!")
        self.img_voxel_height = vh
        if "PMatrixFile" in cfg:
            self._parse_pmatrix(cfg)
        else:
            k, off = first_key(cfg, "SliceOffCenter", "DetectorOffsetVertical")
            if k is None:
                off = 0
                print("Warning: Can not find vertical detector offset for cone beam recon; Using default value 0")
            if not is_num(off):
                fail("ERROR: DetectorOffsetVertical (SliceOffCenter) must be a number!")
            self.det_offset_vertical = off
        if "ImageCenterZ" in cfg:
            self.img_center_z = cfg["ImageCenterZ"]
            if not is_num(self.img_center_z):
                fail("ERROR: ImageCenterZ must be a number!")
            self.img_center_z_auto_set_from_fbp = False
        else:
            center_row = (self.det_elem_vertical_recon_range_end + self.det_elem_vertical_recon_range_begin) / 2
            dist = (center_row - (self.det_elem_count_vertical - 1) / 2) * self.det_elem_height
            if self.first_slice_top_row:
                dist = dist * (-1)
            self.img_center_z = (self.det_offset_vertical + dist) * self.source_isocenter_dis / self.source_det_dis
            self.img_center_z_auto_set_from_fbp = True
            cfg["ImageCenterZ"] = self.img_center_z

    # ------------------------------------------------------------ geometry
    def geometry_spec(self) -> geo.GeometrySpec:
        """The arrays of InitializeArrays (run_mgfbp.py:1244-1259)."""
        u = geo.det_pos_array(self.det_elem_count_horizontal, self.det_elem_count_horizontal,
                              self.positive_u_is_positive_y * self.det_elem_width,
                              (-1) * self.det_offset_horizontal, 0)
        if self.cone_beam:
            v = geo.det_pos_array(self.det_elem_count_vertical, self.det_elem_count_vertical_actual,
                                  self.positive_v_is_positive_z * self.det_elem_height,
                                  self.det_offset_vertical, self.det_elem_vertical_recon_range_begin)
        else:
            v = geo.det_pos_array(self.det_elem_count_vertical, self.det_elem_count_vertical_actual,
                                  0, 0, 0)
        ang = geo.angle_array(self.view_num, self.total_scan_angle, self.img_rot)
        return geo.GeometrySpec(
            img_dim=self.img_dim, img_dim_z=self.img_dim_z, pix=self.img_pix_size,
            vox_h=self.img_voxel_height, n_views=self.view_num, sid=self.source_isocenter_dis,
            sdd=self.source_det_dis, total_scan_angle=self.total_scan_angle, angles=ang, u=u, v=v,
            cx=self.img_center_x, cy=self.img_center_y, cz=self.img_center_z, img_rot=self.img_rot,
            view_mode=self.recon_view_mode, short_scan=bool(self.short_scan), cone=self.cone_beam,
            curved=self.curved_dect, n_v_full=self.det_elem_count_vertical, det_h=self.det_elem_height,
            pmatrix=self.array_pmatrix if self.bool_apply_pmatrix else None)

    def recon_kernel(self) -> np.ndarray | None:
        """InitializeReconKernel (run_mgfbp.py:1261-1279)."""
        n, w = self.det_elem_count_horizontal, self.det_elem_width
        name = self.kernel_name  # AttributeError when no filter key was given
        if name == "HammingFilter":
            return geo.hamming_taps(n, w, self.kernel_param, self.source_det_dis, self.curved_dect,
                                    bool(self.dbt_or_not))
        if name == "GaussianApodizedRamp":
            return geo.gaussian_apodized_ramp_taps(n, w, self.kernel_param, self.source_det_dis,
                                                   self.curved_dect, bool(self.dbt_or_not))
        if name == "SheppLoganFilter":
            return geo.shepp_logan_taps(n, w)
        return None

    def vertical_kernel(self) -> np.ndarray:
        return geo.gaussian_taps(self.det_elem_count_vertical_actual, self.det_elem_height or 1.0,
                                 self.det_elem_vertical_gauss_filter_size)

    # ------------------------------------------------------------ device
    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = Context(self._device)
        return self._ctx

    def plan(self) -> Plan:
        if self._plan is None:
            self._plan = Plan(self.ctx(), self.geometry_spec())
        return self._plan

    def tap_layout(self) -> str:
        """Packed tap layout of the back-projector: ROWS (each element with its
        right neighbour's difference, rows padded by a zero row either side)
        for axial cone beam, QUAD (a cell's four bilinear coefficients) for the
        P-matrix and coronal/sagittal cone paths, PAIR (the same as ROWS,
        row-major over views) for fan beam."""
        if not self.cone_beam:
            return "pair"
        # the P-matrix and coronal/sagittal views run the one-thread-per-voxel
        # kernel, which reads a QUAD cell in one load (ROWS: two)
        return "quad" if self.bool_apply_pmatrix or self.recon_view_mode != 1 else "rows"

    def weight_and_filter(self, sgm: torch.Tensor, view_begin: int | None = None) -> torch.Tensor:
        """BHCorrection + WeightSgm + FilterSinogram (run_mgfbp.py:63-70,
        1282-1292) on a device FP32 sinogram (rows, views, n_u); with
        view_begin, sgm is the block of views [view_begin, view_begin + views)
        of the scan.  Returns the filtered sinogram.  With an MFMA ramp filter
        and no vertical Gaussian the weighting runs inside the filter's input
        pass (one read of the sinogram, the same bits) and sgm is left as
        given; otherwise sgm is weighted in place."""
        ctx, plan = self.ctx(), self.plan()
        bh = self.array_bh_coefficients if self.bool_bh_correction else None
        taps = self.recon_kernel()
        if taps is not None and not self.apply_gauss_vertical and self.filter_method != "f32":
            filt = torch.empty_like(sgm)
            plan.weight_filter(sgm, filt, taps, self.det_elem_width, self.filter_method,
                               0 if view_begin is None else view_begin, bh)
            return filt
        if view_begin is None:
            plan.weight(sgm, bh)
        else:
            plan.weight_views(sgm, view_begin, bh)
        if taps is None:
            # run_mgfbp.py:1283-1285: kernel 'None' filters nothing, the
            # vertical Gaussian included
            return sgm
        src = sgm
        if self.apply_gauss_vertical:
            src = torch.empty_like(sgm)
            ctx.filter_vertical(sgm, src, self.vertical_kernel(), self.det_elem_height)
        filt = torch.empty_like(src)
        ctx.filter_rows(src, filt, taps, self.det_elem_width, self.filter_method)
        return filt

    def fused_filter_pack(self) -> bool:
        """Whether weighting, MFMA ramp filter and the FP32 ROWS repack run as
        one pass (mct_weight_filter_pack): cone beam, FP32 storage, an MFMA
        filter method, no vertical Gaussian, no filtered sinogram to save, and
        a row length the fused epilogue takes."""
        return (self.cone_beam and self.storage == "f32" and self.filter_method != "f32"
                and getattr(self, "kernel_name", None) in ("HammingFilter", "GaussianApodizedRamp",
                                                           "SheppLoganFilter")
                and not self.apply_gauss_vertical
                and not self.save_filtered_sinogram and self.tap_layout() == "rows"
                and self.det_elem_count_vertical_actual >= 2
                and (self.det_elem_count_horizontal - 1) % 64 != 0)

    def reconstruct(self, sgm: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """The device pipeline of MainFunction (run_mgfbp.py:63-81) on a
        resident FP32 sinogram (rows, views, n_u); sgm is weighted in place
        (as the reference mutates its ndarray argument) unless the weighting
        runs inside the ramp filter's input pass (weight_and_filter)."""
        ctx, plan = self.ctx(), self.plan()
        layout = self.tap_layout()
        if self.fused_filter_pack():
            # the MFMA filter writes the back-projector's ROWS taps itself
            # (no FP32 filtered sinogram; the same bits as filter then pack)
            nbytes = _packed_bytes(sgm.shape[0], sgm.shape[1], sgm.shape[2], "f32", "rows")
            packed = torch.empty(nbytes, dtype=torch.uint8, device=sgm.device)
            bh = self.array_bh_coefficients if self.bool_bh_correction else None
            plan.weight_filter_pack(sgm, packed, self.recon_kernel(), self.det_elem_width,
                                    self.filter_method, 0, bh)
            self.last_filtered = None
        else:
            filt = self.weight_and_filter(sgm)
            self.last_filtered = filt
            packed = ctx.pack(filt, self.storage, layout)
        if out is None:
            out = torch.empty((self.img_dim_z, self.img_dim, self.img_dim), dtype=torch.float32,
                              device=sgm.device)
        if (self.autotune and self.tuned is None and self.recon_view_mode == 1
                and not self.bool_apply_pmatrix):
            from .autotune import Autotuner

            self.tuner = self.tuner or Autotuner(ctx)
            self.tuned = self.tuner.tune_backprojector(plan, packed, self.storage, layout, out)
        plan.backproject_fbp(packed, self.storage, layout, out)
        return out

    # ------------------------------------------------------------ files
    def ReadSinogram(self, file):
        """run_mgfbp.py:1189-1228."""
        self.input_path = os.path.join(self.input_dir, file)
        self.output_file = file
        for i in range(int(len(self.output_file_replace) / 2)):
            self.output_file = re.sub(self.output_file_replace[2 * i], self.output_file_replace[2 * i + 1],
                                      self.output_file)
        if self.output_file == file:
            print(f"ERROR: did not file string to replace in '{self.output_file}'")
            sys.exit()
        print("\nLoading %s to RAM ..." % self.input_path)
        if self.output_file_format in ("tif", "tiff"):
            self.output_file = re.sub(".raw", ".tif", self.output_file)
        self.output_path = os.path.join(self.output_dir, self.output_file_prefix + self.output_file)
        self.load_rows()
        return True

    def load_rows(self) -> None:
        """The first view_num views of the file's recon-range rows into
        self.img_sgm (rows, view_num, n_u)."""
        self.img_sgm = np.ascontiguousarray(self.read_file_rows()[:, 0:self.view_num, :])

    def read_file_rows(self) -> np.ndarray:
        """The detector rows of the recon range of self.input_path as
        (rows, sgm_height, n_u) FP32 (run_mgfbp.py:1206-1222): a sinogram file
        is [rows][views][n_u], a post-log image file [views][rows][n_u]."""
        nu, rows = self.det_elem_count_horizontal, self.det_elem_count_vertical_actual
        if self.input_file_form == "sinogram":
            off = self.det_elem_vertical_recon_range_begin * 4 * self.sgm_height * nu
            return np.fromfile(self.input_path, dtype=np.float32, offset=off,
                               count=rows * self.sgm_height * nu).reshape(rows, self.sgm_height, nu)
        off = self.det_elem_vertical_recon_range_begin * nu * 4
        gap = (self.det_elem_vertical_recon_range_begin + self.det_elem_count_vertical - 1
               - self.det_elem_vertical_recon_range_end) * nu * 4
        return imreadRaw(self.input_path, height=rows, width=nu, offset=off, gap=gap,
                         nSlice=self.sgm_height).transpose(1, 0, 2)

    def SaveFilteredSinogram(self):
        if self.save_filtered_sinogram:
            f = self.last_filtered.cpu().numpy().astype(np.float32)
            imwriteTiff(f.transpose(1, 0, 2), self.output_dir + "/" + "sgm_filtered_" + self.output_file)

    def SaveReconImg(self):
        """run_mgfbp.py:1302-1310."""
        if self.convert_to_HU:
            self.img_recon = (self.img_recon / self.water_mu - 1) * 1000
        if self.output_file_format == "raw":
            imwriteRaw(self.img_recon, self.output_path)
        else:
            imwriteTiff(self.img_recon, self.output_path)

    def MainFunction(self):
        """run_mgfbp.py:52-85."""
        self.file_processed_count = 0
        dev = torch.device("cuda", self._device)
        for file in os.listdir(self.input_dir):
            if re.match(self.input_files_pattern, file) and self.ReadSinogram(file):
                print("Reconstructing %s ..." % self.input_path)
                self.file_processed_count += 1
                sgm = torch.from_numpy(self.img_sgm).to(dev)
                vol = self.reconstruct(sgm)
                self.SaveFilteredSinogram()
                self.img_recon = vol.cpu().numpy()
                print("Saving to %s !" % self.output_path)
                self.SaveReconImg()
        return self.img_recon
