"""View-by-view forward projection: drop-in for the reference's run_mgfpj_v3
entry and Mgfpj_v3 class (run_mgfpj_v3.py:45-391), built on fpj_v2.Mgfpj.

What changes against v2, as the reference computes it:
  * the step count is floor(D / (step * diag)) (run_mgfpj_v3.py:252);
  * the image centre along z may differ per view ("ImageCenterZFile",
    run_mgfpj_v3.py:153-172);
  * the u array is the forward-projection one, (i - (N*os - 1)/2) (-w/os)
    - offset, whose first N entries Mgfbp.InitializeArrays then overwrites
    with the FBP u (run_mgfpj_v3.py:86-90); v is the FBP v of the vertical
    range for the same reason; the angles are regenerated uniformly by
    InitializeArrays, so "ScanAngleFile" is read and checked but does not
    reach the projection;
  * binning and noise run over every detector row (run_mgfpj_v3.py:112-115),
    so rows outside the vertical range carry noise around 0.
The projection is one launch over all rows and views, not one per view.
"""
from __future__ import annotations

import numpy as np

from . import geometry as geo
from .config import PI, ReadConfigFile, fail
from .fpj_v2 import Mgfpj, _run


def run_mgfpj_v3(file_path):
    """run_mgfpj_v3.py:45-78."""
    return _run(file_path, Mgfpj_v3, "Performing FPJ from MandoCT (MI355X HIP, ver 3.0) ...")


def _value_list(path, name, n):
    d = ReadConfigFile(path)
    if "Value" not in d:
        fail(f"ERROR: {name} has no member named 'Value'!")
    v = d["Value"]
    if not isinstance(v, list):
        fail(f"ERROR: {name}.Value is not an array")
    if len(v) != n:
        fail(f"ERROR: view number is {n:d} while {name} has {len(v):d} elements!")
    return v


class Mgfpj_v3(Mgfpj):
    """run_mgfpj_v3.py:81."""

    count_mode = 1
    noise_all_rows = True

    def __init__(self, config_dict, device: int = 0, ctx=None):
        super().__init__(config_dict, device, ctx)
        cfg = config_dict
        if "ScanAngleFile" in cfg:
            v = _value_list(cfg["ScanAngleFile"], "ScanAngleFile", self.view_num)
            self.array_angle = (np.array(v, dtype=np.float32) / np.float32(180.0) * np.float32(PI)
                                + np.float32(self.img_rot)).astype(np.float32)
            self.bool_uneven_scan_angle = 1
            print("--Scan Angles From File (Original total scan angle value is discarded!)")
        else:
            self.bool_uneven_scan_angle = False
        if "ImageCenterZFile" in cfg:
            v = _value_list(cfg["ImageCenterZFile"], "ImageCenterZFile", self.view_num)
            self.array_img_center_z = np.array(v, dtype=np.float32)
            self.bool_image_center_z_from_file = True
            print("--Image Center Z From File (defalt value is discarded)")
        else:
            self.array_img_center_z = np.full(self.view_num, self.img_center_z, dtype=np.float32)
            self.bool_image_center_z_from_file = False

    def _u_array(self) -> np.ndarray:
        n, os_ = self.det_elem_count_horizontal, self.oversample_size
        u = geo.det_pos_array(n * os_, n * os_, -self.det_elem_width / os_, -self.det_offset_horizontal, 0)
        u[:n] = geo.det_pos_array(n, n, self.positive_u_is_positive_y * self.det_elem_width,
                                  (-1) * self.det_offset_horizontal, 0)
        return u

    def _center_z(self):
        return self.array_img_center_z
