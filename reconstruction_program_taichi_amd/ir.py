"""IRN-CGLS with TV (and PICCS): drop-ins for the reference's run_mgfbp_ir
entries and the Mgfbp_ir / Mgfbp_ir_piccs classes
(run_mgfbp_iterative_recon_IRN_CGLS.py:14-286,
run_mgfbp_iterative_recon_IRN_CGLS_piccs.py:13-161).

The reference runs the CG recursion in float32 NumPy on the host and, in
every iteration, launches a forward and a back-projection per view with two
host-to-device copies each.  Here every operand stays on the device: one
launch forward-projects all views (rays outside the effective-ray map left
0), one launch back-projects all views with the truncation mask, the TV
terms are device kernels, and the CG scalars <r,r>, <d,q>, alpha and beta
live in FP64 on the device (no host round trip inside an iteration).

System solved per reweighting pass (TikhonovSol):
  (P^T P + c1 D^T W D + c2 D^T W_p D) x = P^T b + c2 D^T W_p D prior
with c1 = Lambda * ratio, c2 = 0 (IRN-TV), or c1 = Lambda (1 - Alpha) ratio,
c2 = Lambda * Alpha * ratio (PICCS); ratio = #sinogram / #image samples.

Additive config key (no reference counterpart):
  "TVCoupleSlices": true|false   couple slices through the z difference when
                                 Nz > 2 (true, the reference's behaviour) or
                                 treat the slices as independent images
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
from .config import PI, ReadConfigFile, fail, imreadRaw, imwriteRaw, imwriteTiff, is_num
from .fbp import Mgfbp


def run_mgfbp_ir(file_path):
    """run_mgfbp_iterative_recon_IRN_CGLS.py:14-49: FBP seed, then IRN-CGLS-TV;
    returns the last reconstruction (HU when WaterMu is given)."""
    print("Performing Iterative Recon from MandoCT (MI355X HIP) ...")
    if not os.path.exists(file_path):
        print(f"ERROR: Config File {file_path} does not exist!")
        sys.exit()
    config_dict = ReadConfigFile(file_path)
    print("Generating seed image from FBP ...")
    fbp = Mgfbp(config_dict)
    seed = fbp.MainFunction()
    start = time.time()
    print("\nPerform Iterative Recon ...")
    ir = Mgfbp_ir(config_dict, ctx=fbp.ctx())
    img = ir.MainFunction(seed)
    _report(ir, time.time() - start)
    return img


def run_mgfbp_ir_piccs(file_path):
    """run_mgfbp_iterative_recon_IRN_CGLS_piccs.py:13-45: PICCS from the prior
    image; returns the last reconstruction."""
    print("Performing Iterative Recon from MandoCT (MI355X HIP, PICCS) ...")
    if not os.path.exists(file_path):
        print(f"ERROR: Config File {file_path} does not exist!")
        sys.exit()
    config_dict = ReadConfigFile(file_path)
    start = time.time()
    ir = Mgfbp_ir_piccs(config_dict)
    img = ir.MainFunction()
    _report(ir, time.time() - start)
    return img


def _report(ir, elapsed):
    if ir.file_processed_count > 0:
        per = elapsed / max(ir.num_iter_runned, 1) / ir.file_processed_count
        print(f"\nA total of {ir.file_processed_count:d} file(s) are reconstructed!")
        print(f"Time cost: {elapsed:.3} sec ({per:.3} sec per iteration). \n")
    else:
        print(f"\nWarning: Did not find files like {ir.input_files_pattern:s} in {ir.input_dir:s}.")
        print("No images are reconstructed!")
    gc.collect()


class Mgfbp_ir(Mgfbp):
    """run_mgfbp_iterative_recon_IRN_CGLS.py:53 (Mgfbp_ir, ver 0.3)."""

    def __init__(self, config_dict, device: int = 0, ctx=None):
        super().__init__(config_dict, device, ctx)
        cfg = config_dict
        # Mgfpj(v2) keys (run_mgfpj_v2.py:113-214)
        self.oversample_size = cfg.get("OversampleSize", 1)
        if not is_num(self.oversample_size) or self.oversample_size % 1 != 0 or self.oversample_size < 0:
            fail("ERROR: OversampleSize must be a positive integer!")
        if self.oversample_size != 1:
            fail("ERROR: OversampleSize must be 1 for iterative reconstruction!")
        self.fpj_step_size = cfg.get("ForwardProjectionStepSize", 0.2)
        if self.bool_apply_pmatrix:  # run_mgfpj_v2.py:211-214
            self.matrix_A_each_view, self.x_s_each_view = pmx.matrix_a_and_source(self.array_pmatrix,
                                                                                  self.view_num)
        if "HelicalPitch" in cfg:
            self.helical_scan, self.helical_pitch = True, cfg["HelicalPitch"]
        else:
            self.helical_scan, self.helical_pitch = False, 0.0
        if self.helical_scan and ("ImageCenterZ" not in cfg or getattr(self, "img_center_z_auto_set_from_fbp", False)):
            self.img_center_z = self.img_voxel_height * (self.img_dim_z - 1) / 2.0 * np.sign(self.helical_pitch)
        if "StartAngle" in cfg:
            # run_mgfpj_v2.py:203-207 keeps StartAngle in degrees
            self.img_rot = cfg["StartAngle"]
            if not is_num(self.img_rot):
                fail("ERROR: StartAngle must be a number!")
        # Mgfpj_v3 keys (run_mgfpj_v3.py:131-172)
        self.array_angle = None
        if "ScanAngleFile" in cfg:
            d = ReadConfigFile(cfg["ScanAngleFile"])
            if "Value" not in d:
                fail("ERROR: ScanAngleFile has no member named 'Value'!")
            if not isinstance(d["Value"], list):
                fail("ERROR: ScanAngleFile.Value is not an array")
            if len(d["Value"]) != self.view_num:
                fail(f"ERROR: view number is {self.view_num:d} while ScanAngleFile has {len(d['Value']):d} elements!")
            self.array_angle = (np.array(d["Value"], dtype=np.float32) / np.float32(180.0) * np.float32(PI)
                                + np.float32(self.img_rot)).astype(np.float32)
        self.array_img_center_z = None
        if "ImageCenterZFile" in cfg:
            d = ReadConfigFile(cfg["ImageCenterZFile"])
            if "Value" not in d:
                fail("ERROR: ImageCenterZFile has no member named 'Value'!")
            if not isinstance(d["Value"], list):
                fail("ERROR: ImageCenterZFile.Value is not an array")
            if len(d["Value"]) != self.view_num:
                fail(f"ERROR: view number is {self.view_num:d} while ImageCenterZFile has {len(d['Value']):d} elements!")
            self.array_img_center_z = np.array(d["Value"], dtype=np.float32)
        # Mgfbp_ir keys (run_mgfbp_iterative_recon_IRN_CGLS.py:57-91)
        self.coef_lambda = self._num_key(cfg, "Lambda", 1e-5, "ERROR: Lambda must be a positive number!")
        self.beta_tv = self._num_key(cfg, "BetaTV", 1e-6, "ERROR: BetaTV must be a positive number!")
        self.num_irn_iter = self._int_key(cfg, "NumberOfIRNIterations", 15,
                                          "ERROR: NumberOfIRNIterations must be a positive integer!")
        self.num_iter = self._int_key(cfg, "NumberOfIterations", 100,
                                      "ERROR: NumberOfIterations must be a positive integer!")
        self.tv_couple_slices = cfg.get("TVCoupleSlices", True)
        if not isinstance(self.tv_couple_slices, bool):
            fail("ERROR: TVCoupleSlices must be true or false!")
        if self.det_elem_count_vertical_actual != self.det_elem_count_vertical:
            fail("ERROR: iterative reconstruction needs the full detector row range!")
        self.sgm_total_pixel_count = self.det_elem_count_vertical_actual * self.view_num * self.det_elem_count_horizontal
        self.img_total_pixel_count = self.img_dim_z * self.img_dim * self.img_dim
        self.pixel_count_ratio = float(self.sgm_total_pixel_count / self.img_total_pixel_count)
        self.loss = np.empty(0)
        self.num_iter_runned = 0
        self._ir_plan = None

    @staticmethod
    def _num_key(cfg, key, default, msg):
        if key not in cfg:
            print(f"Warning: Did not find {key}! Use default value {default}. ")
            return default
        v = cfg[key]
        if not is_num(v) or v < 0.0:
            fail(msg)
        return v

    @staticmethod
    def _int_key(cfg, key, default, msg):
        if key not in cfg:
            print(f"Warning: Did not find {key}! Use default value {default}. ")
            return default
        v = cfg[key]
        if not isinstance(v, int) or v < 0:
            fail(msg)
        return v

    # ------------------------------------------------------------ geometry
    def ir_geometry_spec(self) -> geo.GeometrySpec:
        """Arrays of Mgfbp_ir.MainFunction (run_mgfbp_iterative_recon_IRN_CGLS.py:140-145):
        FBP angles (or ScanAngleFile), FPJ-convention u and v."""
        if self.array_angle is not None:
            ang = self.array_angle
        else:
            ang = geo.angle_array(self.view_num, self.total_scan_angle, self.img_rot)
        v = geo.det_pos_array(self.det_elem_count_vertical, self.det_elem_count_vertical,
                              -self.det_elem_height, self.det_offset_vertical)
        u = geo.det_pos_array(self.det_elem_count_horizontal, self.det_elem_count_horizontal,
                              -self.det_elem_width, -self.det_offset_horizontal)
        return geo.GeometrySpec(
            img_dim=self.img_dim, img_dim_z=self.img_dim_z, pix=self.img_pix_size,
            vox_h=self.img_voxel_height, n_views=self.view_num, sid=self.source_isocenter_dis,
            sdd=self.source_det_dis, total_scan_angle=self.total_scan_angle, angles=ang, u=u, v=v,
            cx=self.img_center_x, cy=self.img_center_y, cz=self.img_center_z, img_rot=self.img_rot,
            short_scan=bool(self.short_scan), cone=self.cone_beam, curved=self.curved_dect,
            n_v_full=self.det_elem_count_vertical, det_h=self.det_elem_height,
            fpj_step=self.fpj_step_size, fpj_count_mode=0, helical=self.helical_scan,
            helical_pitch=self.helical_pitch, center_z=self.array_img_center_z,
            pmatrix=self.array_pmatrix if self.bool_apply_pmatrix else None,
            matrix_a=self.matrix_A_each_view if self.bool_apply_pmatrix else None,
            x_s=self.x_s_each_view if self.bool_apply_pmatrix else None,
            view_mode=self.recon_view_mode)

    def ir_plan(self):
        from .device import Plan

        if self._ir_plan is None:
            self._ir_plan = Plan(self.ctx(), self.ir_geometry_spec())
        return self._ir_plan

    # ------------------------------------------------------------ device solver
    def _buffers(self, dev):
        vshape = (self.img_dim_z, self.img_dim, self.img_dim)
        sshape = (self.det_elem_count_vertical_actual, self.view_num, self.det_elem_count_horizontal)
        B = {k: torch.empty(vshape, dtype=torch.float32, device=dev)
             for k in ("x", "bpb", "d", "r", "q", "t", "w", "wp")}
        B["fp"] = torch.empty(sshape, dtype=torch.float32, device=dev)
        B["map"] = torch.empty(sshape, dtype=torch.float32, device=dev)
        B["s"] = torch.zeros(4, dtype=torch.float64, device=dev)
        B["l"] = torch.zeros(4, dtype=torch.float64, device=dev)
        return B

    def fpbp(self, x, out, B):
        """ForwardProjectionAndBackProjection (:288-316): masked FP of all views,
        then the truncation-masked BP of all views."""
        plan = self.ir_plan()
        plan.forward_project(x, B["fp"], B["map"], 1)
        plan.backproject_ir(B["fp"], None, out)
        self.num_iter_runned += 1
        return out

    def apply_A(self, x, out, B, c1, c2):
        """out = P^T P x + c1 D^T W D x + c2 D^T W_p D x."""
        ctx, cz = self.ctx(), self.tv_couple_slices
        self.fpbp(x, B["t"], B)
        ctx.dtwd(x, B["w"], c1, B["t"], out, cz)
        if c2 != 0.0:
            ctx.dtwd(x, B["wp"], c2, out, out, cz)
        return out

    def tikhonov_sol(self, B, c1, c2, prior=None):
        """TikhonovSol (:188-232): CG on the normal equations from B['x'];
        scalars in FP64 on the device."""
        ctx = self.ctx()
        x, d, r, q, s = B["x"], B["d"], B["r"], B["q"], B["s"]
        self.apply_A(x, q, B, c1, c2)
        ctx.axpby(1.0, B["bpb"], -1.0, q, d)  # d = P^T b - A x
        if c2 != 0.0 and prior is not None:
            ctx.dtwd(prior, B["wp"], c2, d, d, self.tv_couple_slices)
        r.copy_(d)
        for _ in range(self.num_iter):
            self.apply_A(d, q, B, c1, c2)
            ctx.dot(r, r, s[0:1])
            ctx.dot(d, q, s[1:2])
            ctx.cg_update(x, r, d, q, s)
            ctx.cg_direction(d, r, s)
        return x

    def loss_value(self, B, sgm, prior=None, alpha=0.0):
        """LossValCalc (:184-186; PICCS :158-161), FP64."""
        ctx, plan, cz = self.ctx(), self.ir_plan(), self.tv_couple_slices
        plan.forward_project(B["x"], B["fp"], B["map"], 1)
        l = B["l"]
        ctx.sq_diff(B["fp"], sgm, B["map"], l[0:1])
        ctx.tv_value(B["x"], None, self.beta_tv, l[1:2], cz)
        if prior is not None:
            ctx.tv_value(B["x"], prior, self.beta_tv, l[2:3], cz)
        lh = l.cpu().numpy()
        val = 0.5 * lh[0] / self.sgm_total_pixel_count
        if prior is None:
            return val + self.coef_lambda * lh[1] / self.img_total_pixel_count
        return (val + self.coef_lambda * (1 - alpha) * lh[1] / self.img_total_pixel_count
                + self.coef_lambda * alpha * lh[2] / self.img_total_pixel_count)

    def solve(self, sgm: torch.Tensor, x0: torch.Tensor) -> torch.Tensor:
        """MainFunction's per-file body (:157-171) on resident data (mu units)."""
        ctx, plan = self.ctx(), self.ir_plan()
        B = self._buffers(sgm.device)
        plan.ray_mask(B["map"])
        plan.backproject_ir(sgm, B["map"], B["bpb"])
        B["x"].copy_(x0)
        c1 = self.coef_lambda * self.pixel_count_ratio
        B["w"].fill_(1.0)
        self.tikhonov_sol(B, c1, 0.0)
        self.loss = np.append(self.loss, self.loss_value(B, sgm))
        for _ in range(self.num_irn_iter):
            ctx.tv_weight(B["x"], None, self.beta_tv, B["w"], self.tv_couple_slices)
            self.tikhonov_sol(B, c1, 0.0)
            self.loss = np.append(self.loss, self.loss_value(B, sgm))
        self._last_buffers = B
        return B["x"]

    def SaveReconImg(self):
        img = self.img_recon
        if self.convert_to_HU:
            img = (img / self.water_mu - 1) * 1000
        if self.output_file_format == "raw":
            imwriteRaw(img, self.output_path)
        else:
            imwriteTiff(img, self.output_path)

    def MainFunction(self, img_recon_seed):
        """run_mgfbp_iterative_recon_IRN_CGLS.py:138-176."""
        self.file_processed_count = 0
        dev = torch.device("cuda", self._device)
        for file in os.listdir(self.input_dir):
            if re.match(self.input_files_pattern, file) and self.ReadSinogram(file):
                self.file_processed_count += 1
                print("Reconstructing %s ..." % self.input_path)
                if self.convert_to_HU:
                    img_recon_seed = (img_recon_seed / 1000 + 1) * self.water_mu
                sgm = torch.from_numpy(self.img_sgm).to(dev)
                x0 = torch.from_numpy(np.ascontiguousarray(img_recon_seed, dtype=np.float32)).to(dev)
                self.img_recon = self.solve(sgm, x0).cpu().numpy()
                self.img_x = self.img_recon
                print("\nSaving to %s !" % self.output_path)
                self.SaveReconImg()
                if self.convert_to_HU:
                    self.img_recon = (self.img_recon / self.water_mu - 1) * 1000
        return self.img_recon


class Mgfbp_ir_piccs(Mgfbp_ir):
    """run_mgfbp_iterative_recon_IRN_CGLS_piccs.py:49: prior-image constrained
    CS; the CG starts from the prior image."""

    def __init__(self, config_dict, device: int = 0, ctx=None):
        super().__init__(config_dict, device, ctx)
        cfg = config_dict
        if "Alpha" in cfg:
            self.coef_alpha = cfg["Alpha"]
            if not is_num(self.coef_alpha) or self.coef_alpha < 0.0:
                fail("ERROR: Alpha must be a positive number!")
        else:
            self.coef_alpha = 1e-5
            print("Warning: Did not find Alpha! Use default value 1e-5. ")
        if "PriorImageFile" not in cfg:
            fail("ERROR: Can not find PriorImageFile!")
        prior = imreadRaw(cfg["PriorImageFile"], width=self.img_dim, height=self.img_dim, nSlice=self.img_dim_z)
        prior = prior.reshape((self.img_dim_z, self.img_dim, self.img_dim))
        if self.convert_to_HU:
            prior = (prior / 1000 + 1) * self.water_mu
        self.img_prior = np.ascontiguousarray(prior, dtype=np.float32)

    def solve(self, sgm: torch.Tensor, prior: torch.Tensor) -> torch.Tensor:
        """MainFunction's per-file body (piccs.py:83-107)."""
        ctx, plan, cz = self.ctx(), self.ir_plan(), self.tv_couple_slices
        B = self._buffers(sgm.device)
        plan.ray_mask(B["map"])
        plan.backproject_ir(sgm, B["map"], B["bpb"])
        B["x"].copy_(prior)
        a = self.coef_alpha
        c1 = self.coef_lambda * (1 - a) * self.pixel_count_ratio
        c2 = self.coef_lambda * a * self.pixel_count_ratio
        B["w"].fill_(1.0)
        B["wp"].fill_(1.0)
        self.tikhonov_sol(B, c1, c2, prior)
        self.loss = np.append(self.loss, self.loss_value(B, sgm, prior, a))
        for _ in range(self.num_irn_iter):
            ctx.tv_weight(B["x"], None, self.beta_tv, B["w"], cz)
            ctx.tv_weight(B["x"], prior, self.beta_tv, B["wp"], cz)
            self.tikhonov_sol(B, c1, c2, prior)
            self.loss = np.append(self.loss, self.loss_value(B, sgm, prior, a))
        self._last_buffers = B
        return B["x"]

    def MainFunction(self):
        """piccs.py:70-110."""
        self.file_processed_count = 0
        dev = torch.device("cuda", self._device)
        for file in os.listdir(self.input_dir):
            if re.match(self.input_files_pattern, file) and self.ReadSinogram(file):
                self.file_processed_count += 1
                sgm = torch.from_numpy(self.img_sgm).to(dev)
                prior = torch.from_numpy(self.img_prior).to(dev)
                self.img_recon = self.solve(sgm, prior).cpu().numpy()
                self.img_x = self.img_recon
                self.SaveReconImg()
                if self.convert_to_HU:
                    self.img_recon = (self.img_recon / self.water_mu - 1) * 1000
        return self.img_recon
