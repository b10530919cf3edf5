"""MI355X-native reconstruction hot path of MandoCT-Taichi
(SEU-CT-Recon/Reconstruction_program_taichi): sinogram weighting, ramp
filter, pixel-driven fan/cone back-projection, bilinear ray-driven forward
projection and the IRN-CGLS / TV / PICCS loop, as hand-written HIP kernels
for gfx950 behind the C ABI of include/mct.h (libmct.so), driven by
config-JSONC drop-ins for run_mgfbp / run_mgfpj / run_mgfbp_ir.
"""
from .config import (PI, ReadConfigFile, XuLeastSquareSolution, imreadRaw, load_jsonc,
                     remove_comments, save_jsonc)

__all__ = ["PI", "ReadConfigFile", "XuLeastSquareSolution", "imreadRaw", "load_jsonc",
           "remove_comments", "save_jsonc", "run_mgfbp", "Mgfbp", "run_mgfpj", "Mgfpj",
           "run_mgfpj_v2", "run_mgfpj_v3", "Mgfpj_v3", "run_mgfbp_ir", "Mgfbp_ir", "Mgfbp_ir_piccs",
           "run_mgfbp_v2", "Mgfbp_v2", "run_mgfbp_helical", "Mgfbp_helical"]


def __getattr__(name):
    if name in ("run_mgfbp", "Mgfbp"):
        from . import fbp
        return getattr(fbp, name)
    if name in ("run_mgfpj", "Mgfpj"):
        from . import fpj
        return getattr(fpj, name)
    if name == "run_mgfpj_v2":
        from . import fpj_v2
        return fpj_v2.run_mgfpj_v2
    if name in ("run_mgfpj_v3", "Mgfpj_v3"):
        from . import fpj_v3
        return getattr(fpj_v3, name)
    if name in ("run_mgfbp_v2", "Mgfbp_v2"):
        from . import fbp_v2
        return getattr(fbp_v2, name)
    if name in ("run_mgfbp_helical", "Mgfbp_helical"):
        from . import fbp_helical
        return getattr(fbp_helical, name)
    if name in ("run_mgfbp_ir", "Mgfbp_ir", "Mgfbp_ir_piccs"):
        from . import ir
        return getattr(ir, name)
    raise AttributeError(name)
