"""Drop-in for the reference module run_mgfbp_iterative_recon_IRN_CGLS
(run_mgfbp_iterative_recon_IRN_CGLS.py:9-53): everything run_mgfbp and
run_mgfpj_v3 export, plus run_mgfbp_ir (FBP seed, then IRN-CGLS-TV) and
Mgfbp_ir."""
from run_mgfbp import *  # noqa: F401,F403
from run_mgfpj_v3 import *  # noqa: F401,F403

from reconstruction_program_taichi_amd.ir import Mgfbp_ir, run_mgfbp_ir  # noqa: F401
