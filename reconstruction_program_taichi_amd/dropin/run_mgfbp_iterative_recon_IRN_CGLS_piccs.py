"""Drop-in for the reference module run_mgfbp_iterative_recon_IRN_CGLS_piccs
(…_piccs.py:8-49): everything the IRN-CGLS module exports, with
run_mgfbp_ir rebound to the PICCS entry (the reference module redefines it),
plus Mgfbp_ir_piccs."""
from run_mgfbp import *  # noqa: F401,F403
from run_mgfbp_iterative_recon_IRN_CGLS import *  # noqa: F401,F403

from reconstruction_program_taichi_amd.ir import Mgfbp_ir_piccs  # noqa: F401
from reconstruction_program_taichi_amd.ir import run_mgfbp_ir_piccs as run_mgfbp_ir  # noqa: F401
