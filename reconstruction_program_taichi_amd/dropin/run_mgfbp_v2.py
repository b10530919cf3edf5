"""Drop-in for the reference module run_mgfbp_v2 (run_mgfbp_v2.py:20-61):
everything run_mgfbp exports, plus run_mgfbp_v2 and Mgfbp_v2 (view-block
streaming FBP)."""
from run_mgfbp import *  # noqa: F401,F403

from reconstruction_program_taichi_amd.fbp_v2 import Mgfbp_v2, run_mgfbp_v2  # noqa: F401
