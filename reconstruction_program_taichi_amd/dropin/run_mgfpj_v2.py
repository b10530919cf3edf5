"""Drop-in for the reference module run_mgfpj_v2 (run_mgfpj_v2.py:31-70):
everything run_mgfbp exports, plus run_mgfpj_v2 and its Mgfpj, the forward
projector built on Mgfbp's geometry."""
from run_mgfbp import *  # noqa: F401,F403

from reconstruction_program_taichi_amd.fpj_v2 import Mgfpj, run_mgfpj_v2  # noqa: F401
