"""Drop-in for the reference module run_mgfpj_v3 (run_mgfpj_v3.py:40-81):
everything run_mgfbp and run_mgfpj_v2 export, plus run_mgfpj_v3 and
Mgfpj_v3."""
from run_mgfbp import *  # noqa: F401,F403
from run_mgfpj_v2 import *  # noqa: F401,F403

from reconstruction_program_taichi_amd.fpj_v3 import Mgfpj_v3, run_mgfpj_v3  # noqa: F401
