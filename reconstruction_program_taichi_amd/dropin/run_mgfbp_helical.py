"""Drop-in for the reference module run_mgfbp_helical
(run_mgfbp_helical.py:21-58): everything run_mgfbp exports, imaddRaw,
run_mgfbp_helical and Mgfbp_helical."""
from run_mgfbp import *  # noqa: F401,F403
from run_mgfpj import imaddRaw  # noqa: F401

from reconstruction_program_taichi_amd.fbp_helical import Mgfbp_helical, run_mgfbp_helical  # noqa: F401
