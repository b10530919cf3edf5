"""Drop-in for the reference module run_mgfbp (run_mgfbp.py:16-51,
1313-1359): ``from run_mgfbp import *`` gives run_mgfbp, Mgfbp, PI and the
config/raw helpers, backed by the HIP kernels of libmct.so."""
from _mct_dropin import *  # noqa: F401,F403
from _mct_dropin import PI  # noqa: F401

from reconstruction_program_taichi_amd.fbp import Mgfbp, run_mgfbp  # noqa: F401
