"""Drop-in for the reference module run_mgfpj (run_mgfpj.py:21-56,
671-729): run_mgfpj, Mgfpj, PI, imaddRaw and the config/raw helpers."""
from _mct_dropin import *  # noqa: F401,F403
from _mct_dropin import PI  # noqa: F401

from reconstruction_program_taichi_amd.fpj import Mgfpj, run_mgfpj  # noqa: F401
