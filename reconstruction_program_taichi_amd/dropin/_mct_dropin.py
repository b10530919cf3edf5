"""Names every reference module's ``import *`` exposes besides its own: the
standard-library and NumPy modules the reference imports at top level and
its config/raw-file helpers (run_mgfbp.py:1-16, 1313-1359;
run_mgfpj.py:8-21, 671-729).  Taichi, crip, tqdm and matplotlib are not
part of this build; crip's imwriteRaw/imwriteTiff come from the package.

The directory holding this file is the one a caller puts on sys.path in
place of the reference's; the package itself sits two levels up and is put
on sys.path here when it is not importable already."""
import importlib.util as _ilu
import sys as _sys
from pathlib import Path as _Path

if _ilu.find_spec("reconstruction_program_taichi_amd") is None:
    _sys.path.insert(0, str(_Path(__file__).resolve().parents[2]))
import gc  # noqa: F401
import json  # noqa: F401
import math  # noqa: F401
import os  # noqa: F401
import re  # noqa: F401
import sys  # noqa: F401
import time  # noqa: F401
import warnings  # noqa: F401

import numpy as np  # noqa: F401

from reconstruction_program_taichi_amd.config import (  # noqa: F401
    PI, ReadConfigFile, XuLeastSquareSolution, imaddRaw, imreadRaw, imwriteRaw, imwriteTiff,
    load_jsonc, remove_comments, save_jsonc)
