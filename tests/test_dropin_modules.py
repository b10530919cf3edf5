"""The reference-named drop-in modules (reconstruction_program_taichi_amd/
dropin/): with that directory on sys.path in place of the reference's, an
unchanged caller's ``from run_mgfbp import *`` (and the other seven) gets the
names the reference module gives it (run_mgfbp.py:1-16, 18, 51, 1313-1359;
run_mgfpj.py:21-56, 671-729; run_mgfpj_v2.py:31-70; run_mgfpj_v3.py:40-81;
run_mgfbp_v2.py:20-61; run_mgfbp_helical.py:17-58;
run_mgfbp_iterative_recon_IRN_CGLS.py:5-53; …_piccs.py:4-49), bound to this
package's implementations.  Checked in a fresh interpreter so that nothing
already imported helps.
"""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
DROPIN = ROOT / "reconstruction_program_taichi_amd" / "dropin"

HELPERS = ["PI", "ReadConfigFile", "load_jsonc", "save_jsonc", "remove_comments", "imreadRaw",
           "XuLeastSquareSolution", "imwriteRaw", "imwriteTiff", "np", "os", "sys", "re", "json",
           "time", "gc"]
# module -> {exported name: "package_module.attribute" it must be}
EXPECT = {
    "run_mgfbp": {"run_mgfbp": "fbp.run_mgfbp", "Mgfbp": "fbp.Mgfbp"},
    "run_mgfpj": {"run_mgfpj": "fpj.run_mgfpj", "Mgfpj": "fpj.Mgfpj", "imaddRaw": "config.imaddRaw"},
    "run_mgfpj_v2": {"run_mgfpj_v2": "fpj_v2.run_mgfpj_v2", "Mgfpj": "fpj_v2.Mgfpj",
                     "run_mgfbp": "fbp.run_mgfbp", "Mgfbp": "fbp.Mgfbp"},
    "run_mgfpj_v3": {"run_mgfpj_v3": "fpj_v3.run_mgfpj_v3", "Mgfpj_v3": "fpj_v3.Mgfpj_v3",
                     "run_mgfpj_v2": "fpj_v2.run_mgfpj_v2", "Mgfpj": "fpj_v2.Mgfpj",
                     "Mgfbp": "fbp.Mgfbp"},
    "run_mgfbp_v2": {"run_mgfbp_v2": "fbp_v2.run_mgfbp_v2", "Mgfbp_v2": "fbp_v2.Mgfbp_v2",
                     "run_mgfbp": "fbp.run_mgfbp"},
    "run_mgfbp_helical": {"run_mgfbp_helical": "fbp_helical.run_mgfbp_helical",
                          "Mgfbp_helical": "fbp_helical.Mgfbp_helical", "imaddRaw": "config.imaddRaw",
                          "Mgfbp": "fbp.Mgfbp"},
    "run_mgfbp_iterative_recon_IRN_CGLS": {
        "run_mgfbp_ir": "ir.run_mgfbp_ir", "Mgfbp_ir": "ir.Mgfbp_ir", "Mgfpj_v3": "fpj_v3.Mgfpj_v3",
        "run_mgfpj_v3": "fpj_v3.run_mgfpj_v3", "run_mgfbp": "fbp.run_mgfbp"},
    "run_mgfbp_iterative_recon_IRN_CGLS_piccs": {
        "run_mgfbp_ir": "ir.run_mgfbp_ir_piccs", "Mgfbp_ir_piccs": "ir.Mgfbp_ir_piccs",
        "Mgfbp_ir": "ir.Mgfbp_ir", "Mgfbp": "fbp.Mgfbp"},
}

PROBE = r"""
import importlib, json, sys
sys.path.insert(0, sys.argv[1])
expect, helpers = json.loads(sys.argv[2]), json.loads(sys.argv[3])
report = {}
for mod, names in expect.items():
    ns = {}
    exec(f"from {mod} import *", ns)
    bad = [h for h in helpers if h not in ns]
    for name, target in names.items():
        pkg_mod, attr = target.split(".")
        obj = getattr(importlib.import_module("reconstruction_program_taichi_amd." + pkg_mod), attr)
        if ns.get(name) is not obj:
            bad.append(name)
    if ns.get("PI") != 3.1415926536:
        bad.append("PI value")
    report[mod] = bad
print(json.dumps(report))
"""


def test_reference_named_modules_export_the_reference_names():
    r = subprocess.run([sys.executable, "-c", PROBE, str(DROPIN), json.dumps(EXPECT), json.dumps(HELPERS)],
                       capture_output=True, text=True, cwd="/", timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    report = json.loads(r.stdout.strip().splitlines()[-1])
    assert report == {m: [] for m in EXPECT}, report


def test_xu_least_squares_and_imaddraw(tmp_path):
    from reconstruction_program_taichi_amd.config import XuLeastSquareSolution, imaddRaw, imreadRaw
    rng = np.random.default_rng(0)
    B = rng.standard_normal((9, 3))
    y = rng.standard_normal(9)
    assert np.allclose(XuLeastSquareSolution(B, y), np.linalg.lstsq(B, y, rcond=None)[0])
    p = tmp_path / "a.raw"
    a = rng.standard_normal((2, 3, 4)).astype(np.float32)
    imaddRaw(a[0], str(p), idx=0)
    imaddRaw(a[1], str(p))
    assert np.array_equal(imreadRaw(str(p), 3, 4, nSlice=2), a)
    imaddRaw(a[0], str(p), dtype=np.float64, idx=0)
    assert np.array_equal(np.fromfile(p, np.float64).reshape(3, 4), a[0].astype(np.float64))

