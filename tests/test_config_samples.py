"""The reference's own config files as fixtures (tests/golden/configs, copied
unchanged from the reference's top level and modified_config/).

* config_mgfbp_sample.jsonc: the documented FBP sample.  It parses, and its
  DetectorElementVerticalReconRange [500, 700] lies outside its 400 detector
  rows, so Mgfbp exits with the reference's message (run_mgfbp.py:274-295).
* config_mgfpj_sample.jsonc: the FP sample has no comma before its last key
  (lines 83-85), so the JSON parser rejects it, in the reference as here.
* config_mgfbp_temp.jsonc -> modified_config/config_mgfbp.jsonc: the
  reference's SaveModifiedConfigFolder output for that input
  (run_mgfbp.py:644-663, 690-696): the key removed, ImageCenterZ added,
  json.dump formatting.  Mgfbp must write the same bytes.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from reconstruction_program_taichi_amd import geometry as geo
from reconstruction_program_taichi_amd.config import PI, load_jsonc
from reconstruction_program_taichi_amd.fbp import Mgfbp

GOLD = Path(__file__).resolve().parent / "golden" / "configs"


def test_mgfbp_sample_parses_and_exits_on_row_range(tmp_path, monkeypatch, capsys):
    cfg = load_jsonc(str(GOLD / "config_mgfbp_sample.jsonc"))
    assert cfg["DetectorElementVerticalReconRange"] == [500, 700]
    assert cfg["ReconViewMode"] == "Coronal"
    assert cfg["BeamHardeningCorrectionCoefficients"] == [1, 0.03, 0.02]
    assert cfg["InputDir"] == "./sgm" and cfg["HammingFilter"] == 1 and cfg["DBTMode"] is True
    monkeypatch.chdir(tmp_path)
    with pytest.raises(SystemExit):
        Mgfbp(cfg)
    assert "ERROR: Out of detector row range!" in capsys.readouterr().out


def test_mgfpj_sample_is_rejected_by_the_json_parser():
    with pytest.raises(json.JSONDecodeError) as e:
        load_jsonc(str(GOLD / "config_mgfpj_sample.jsonc"))
    assert e.value.msg == "Expecting ',' delimiter"
    assert e.value.doc[e.value.pos:].lstrip().startswith('"OutputFileForm"')


def test_modified_config_written_as_the_reference_writes_it(tmp_path, monkeypatch):
    cfg = load_jsonc(str(GOLD / "config_mgfbp_temp.jsonc"))
    monkeypatch.chdir(tmp_path)
    fbp = Mgfbp(cfg)
    written = (tmp_path / "modified_config" / "config_mgfbp.jsonc").read_bytes()
    assert written == (GOLD / "modified_config_mgfbp.jsonc").read_bytes()
    # the derived geometry of this 4-turn (1440 degree) helical-style cone scan
    assert fbp.view_num == 4320 and fbp.cone_beam and fbp.img_dim == 256 and fbp.img_dim_z == 256
    assert fbp.total_scan_angle == 1440 / 180.0 * PI and fbp.short_scan == 0
    assert geo.div_factor(fbp.total_scan_angle, False) == np.float32(1) / np.float32(8)
    assert fbp.first_slice_top_row and fbp.positive_v_is_positive_z == -1
    assert fbp.kernel_name == "GaussianApodizedRamp" and fbp.kernel_param == 1
    assert fbp.img_center_z == 0.0
    spec = fbp.geometry_spec()
    assert spec.n_u == 1024 and spec.n_v == 64 and spec.n_views == 4320
    # u decreases across the detector and carries the -5.69 mm offset
    # (u_i = -(i - (N-1)/2) w - offset; run_mgfbp.py:247, 1247-1248)
    assert spec.u[0] > spec.u[-1]
    assert np.isclose(spec.u[0], np.float32(511.5) * np.float32(0.388) + np.float32(5.69), rtol=1e-6)


def test_modified_config_reads_back_with_the_same_geometry(tmp_path, monkeypatch):
    """The modified config (ImageCenterZ now given) builds the same geometry
    as its source config."""
    monkeypatch.chdir(tmp_path)
    a = Mgfbp(load_jsonc(str(GOLD / "config_mgfbp_temp.jsonc"))).geometry_spec()
    b = Mgfbp(load_jsonc(str(GOLD / "modified_config_mgfbp.jsonc"))).geometry_spec()
    for k in ("angles", "u", "v"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert (a.cz, a.sid, a.sdd, a.img_dim_z) == (b.cz, b.sid, b.sdd, b.img_dim_z)
