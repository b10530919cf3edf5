"""CPU tests of bench.py's host logic: the counter record it reports the
back-projection's HBM traffic and VALU issue from is found only when its
source hash matches the sources at hand, and the derived figures follow
from the counters."""
import sys
import types
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402
from reconstruction_program_taichi_amd import _build  # noqa: E402

ARGS = types.SimpleNamespace(storage="f32", bp_tile=0, bp_batch=0)


def test_committed_counter_record_matches_the_sources():
    rec, where = bench.counter_record("bp_cone_lds_kernel", bench.counter_workload(ARGS), "mct_bp.hip")
    assert rec is not None, where
    assert where.startswith("profiles/")
    assert rec["source_hash"] == _build.kernel_source_hash("mct_bp.hip")


def test_counter_record_of_other_sources_is_refused(monkeypatch):
    monkeypatch.setattr(_build, "kernel_source_hash", lambda source: "0" * 16)
    rec, why = bench.counter_record("bp_cone_lds_kernel", bench.counter_workload(ARGS), "mct_bp.hip")
    assert rec is None and "other sources" in why
    rec, why = bench.counter_record("bp_cone_lds_kernel", "another workload", "mct_bp.hip")
    assert rec is None and "no counter record" in why


def test_source_hash_follows_the_kernel_file(tmp_path, monkeypatch):
    csrc = tmp_path / "csrc"
    csrc.mkdir()
    (tmp_path / "include").mkdir()
    (tmp_path / "include" / "mct.h").write_text("h")
    (csrc / "a.hip").write_text("kernel 1")
    (csrc / "b.hip").write_text("other")
    (csrc / "x.h").write_text("header")
    monkeypatch.setattr(_build, "CSRC", csrc)
    monkeypatch.setattr(_build, "PKG", tmp_path / "pkg")
    h0 = _build.kernel_source_hash("a.hip")
    (csrc / "b.hip").write_text("other, edited")
    assert _build.kernel_source_hash("a.hip") == h0  # another file's kernels
    (csrc / "x.h").write_text("header, edited")
    h1 = _build.kernel_source_hash("a.hip")
    assert h1 != h0
    (csrc / "a.hip").write_text("kernel 2")
    assert _build.kernel_source_hash("a.hip") not in (h0, h1)


def test_measured_bounds_arithmetic():
    rec = {"source_hash": "x",
           "per_dispatch": {"FETCH_SIZE_KiB": 1.0e7, "WRITE_SIZE_KiB": 1.0e5, "SQ_INSTS_VALU": 2.4e9,
                            "GRBM_GUI_ACTIVE": 1.2e7},
           "calibration": {"fetch_factor": 2.0, "write_factor": 1.0}}
    rec["per_dispatch"]["SQ_ACTIVE_INST_VALU2"] = 3.0e8
    m = bench.measured_bounds(rec, "profiles/x.md", 10.0, 100, {"bp_slice_mix": 4.0})
    assert m["hbm_bytes"] == pytest.approx(1.0e7 * 1024 * 2 + 1.0e5 * 1024)
    assert m["hbm_frac"] == pytest.approx(m["hbm_bytes"] / 0.01 / 8e12, rel=1e-3)
    # 2.4e9 wave instructions over 100 CUs and 10 ms: 2.4 per CU per ns, one
    # per CU per clock at 2.4 GHz, half the two-per-clock dual-issue ceiling
    assert m["valu_insts_per_cu_per_ns"] == pytest.approx(2.4)
    assert m["valu_single_issue_frac"] == pytest.approx(1.0)
    assert m["valu_issue_frac"] == pytest.approx(0.5)
    assert m["valu_issue_frac_busy_cycles"] == pytest.approx(2.4e9 / (2 * 100 * 1.2e7))
    # 3e8 dual-issue SIMD cycles over 100 CUs x 1.2e7 clocks (4 SIMDs x 1/4)
    assert m["valu_dual_issue_cycles_frac"] == pytest.approx(3.0e8 / (100 * 1.2e7))
    assert m["valu_frac_of_slice_mix_rate"] == pytest.approx(2.4 / 4.0)
    assert "valu_frac_of_slice_mix_rate" not in bench.measured_bounds(rec, "profiles/x.md", 10.0, 100, None)
