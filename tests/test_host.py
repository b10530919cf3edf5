"""CPU tests of the host side: the C ABI library loads and exports every
symbol include/mct.h declares (no GPU call), the JSONC loader and config
parsing behave as the reference's, and the geometry arrays follow their
closed forms."""
import json
import re
from pathlib import Path

import numpy as np
import pytest

from reconstruction_program_taichi_amd import _abi, config, geometry as geo

ROOT = Path(__file__).resolve().parent.parent


def header_functions():
    text = (ROOT / "include" / "mct.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mct_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
        assert n in _abi.SIGNATURES, n
    assert set(_abi.SIGNATURES) == set(names)
    assert lib.mct_abi_version() == _abi.ABI_VERSION


def test_packed_bytes_host_helper():
    lib = _abi.load()
    # QUAD: cell rows -1 .. rows - 1 (the detector's 255 and one zero-padded row either side)
    assert lib.mct_packed_bytes(256, 1080, 1024, _abi.MCT_F32, _abi.MCT_QUAD) == 257 * 1080 * 1023 * 16
    # ROWS: element rows -1 .. rows (one zero-padded row either side), PAIR taps
    assert lib.mct_packed_bytes(256, 1080, 1024, _abi.MCT_F32, _abi.MCT_ROWS) == 258 * 1080 * 1023 * 8
    assert lib.mct_packed_bytes(256, 1080, 1024, _abi.MCT_BF16, _abi.MCT_ROWS) == 258 * 1080 * 1023 * 4
    assert lib.mct_packed_bytes(1, 10, 8, _abi.MCT_F32, _abi.MCT_ROWS) == -1
    assert lib.mct_packed_bytes(1, 720, 800, _abi.MCT_F16, _abi.MCT_PAIR) == 720 * 799 * 4
    assert lib.mct_packed_bytes(3, 10, 8, _abi.MCT_BF16, _abi.MCT_PLAIN) == 3 * 10 * 8 * 2
    assert lib.mct_packed_bytes(1, 10, 8, _abi.MCT_F32, _abi.MCT_QUAD) == -1
    assert lib.mct_packed_bytes(0, 10, 8, _abi.MCT_F32, _abi.MCT_PLAIN) == -1


def test_null_arguments_return_errors_not_crashes():
    lib = _abi.load()
    assert lib.mct_create(0, None) == _abi.MCT_ERR_INVALID
    assert b"NULL" in lib.mct_last_error()
    assert lib.mct_plan_create(None, None, None) == _abi.MCT_ERR_INVALID
    assert lib.mct_set_bp_tile(None, 8, 0) == _abi.MCT_ERR_INVALID
    assert lib.mct_destroy(None) == 0


# ------------------------------------------------------------------ JSONC
def test_remove_comments_matches_reference_regex():
    s = '{\n// line\n"a": 1, /* block\n spans */ "b": "x//y"\n}'
    out = config.remove_comments(s)
    assert "line" not in out and "spans" not in out
    # the reference's regex also strips // inside strings
    assert '"x' in out and "y\"" not in out


def test_load_jsonc(tmp_path):
    p = tmp_path / "c.jsonc"
    p.write_text('// header\n{\n  "Views": 720, // views\n  /* c */ "ConeBeam": false\n}\n')
    assert config.ReadConfigFile(str(p)) == {"Views": 720, "ConeBeam": False}
    out = tmp_path / "o.jsonc"
    config.save_jsonc(str(out), {"a": [1, 2]})
    assert json.loads(out.read_text()) == {"a": [1, 2]}
    with pytest.raises(AssertionError):
        config.save_jsonc(str(tmp_path / "o.json"), {})


def test_imread_raw_offset_gap(tmp_path):
    a = np.arange(5 * 4 * 3, dtype=np.float32).reshape(5, 4, 3)
    p = tmp_path / "a.raw"
    a.tofile(p)
    r = config.imreadRaw(str(p), height=2, width=3, nSlice=5, offset=3 * 4, gap=2 * 3 * 4)
    assert np.array_equal(r, a[:, 1:3, :])


def _fbp_cfg(tmp_path, **kw):
    c = {"InputDir": str(tmp_path), "OutputDir": str(tmp_path / "out"), "InputFiles": "sgm_.*",
         "OutputFilePrefix": "", "OutputFileReplace": ["sgm", "rec"],
         "DetectorElementCountHorizontal": 100, "DetectorElementWidth": 0.8,
         "DetectorOffsetHorizontal": 0.5, "DetectorElementCountVertical": 2, "Views": 90,
         "SourceIsocenterDistance": 250, "SourceDetectorDistance": 500, "ImageDimension": 64,
         "PixelSize": 0.6, "HammingFilter": 1}
    c.update(kw)
    return c


def test_mgfbp_config_defaults_and_arrays(tmp_path):
    from reconstruction_program_taichi_amd.fbp import Mgfbp

    m = Mgfbp(_fbp_cfg(tmp_path, ImageRotation=10.0))
    assert m.short_scan == 0 and m.img_dim_z == 2 and m.output_file_format == "raw"
    spec = m.geometry_spec()
    i = np.arange(100)
    assert np.allclose(spec.u, -(i - 49.5) * 0.8 - 0.5, atol=1e-5)
    assert np.allclose(spec.angles, 2 * geo.PI / 90 * np.arange(90) + 10 / 180 * geo.PI, atol=1e-6)
    assert np.all(spec.v == 0)
    assert spec.u.dtype == np.float32 and spec.angles.dtype == np.float32


def test_mgfbp_cone_center_z_default(tmp_path):
    from reconstruction_program_taichi_amd.fbp import Mgfbp

    m = Mgfbp(_fbp_cfg(tmp_path, ConeBeam=True, DetectorElementHeight=0.4, ImageDimensionZ=8,
                       VoxelHeight=0.3, DetectorElementCountVertical=20, DetectorOffsetVertical=1.0,
                       DetectorElementVerticalReconRange=[4, 11]))
    # run_mgfbp.py:651-656: centre row of the range, magnified back to the isocentre
    expect = (1.0 + ((4 + 11) / 2 - 19 / 2) * 0.4) * 250 / 500
    assert abs(m.img_center_z - expect) < 1e-12
    spec = m.geometry_spec()
    assert spec.n_v == 8
    assert np.allclose(spec.v, (np.arange(8) + 4 - 9.5) * 0.4 + 1.0, atol=1e-6)


@pytest.mark.parametrize("kw,msg", [
    ({"ImageDimension": -3}, "ImageDimension must be a positive integer"),
    ({"OutputFileFormat": "png"}, "Output file format"),
    ({"ReconViewMode": "Coronal"}, "ReconViewMode can only be"),
    ({"DetectorElementVerticalReconRange": [500, 700]}, "Out of detector row range"),
    ({"ConeBeam": "yes"}, "ConeBeam must be true or false"),
])
def test_mgfbp_config_errors_exit(tmp_path, capsys, kw, msg):
    from reconstruction_program_taichi_amd.fbp import Mgfbp

    with pytest.raises(SystemExit):
        Mgfbp(_fbp_cfg(tmp_path, **kw))
    assert msg in capsys.readouterr().out


def test_missing_horizontal_offset_raises_attribute_error(tmp_path):
    from reconstruction_program_taichi_amd.fbp import Mgfbp

    c = _fbp_cfg(tmp_path)
    del c["DetectorOffsetHorizontal"]
    with pytest.raises(AttributeError):
        Mgfbp(c)


def test_missing_filter_key_fails_at_filtering(tmp_path):
    from reconstruction_program_taichi_amd.fbp import Mgfbp

    c = _fbp_cfg(tmp_path)
    del c["HammingFilter"]
    m = Mgfbp(c)
    with pytest.raises(AttributeError):
        m.recon_kernel()


def test_div_factor_cases():
    assert geo.div_factor(2 * geo.PI, False) == pytest.approx(0.5)
    assert geo.div_factor(4 * geo.PI, False) == pytest.approx(0.25)
    assert geo.div_factor(geo.PI + 0.3, True) == 1.0
    assert geo.div_factor(2 * geo.PI + 1.0, True) == pytest.approx(0.5)
    assert geo.div_factor(2 * geo.PI + 4.0, True) == pytest.approx(1 / 3)


def test_plan_methods_refuse_buffers_of_another_shape():
    """Every Plan method checks its tensors against the plan's geometry before
    any device call (the kernels size their launches from the geometry)."""
    import torch
    from reconstruction_program_taichi_amd.device import Plan
    from tests._cases import cone_case

    spec = cone_case(nx=16, nz=8, nv=12, nu=20, nrows=10)
    plan = Plan.__new__(Plan)
    plan.spec, plan.h, plan.ctx = spec, None, None
    sgm_ok = torch.zeros((10, 12, 20))
    vol_ok = torch.zeros((8, 16, 16))
    bad_sgm = [torch.zeros((10, 12, 19)), torch.zeros((12, 10, 20)), torch.zeros((10, 11, 20))]
    bad_vol = [torch.zeros((8, 16, 15)), torch.zeros((16, 16, 8))]
    for b in bad_sgm:
        with pytest.raises(ValueError, match="shape"):
            plan.weight(b)
        with pytest.raises(ValueError, match="shape"):
            plan.forward_project(vol_ok, b)
        with pytest.raises(ValueError, match="shape"):
            plan.backproject_ir(b, None, vol_ok)
        with pytest.raises(ValueError, match="shape"):
            plan.ray_mask(b)
    for b in bad_vol:
        with pytest.raises(ValueError, match="shape"):
            plan.forward_project(b, sgm_ok)
        with pytest.raises(ValueError, match="shape"):
            plan.backproject_ir(sgm_ok, None, b)
    with pytest.raises(ValueError, match="shape"):
        plan.forward_project(vol_ok, sgm_ok, ray_mask=torch.zeros((10, 12, 10)))
    with pytest.raises(ValueError, match="multiple"):
        plan.forward_project(vol_ok, sgm_ok, oversample=3)
    with pytest.raises(ValueError, match="shape"):
        plan.forward_project(vol_ok, sgm_ok, oversample=2)  # out must have n_u / 2 columns
    small = torch.zeros(100, dtype=torch.uint8)
    with pytest.raises(ValueError, match="bytes"):
        plan.backproject_fbp(small, "f32", "quad", vol_ok)
    with pytest.raises(ValueError, match="views"):
        plan.weight_views(torch.zeros((10, 5, 20)), 9)
    with pytest.raises(ValueError, match="views"):
        plan.backproject_views(small, "f32", "quad", 10, 4, vol_ok)
    with pytest.raises(ValueError, match="shape"):
        plan.weight_filter(torch.zeros((10, 4, 20)), torch.zeros((10, 4, 21)), np.zeros(39), 1.0, "bf16x3")
    # right shapes on the host: refused as host memory, still before any call
    with pytest.raises(ValueError, match="device"):
        plan.weight(sgm_ok)
    with pytest.raises(ValueError, match="float32"):
        plan.ray_mask(torch.zeros((10, 12, 20), dtype=torch.float64))


@pytest.mark.parametrize("n", [1, 2, 37, 800])
def test_ramp_filter_difference_form_is_exact(n):
    """The MFMA ramp filter's operands: D_m = x_m - x_{m-1} for m = 0..n
    (x_{-1} = x_n = 0) and G[i] = taps[0] + ... + taps[i-1] (2n entries,
    G[0] = 0); sum_{m<=n} D_m G[k + n - m] equals the reference's
    sum_{m<n} x_m taps[k + n-1 - m] (summation by parts), in FP64."""
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 50.0
    taps = rng.standard_normal(2 * n - 1)
    D = np.diff(np.concatenate([[0.0], x, [0.0]]))  # D_0 .. D_n
    G = np.concatenate([[0.0], np.cumsum(taps)])
    k = np.arange(n)
    y = np.array([sum(D[m] * G[kk + n - m] for m in range(n + 1)) for kk in k])
    want = np.convolve(x, taps)[n - 1:2 * n - 1]
    assert np.allclose(y, want, rtol=1e-10, atol=1e-9 * np.abs(want).max())


def test_kernel_fallbacks_are_reported_once_per_plan():
    """When the automatic kernel choice has to fall back to a direct kernel
    (tables or records too large for the geometry), the plan says so once,
    as a RuntimeWarning; not when the caller asked for the direct kernel."""
    import types
    import warnings

    from reconstruction_program_taichi_amd.device import Plan
    from tests._cases import cone_case, fan_case

    def plan_for(spec, bp_kernel="auto", fp_kernel="auto", last="direct"):
        p = Plan.__new__(Plan)
        p.spec, p.h = spec, None
        p.ctx = types.SimpleNamespace(bp_kernel=bp_kernel, fp_kernel=fp_kernel,
                                      bp_last_kernel=lambda: last, fp_last_kernel=lambda: last)
        return p

    cone = cone_case(nx=16, nz=8, nv=12, nu=20, nrows=10)
    p = plan_for(cone)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        p._report_fallback_bp("quad")
        p._report_fallback_bp("quad")
        p._report_fallback_fp()
        p._report_fallback_fp()
    msgs = [str(x.message) for x in w if issubclass(x.category, RuntimeWarning)]
    assert len(msgs) == 2 and "direct-gather" in msgs[0] and "records" in msgs[1]
    quiet = [plan_for(cone, bp_kernel="direct", fp_kernel="direct"), plan_for(cone, last="lds"),
             plan_for(fan_case(nx=16, nv=8, nu=20))]
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        for q in quiet[:2]:
            q._report_fallback_bp("quad")
        quiet[0]._report_fallback_fp()
        quiet[2]._report_fallback_bp("pair")
    assert not [x for x in w if issubclass(x.category, RuntimeWarning)]


@pytest.mark.parametrize("n, nz, nv, nu", [(512, 256, 256, 1024), (1024, 1024, 4096, 4096)])
def test_lds_window_margin_covers_fp32_rounding(n, nz, nv, nu):
    """The LDS-staged cone back-projector's windows (csrc/mct_bp.hip,
    bp_window_kernel) hold the cells of [min - e, max + e] of each tile's
    detector coordinates (FP64, e = 1/64 + 2^-18 |coordinate|), and the slice
    loop reads its row without a clamp: every voxel's FP32 row coordinate, as
    the kernel computes it (products and sums rounded one by one, 1/L from a
    reciprocal within 1 ulp), must stay within e of the exact one.  Checked
    on random voxels and views of the benchmark geometry and of a 1024^3
    volume on a 4096 x 4096 detector."""
    rng = np.random.default_rng(7)
    f32 = np.float32
    pix, vh, du, dv, sid, sdd = 0.3, 0.3, 0.4, 0.4, 750.0, 1250.0
    pix = pix * 512 / n
    u0, v0 = -(nu - 1) / 2 * du, -(nv - 1) / 2 * dv
    m = 200_000
    ix, iy = rng.integers(0, n, m), rng.integers(0, n, m)
    iz, beta = rng.integers(0, nz, m), rng.uniform(0, 2 * np.pi, m)
    x = pix * (ix - (n - 1) / 2)
    y = -pix * (iy - (n - 1) / 2)
    z = (iz - (nz - 1) / 2) * vh
    c, s = np.cos(beta), np.sin(beta)
    # exact (FP64)
    L = sid - x * c - y * s
    t64 = (sdd / L * z - v0) / dv
    u64 = (sdd / L * (-x * s + y * c) - u0) / du
    # the kernel's FP32 order; the reciprocal off by up to one ulp either way
    xf, yf, cf, sf = f32(x), f32(y), f32(c), f32(s)
    Lf = f32(f32(f32(sid) - f32(xf * cf)) - f32(yf * sf))
    rL = f32(f32(1.0) / Lf)
    rL = np.where(rng.integers(0, 2, m) == 1, np.nextafter(rL, f32(np.inf)), np.nextafter(rL, f32(0)))
    mag = f32(f32(sdd) * rL)
    zf = f32(f32(f32(iz) - f32((nz - 1) / 2)) * f32(vh))
    t32 = f32(f32(f32(mag * zf) - f32(v0)) * f32(1.0 / dv))
    yr = f32(f32(f32(-xf) * sf) + f32(yf * cf))
    u32 = f32(f32(f32(mag * yr) - f32(u0)) * f32(1.0 / du))
    et = 1.0 / 64 + np.abs(t64) * 2.0 ** -18
    eu = 1.0 / 64 + np.abs(u64) * 2.0 ** -18
    assert np.max(np.abs(t32 - t64) / et) < 0.1
    assert np.max(np.abs(u32 - u64) / eu) < 0.1
