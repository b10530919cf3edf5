"""GPU parity of the FBP/FDK path (weighting, ramp filter, back-projection)
against the oracle, through the C ABI.

Tolerances: FP32 kernels vs the FP32 reference-order oracle rel-L2 <= 1e-5
with identical masks (voxels zeroed by the detector-range test); FP16/BF16
storage vs the FP64 oracle as stated per test.
"""
import numpy as np
import pytest
import torch

from oracle import recon
from reconstruction_program_taichi_amd.device import Plan
from tests._cases import PI, cone_case, fan_case, oracle_dict, rel_l2

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("kind", ["fan_full", "fan_short", "fan_short_rot", "cone"])
def test_weight_matches_oracle(ctx, kind):
    if kind == "cone":
        spec = cone_case()
    elif kind == "fan_full":
        spec = fan_case()
    elif kind == "fan_short":
        spec = fan_case(scan_deg=220.0)
    else:
        spec = fan_case(scan_deg=-250.0, rot_deg=30.0, off=1.3)
    rng = np.random.default_rng(3)
    sgm = rng.uniform(0, 2, (spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    plan = Plan(ctx, spec)
    t = dev(sgm)
    plan.weight(t)
    got = host(t)
    ref = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle,
                                spec.short_scan, dtype=np.float32)
    assert rel_l2(got, ref) <= 1e-6


def test_weight_bh(ctx):
    spec = fan_case()
    rng = np.random.default_rng(4)
    sgm = rng.uniform(0, 2, (spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    bh = [1.0, 0.05, -0.01]
    plan = Plan(ctx, spec)
    t = dev(sgm)
    plan.weight(t, np.array(bh, np.float32))
    ref = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle,
                                False, bh=bh, dtype=np.float32)
    assert rel_l2(host(t), ref) <= 1e-6


@pytest.mark.parametrize("n,rows", [(100, 7), (800, 9), (1024, 4), (37, 1)])
def test_filter_rows_f32(ctx, n, rows):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((rows, n)).astype(np.float32)
    taps = recon.ramlak_taps(n, 0.4, np.float32)
    out = torch.empty((rows, n), dtype=torch.float32, device=DEV)
    ctx.filter_rows(dev(x), out, taps, 0.4, "f32")
    got = host(out)
    ref32 = recon.filter_rows(x, taps, 0.4, np.float32)
    ref64 = recon.filter_rows(x, taps.astype(np.float64), 0.4, np.float64)
    assert rel_l2(got, ref32) <= 2e-6
    assert rel_l2(got, ref64) <= 1e-5


@pytest.mark.parametrize("method,tol", [("bf16", 1e-2), ("bf16x3", 1e-4)])
@pytest.mark.parametrize("n,rows", [(1024, 300), (800, 129), (100, 7), (37, 1)])
def test_filter_rows_mfma(ctx, method, tol, n, rows):
    """Toeplitz GEMM on MFMA against the FP64 oracle: BF16 operands (relative
    rounding 2^-9 per operand: the rows' first differences and the taps'
    running sums) and split-BF16 (about 2^-17)."""
    rng = np.random.default_rng(n + rows)
    x = rng.standard_normal((rows, n)).astype(np.float32)
    taps = recon.ramlak_taps(n, 0.4, np.float32)
    out = torch.empty((rows, n), dtype=torch.float32, device=DEV)
    ctx.filter_rows(dev(x), out, taps, 0.4, method)
    got = host(out)
    ref64 = recon.filter_rows(x.astype(np.float64), taps.astype(np.float64), 0.4, np.float64)
    err = rel_l2(got, ref64)
    print(f"{method} n={n}: rel-L2 vs FP64 {err:.2e}")
    assert err <= tol


@pytest.mark.parametrize("case", ["cone_bf16x3", "fan_short_bh_bf16", "fan_curved_block"])
def test_weight_filter_fused_equals_separate(ctx, case):
    """mct_weight_filter_rows = mct_weight_views then mct_filter_rows, bit for
    bit (the same FP32 weighting, applied where the rows become BF16
    planes), with the input sinogram left unweighted."""
    from dataclasses import replace

    bh, view_begin, count = None, 0, None
    if case == "cone_bf16x3":
        spec, method, w = cone_case(), "bf16x3", 0.6
    elif case == "fan_short_bh_bf16":
        spec, method, w = fan_case(scan_deg=220.0, rows=3), "bf16", 0.8
        bh = [1.0, 0.05, -0.002]
    else:  # views 17 .. 56 of a short scan on a curved detector
        spec, method, w = replace(fan_case(scan_deg=250.0, rows=2), curved=True), "bf16x3", 0.8
        view_begin, count = 17, 40
    count = count or spec.n_views
    rng = np.random.default_rng(21)
    raw = rng.uniform(0, 2, (spec.n_v, count, spec.n_u)).astype(np.float32)
    taps = recon.ramlak_taps(spec.n_u, w, np.float32)
    plan = Plan(ctx, spec)
    sep_in = dev(raw)
    plan.weight_views(sep_in, view_begin, bh)
    ref = torch.empty_like(sep_in)
    ctx.filter_rows(sep_in, ref, taps, w, method)
    sgm = dev(raw)
    got = torch.empty_like(sgm)
    plan.weight_filter(sgm, got, taps, w, method, view_begin, bh)
    torch.cuda.synchronize()
    assert torch.equal(got, ref)
    assert torch.equal(sgm.cpu(), torch.from_numpy(raw))
    assert np.count_nonzero(host(got)) > got.numel() // 2


@pytest.mark.parametrize("case", ["cone_bf16x3", "cone_wide_bh_bf16", "cone_block_curved"])
def test_weight_filter_pack_equals_filter_then_pack(ctx, case):
    """mct_weight_filter_pack = mct_weight_filter_rows then mct_pack_sinogram
    (FP32, ROWS), bit for bit: the filter's epilogue writes each element with
    its right neighbour's difference, the columns at the end of a
    wavefront's 64 are completed afterwards, and the zero rows either side are
    written.  Row lengths with one (80) and several (200, 1024) such columns,
    a partial last row tile, BH coefficients and a block of views."""
    from dataclasses import replace

    bh, view_begin, count = None, 0, None
    if case == "cone_bf16x3":
        spec, method, w = cone_case(nu=80, nrows=29), "bf16x3", 0.6
    elif case == "cone_wide_bh_bf16":
        spec, method, w = cone_case(nu=1024, nrows=7, nv=40), "bf16", 0.4
        bh = [1.0, 0.05, -0.002]
    else:
        spec, method, w = replace(cone_case(nu=200, nrows=13, nv=90), curved=True), "bf16x3", 0.5
        view_begin, count = 23, 41
    count = count or spec.n_views
    rng = np.random.default_rng(5)
    raw = rng.uniform(0, 2, (spec.n_v, count, spec.n_u)).astype(np.float32)
    taps = recon.ramlak_taps(spec.n_u, w, np.float32)
    plan = Plan(ctx, spec)
    sgm = dev(raw)
    filt = torch.empty_like(sgm)
    plan.weight_filter(sgm, filt, taps, w, method, view_begin, bh)
    want = ctx.pack(filt, "f32", "rows")
    got = torch.full_like(want, 0x5A)  # poisoned: every byte must be written
    plan.weight_filter_pack(sgm, got, taps, w, method, view_begin, bh)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
    assert torch.equal(sgm.cpu(), torch.from_numpy(raw))


def test_weight_filter_pack_refuses_row_lengths_it_cannot_take(ctx):
    spec = cone_case(nu=65)  # n_u - 1 = 64: the last element's neighbour has no column
    plan = Plan(ctx, spec)
    sgm = torch.zeros((spec.n_v, spec.n_views, spec.n_u), device=DEV)
    packed = ctx.pack(sgm, "f32", "rows")
    with pytest.raises(RuntimeError, match="multiple of 64"):
        plan.weight_filter_pack(sgm, packed, recon.ramlak_taps(spec.n_u, 0.6, np.float32), 0.6, "bf16x3")


def test_weight_filter_fused_rejects_the_f32_method(ctx):
    spec = fan_case()
    plan = Plan(ctx, spec)
    sgm = torch.zeros((spec.n_v, spec.n_views, spec.n_u), device=DEV)
    with pytest.raises(RuntimeError, match="MFMA"):
        plan.weight_filter(sgm, torch.empty_like(sgm), recon.ramlak_taps(spec.n_u, 0.8, np.float32),
                           0.8, "f32")


@pytest.mark.parametrize("rows", [1, 7, 37, 48])
def test_filter_vertical_matches_oracle(ctx, rows):
    from reconstruction_program_taichi_amd import geometry as geo

    rng = np.random.default_rng(rows)
    x = rng.standard_normal((rows, 9, 33)).astype(np.float32)
    kern = geo.gaussian_taps(rows, 0.6, 1.2)
    out = torch.empty_like(dev(x))
    ctx.filter_vertical(dev(x), out, kern, 0.6)
    ref32 = recon.filter_vertical(x, kern, 0.6, np.float32)
    ref64 = recon.filter_vertical(x.astype(np.float64), kern.astype(np.float64), 0.6, np.float64)
    got = host(out)
    assert rel_l2(got, ref32) <= 1e-6
    assert rel_l2(got, ref64) <= 1e-5


def test_context_destroyed_before_plan():
    """Garbage collection may finalise a context before its plans: destroying
    a plan must not read its (already freed) context."""
    from reconstruction_program_taichi_amd.device import Context

    c = Context(0)
    p = Plan(c, fan_case(nx=16, nv=8, nu=20))
    c.close()
    p.close()


def test_filter_rows_empty(ctx):
    out = torch.empty((0, 16), dtype=torch.float32, device=DEV)
    inp = torch.empty((0, 16), dtype=torch.float32, device=DEV)
    ctx.filter_rows(inp, out, np.zeros(31, np.float32), 1.0, "f32")


def mask_disagreements(got, ref, alive):
    """Voxels the GPU kept although the oracle's range test zeroes them, plus
    voxels the GPU zeroed although the oracle keeps a nonzero value."""
    return int(((got != 0) & ~alive).sum() + ((got == 0) & alive & (ref != 0)).sum())


def _bp_check(ctx, spec, layout, dtype="f32", seed=0):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    plan = Plan(ctx, spec)
    packed = ctx.pack(dev(f), dtype, layout)
    vol = torch.empty((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32, device=DEV)
    plan.backproject_fbp(packed, dtype, layout, vol)
    got = host(vol)
    ref, alive = recon.backproject_fbp(f, oracle_dict(spec), np.float32, return_mask=True)
    return got, ref, alive, f


@pytest.mark.parametrize("kind", ["full", "short", "rot_off"])
def test_bp_fan_pair(ctx, kind):
    spec = {"full": fan_case(), "short": fan_case(scan_deg=200.0),
            "rot_off": fan_case(rot_deg=17.0, off=2.0, scan_deg=300.0)}[kind]
    got, ref, alive, _ = _bp_check(ctx, spec, "pair")
    assert mask_disagreements(got, ref, alive) == 0
    assert rel_l2(got, ref) <= 1e-5


def test_bp_fan_plain_generic(ctx):
    spec = fan_case(nx=40, nv=50, nu=70)
    got, ref, alive, _ = _bp_check(ctx, spec, "plain")
    assert rel_l2(got, ref) <= 1e-5


@pytest.mark.parametrize("layout", ["quad", "rows"])
@pytest.mark.parametrize("zc,batch", [(4, 4), (8, 4), (8, 8), (16, 8), (16, 16), (32, 4), (32, 8),
                                      (32, 16)])
def test_bp_cone_quad_tiles(ctx, zc, batch, layout):
    """Direct-gather kernel at every tile shape, from QUAD cells and from ROWS
    element rows."""
    spec = cone_case(nz=37)  # taller than the detector's cone: top/bottom slices masked
    ctx.set_bp_tile(zc, batch)
    ctx.set_bp_kernel("direct")
    try:
        got, ref, alive, _ = _bp_check(ctx, spec, layout)
    finally:
        ctx.set_bp_tile(0, 0)
        ctx.set_bp_kernel("auto")
    assert 0.02 < (~alive).mean() < 0.98  # the detector-range mask is exercised
    assert mask_disagreements(got, ref, alive) == 0
    assert rel_l2(got, ref) <= 1e-5


def _lds_spec(kind):
    if kind == "tall":
        return cone_case(nz=37)
    if kind == "top_row_offset":
        return cone_case(top_row=True, off_v=1.3, nz=30)
    if kind == "odd_sizes":  # partial tiles in x/y and a partial slice chunk
        return cone_case(nx=41, nz=21, nv=50, nu=83, nrows=29)
    if kind == "rotated_offcentre":
        spec = cone_case(nz=20)
        from dataclasses import replace

        from reconstruction_program_taichi_amd import geometry as geo
        rot = 0.4
        return replace(spec, img_rot=rot, cx=1.7, cy=-2.2,
                       angles=geo.angle_array(spec.n_views, 2 * PI, rot))
    if kind == "curved":
        from dataclasses import replace
        return replace(cone_case(nz=30), curved=True)
    raise ValueError(kind)


@pytest.mark.parametrize("layout", ["quad", "rows"])
@pytest.mark.parametrize("kind", ["tall", "top_row_offset", "odd_sizes", "rotated_offcentre", "curved"])
def test_bp_cone_lds_staged(ctx, kind, layout):
    """LDS-staged kernel against the oracle (identical masks) and against the
    direct-gather kernel.  The two differ only in where the row coordinate is
    rounded (relative to the window's first row, or to detector row 0), about
    1e-6 rel-L2 on these white-noise sinograms."""
    spec = _lds_spec(kind)
    ctx.set_bp_kernel("lds")
    try:
        got, ref, alive, f = _bp_check(ctx, spec, layout, seed=21)
        assert ctx.bp_last_kernel() == "lds"
        ctx.set_bp_kernel("direct")
        direct, _, _, _ = _bp_check(ctx, spec, layout, seed=21)
        assert ctx.bp_last_kernel() == "direct"
    finally:
        ctx.set_bp_kernel("auto")
    print(f"{kind} ({layout}): lds vs oracle {rel_l2(got, ref):.2e}, direct vs oracle {rel_l2(direct, ref):.2e}, "
          f"lds vs direct {rel_l2(got, direct):.2e}")
    assert mask_disagreements(got, ref, alive) == 0
    assert rel_l2(got, ref) <= 1e-5
    assert np.array_equal(got != 0, direct != 0)
    assert rel_l2(got, direct) <= 4e-6


@pytest.mark.parametrize("layout", ["quad", "rows"])
@pytest.mark.parametrize("dtype", ["f16", "bf16"])
def test_bp_cone_lds_half_storage_equals_direct(ctx, dtype, layout):
    spec = cone_case(nz=37)
    rng = np.random.default_rng(13)
    f = (1.0 + 0.3 * rng.standard_normal((spec.n_v, spec.n_views, spec.n_u))).astype(np.float32)
    plan = Plan(ctx, spec)
    packed = ctx.pack(dev(f), dtype, layout)
    out = {}
    for kind in ("lds", "direct"):
        ctx.set_bp_kernel(kind)
        try:
            vol = torch.empty((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32,
                              device=DEV)
            plan.backproject_fbp(packed, dtype, layout, vol)
            out[kind] = host(vol)
        finally:
            ctx.set_bp_kernel("auto")
    assert np.array_equal(out["lds"] != 0, out["direct"] != 0)
    assert rel_l2(out["lds"], out["direct"]) <= 1e-6


@pytest.mark.parametrize("layout", ["quad", "rows"])
@pytest.mark.parametrize("variant", ["top_row", "offset"])
def test_bp_cone_variants(ctx, variant, layout):
    spec = cone_case(top_row=True) if variant == "top_row" else cone_case(off_v=1.7, nz=30)
    got, ref, alive, _ = _bp_check(ctx, spec, layout, seed=5)
    assert mask_disagreements(got, ref, alive) == 0
    assert rel_l2(got, ref) <= 1e-5


def test_bp_cone_plain_generic(ctx):
    spec = cone_case(nx=32, nz=16, nv=40)
    got, ref, alive, _ = _bp_check(ctx, spec, "plain")
    assert mask_disagreements(got, ref, alive) == 0
    assert rel_l2(got, ref) <= 1e-5


@pytest.mark.parametrize("layout", ["quad", "rows"])
@pytest.mark.parametrize("dtype,tol", [("f16", 1e-3), ("bf16", 8e-3)])
def test_bp_cone_half_storage(ctx, dtype, tol, layout):
    spec = cone_case()
    rng = np.random.default_rng(9)
    # smooth positive data, as a filtered sinogram of a phantom is at this scale
    f = (1.0 + 0.3 * rng.standard_normal((spec.n_v, spec.n_views, spec.n_u))).astype(np.float32)
    plan = Plan(ctx, spec)
    packed = ctx.pack(dev(f), dtype, layout)
    vol = torch.empty((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32, device=DEV)
    plan.backproject_fbp(packed, dtype, layout, vol)
    got = host(vol)
    ref, alive = recon.backproject_fbp(f, oracle_dict(spec), np.float64, return_mask=True)
    keep = alive & (got != 0)
    err = rel_l2(got[keep], ref[keep])
    print(f"{dtype} storage ({layout}): rel-L2 vs FP64 oracle {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("layout", ["quad", "rows"])
def test_fdk_pipeline_small(ctx, layout):
    """weight -> ramp filter -> pack -> back-project against the oracle chain."""
    spec = cone_case(nu=80, nrows=32)
    rng = np.random.default_rng(11)
    sgm = rng.uniform(0, 1, (spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    taps = recon.ramlak_taps(spec.n_u, 0.6, np.float32)
    plan = Plan(ctx, spec)
    t = dev(sgm)
    plan.weight(t)
    filt = torch.empty_like(t)
    ctx.filter_rows(t, filt, taps, 0.6, "f32")
    packed = ctx.pack(filt, "f32", layout)
    vol = torch.empty((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32, device=DEV)
    plan.backproject_fbp(packed, "f32", layout, vol)
    got = host(vol)
    w = recon.weight_sinogram(sgm, spec.u, spec.v, spec.angles, spec.sdd, spec.total_scan_angle,
                              False, dtype=np.float32)
    fr = recon.filter_rows(w, taps, 0.6, np.float32)
    ref, alive = recon.backproject_fbp(fr, oracle_dict(spec), np.float32, return_mask=True)
    assert mask_disagreements(got, ref, alive) == 0
    assert rel_l2(got, ref) <= 1e-5
