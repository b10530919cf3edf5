"""GPU parity of the forward projector, ray mask, noise and the IRN-CGLS / TV
kernels against the oracle, through the C ABI.

Tolerances: forward projection vs the FP32 reference-order oracle rel-L2 <=
1e-5; IR back-projection rel-L2 <= 1e-5 with identical truncation masks;
ray mask bit-exact except at most 0.1% of rays whose decisions sit on an
FP32 boundary; FP64 reductions rel 1e-12; noise rel-L2 <= 1e-6 against the
oracle's Philox stream.
"""
import numpy as np
import pytest
import torch

from oracle import recon
from reconstruction_program_taichi_amd import geometry as geo
from reconstruction_program_taichi_amd.device import Plan
from tests._cases import cone_case, fan_case, oracle_dict, rel_l2

pytestmark = pytest.mark.gpu
DEV = "cuda"


def dev(a, dtype=np.float32):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _vol(spec, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(0, 0.02, (spec.img_dim_z, spec.img_dim, spec.img_dim)).astype(np.float32)
    return v


def _fp_spec(kind):
    """fan_rows70 / fan_curved: 70 slices, so the fan path that traces one ray
    per wavefront over 64 slices runs (a full and a partial slice tile);
    cone_helical: gantry travel per view from HelicalPitch."""
    os_ = 2 if kind == "fan_os2" else 1
    if kind in ("fan_rows70", "fan_curved"):
        spec = fan_case(nx=40, nv=24, nu=60, rows=70, rot_deg=11.0, off=0.7)
        spec.curved = kind == "fan_curved"
    elif kind.startswith("fan"):
        spec = fan_case(nx=40, nv=24, nu=60 * os_, w=0.8 / os_, rows=3)
    elif kind == "cone_offcentre":
        spec = cone_case(nx=33, nz=15, nv=20, nu=48, nrows=24)
        spec.cx, spec.cy = 1.3, -0.8
    else:
        spec = cone_case(nx=32, nz=16, nv=20, nu=48, nrows=24, top_row=(kind == "cone_top"))
        if kind == "cone_curved":
            spec.curved = True
        if kind == "cone_helical":
            spec.helical, spec.helical_pitch = True, 0.8
    spec.fpj_step = 0.2
    return spec, os_


def _fp_oracle_dict(spec):
    g = oracle_dict(spec)
    if spec.helical:
        g["zpv"] = recon.helical_z_per_view(spec.angles, spec.v, spec.n_v_full, spec.helical_pitch,
                                            spec.sdd, spec.sid)
    return g


@pytest.mark.parametrize("kernel", ["cells", "direct"])
@pytest.mark.parametrize("kind", ["fan", "cone", "cone_top", "fan_os2", "cone_offcentre", "fan_rows70",
                                  "fan_curved", "cone_curved", "cone_helical"])
def test_forward_project(ctx, kind, kernel):
    spec, os_ = _fp_spec(kind)
    vol = _vol(spec, 1)
    plan = Plan(ctx, spec)
    out = torch.zeros((spec.n_v, spec.n_views, spec.n_u // os_), dtype=torch.float32, device=DEV)
    ctx.set_fp_kernel(kernel)
    try:
        plan.forward_project(dev(vol), out, None, os_)
        assert ctx.fp_last_kernel() == kernel
    finally:
        ctx.set_fp_kernel("auto")
    got = host(out)
    ref = recon.forward_project(vol, _fp_oracle_dict(spec), oversample=os_, dtype=np.float32)
    assert rel_l2(got, ref) <= 1e-5


@pytest.mark.parametrize("kind", ["fan", "cone", "cone_offcentre", "fan_os2", "fan_rows70"])
def test_forward_project_cells_equal_direct(ctx, kind):
    """Same samples and in-bounds tests; the cells hold FP32 bilinear
    coefficients {a, b-a, c-a, (d-c)-(b-a)} where the direct kernel
    interpolates raw voxels, and the 64-slice fan path (fan_rows70) adds each
    cell's steps as one term: equal to rounding, and zero on the same rays."""
    spec, os_ = _fp_spec(kind)
    vol = dev(_vol(spec, 4))
    plan = Plan(ctx, spec)
    outs = {}
    for kernel in ("cells", "direct"):
        ctx.set_fp_kernel(kernel)
        try:
            o = torch.zeros((spec.n_v, spec.n_views, spec.n_u // os_), dtype=torch.float32, device=DEV)
            plan.forward_project(vol, o, None, os_)
            outs[kernel] = host(o)
        finally:
            ctx.set_fp_kernel("auto")
    assert np.array_equal(outs["cells"] != 0, outs["direct"] != 0)
    assert rel_l2(outs["cells"], outs["direct"]) <= 2e-6


def test_forward_project_count_mode_v3(ctx):
    spec = cone_case(nx=24, nz=12, nv=10, nu=40, nrows=16)
    spec.fpj_count_mode = 1
    vol = _vol(spec, 2)
    plan = Plan(ctx, spec)
    out = torch.zeros((spec.n_v, spec.n_views, spec.n_u), dtype=torch.float32, device=DEV)
    plan.forward_project(dev(vol), out)
    ref = recon.forward_project(vol, oracle_dict(spec), count_mode=1, dtype=np.float32)
    assert rel_l2(host(out), ref) <= 1e-5


def test_ray_mask_and_masked_fp(ctx):
    spec = cone_case(nx=32, nz=12, nv=16, nu=64, nrows=40, h=0.5)
    plan = Plan(ctx, spec)
    mask = torch.empty((spec.n_v, spec.n_views, spec.n_u), dtype=torch.float32, device=DEV)
    plan.ray_mask(mask)
    got = host(mask)
    ref = recon.ray_mask(oracle_dict(spec))
    assert 0.05 < (ref == 0).mean() < 0.95
    assert (got != ref).mean() <= 1e-3
    vol = _vol(spec, 3)
    out = torch.zeros_like(mask)
    plan.forward_project(dev(vol), out, mask, 1)
    fp_ref = recon.forward_project(vol, oracle_dict(spec), ray_mask=got, dtype=np.float32)
    assert rel_l2(host(out), fp_ref) <= 1e-5


def test_ray_mask_fan_is_ones(ctx):
    spec = fan_case(nx=16, nv=8, nu=20)
    plan = Plan(ctx, spec)
    mask = torch.zeros((spec.n_v, spec.n_views, spec.n_u), dtype=torch.float32, device=DEV)
    plan.ray_mask(mask)
    assert bool((mask == 1).all())


@pytest.mark.parametrize("kind", ["fan", "cone", "cone_masked", "fan_chunks", "fan_chunks_masked",
                                  "fan_chunks_vh"])
def test_backproject_ir(ctx, kind):
    """fan: 2 slices (one thread per column and view walks the sinogram
    rows); fan_chunks*: 8 slices or more, read from the slice-chunk records
    (11 and 13 slices: a partly filled last chunk; _vh: a voxel height, so a
    weight per slice); the corners of the 48 x 0.6 mm image leave the
    detector (truncated columns)."""
    if kind == "fan":
        spec = fan_case(nx=48, nv=40, nu=70)
    elif kind.startswith("fan_chunks"):
        spec = fan_case(nx=48, nv=40, nu=70, rows={"fan_chunks": 11, "fan_chunks_masked": 13,
                                                   "fan_chunks_vh": 9}[kind], rot_deg=7.0, off=0.9)
        if kind == "fan_chunks_vh":
            spec.vox_h = 0.4
    else:
        spec = cone_case(nx=32, nz=16, nv=30)
    rng = np.random.default_rng(4)
    sgm = rng.standard_normal((spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)
    mask = None
    if kind.endswith("masked"):
        mask = (rng.uniform(0, 1, sgm.shape) > 0.3).astype(np.float32)
    plan = Plan(ctx, spec)
    vol = torch.empty((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32, device=DEV)
    plan.backproject_ir(dev(sgm), None if mask is None else dev(mask), vol)
    got = host(vol)
    ref = recon.backproject_ir(sgm, oracle_dict(spec), ray_mask=mask, dtype=np.float32)
    assert int(((got != 0) & (ref == 0)).sum()) == 0
    if kind.startswith("fan"):
        assert int(((got == 0) & (ref != 0)).sum()) == 0 and 0 < int((ref == 0).sum()) < ref.size // 2
    assert rel_l2(got, ref) <= 1e-5


def test_add_noise_matches_oracle_stream(ctx):
    rng = np.random.default_rng(5)
    p = rng.uniform(0, 4, (3, 50, 70)).astype(np.float32)
    t = dev(p)
    ctx.add_noise(t, 1e5, seed=1234)
    got = host(t)
    ref = recon.add_noise(p, 1e5, 1234)
    assert rel_l2(got, ref) <= 1e-6
    t2 = dev(p)
    ctx.add_noise(t2, 1e5, seed=1234)
    assert np.array_equal(host(t2), got)  # depends only on (seed, index)


def test_dot_and_sq_diff(ctx):
    rng = np.random.default_rng(6)
    a = rng.standard_normal(1_000_003).astype(np.float32)
    b = rng.standard_normal(1_000_003).astype(np.float32)
    m = (rng.uniform(0, 1, a.shape) > 0.5).astype(np.float32)
    out = torch.zeros(1, dtype=torch.float64, device=DEV)
    ctx.dot(dev(a), dev(b), out)
    ref = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
    assert abs(host(out)[0] - ref) <= 1e-12 * np.abs(a.astype(np.float64) * b).sum()
    ctx.sq_diff(dev(a), dev(b), dev(m), out)
    ref = float(np.sum((a.astype(np.float64) - b.astype(np.float64) * m) ** 2))
    assert abs(host(out)[0] / ref - 1) <= 1e-12


def test_dot_deterministic(ctx):
    a = torch.randn(3_000_000, device=DEV)
    o1 = torch.zeros(1, dtype=torch.float64, device=DEV)
    o2 = torch.zeros(1, dtype=torch.float64, device=DEV)
    ctx.dot(a, a, o1)
    torch.randn(10_000_000, device=DEV).sum()  # unrelated work in between
    ctx.dot(a, a, o2)
    assert host(o1)[0] == host(o2)[0]


@pytest.mark.parametrize("nz", [1, 2, 5])
def test_tv_weight_value_dtwd(ctx, nz):
    rng = np.random.default_rng(nz)
    x = rng.standard_normal((nz, 33, 29)).astype(np.float32)
    prior = rng.standard_normal((nz, 33, 29)).astype(np.float32)
    beta = 1e-3
    wr = torch.empty((nz, 33, 29), dtype=torch.float32, device=DEV)
    ctx.tv_weight(dev(x), dev(prior), beta, wr)
    ref_w = recon.tv_weight(x.astype(np.float64) - prior, beta)
    assert rel_l2(host(wr), ref_w) <= 1e-6
    val = torch.zeros(1, dtype=torch.float64, device=DEV)
    ctx.tv_value(dev(x), None, beta, val)
    assert abs(host(val)[0] / recon.tv_value(x.astype(np.float64), beta) - 1) <= 1e-12
    base = rng.standard_normal(x.shape).astype(np.float32)
    out = torch.empty_like(wr)
    wr_np = host(wr)
    ctx.dtwd(dev(x), wr, -0.7, dev(base), out)
    ref = base.astype(np.float64) - 0.7 * recon.dtwd(x.astype(np.float64), wr_np.astype(np.float64))
    assert rel_l2(host(out), ref) <= 1e-6


def test_cg_step_kernels(ctx):
    rng = np.random.default_rng(7)
    n = 100_000
    x, r, d, q = (rng.standard_normal(n).astype(np.float32) for _ in range(4))
    tx, tr, td, tq = dev(x), dev(r), dev(d), dev(q)
    s = torch.zeros(4, dtype=torch.float64, device=DEV)
    ctx.dot(tr, tr, s[0:1])
    ctx.dot(td, tq, s[1:2])
    ctx.cg_update(tx, tr, td, tq, s)
    ctx.cg_direction(td, tr, s)
    r64, d64, q64 = (v.astype(np.float64) for v in (r, d, q))
    alpha = np.dot(r64, r64) / np.dot(d64, q64)
    xn = x + alpha * d64
    rn = (r64 - alpha * q64).astype(np.float32).astype(np.float64)
    beta = np.dot(rn, rn) / np.dot(r64, r64)
    dn = rn + beta * d64
    assert rel_l2(host(tx), xn) <= 1e-7
    assert rel_l2(host(tr), rn) <= 1e-7
    assert rel_l2(host(td), dn) <= 1e-7
    sh = host(s)
    assert abs(sh[0] / np.dot(rn, rn) - 1) <= 1e-12
