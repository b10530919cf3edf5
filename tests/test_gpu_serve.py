"""GPU tests of batched back-projection and the serving loop.

Tolerance: exact.  The batched launch runs the single-job tile code and the
CG reductions are deterministic, so a job's result must be bit-identical
alone and in a batch / in the served mix.
"""
import numpy as np
import pytest
import torch

from oracle import recon
from reconstruction_program_taichi_amd import serve
from reconstruction_program_taichi_amd.device import Context, Plan
from tests._cases import cone_case, fan_case, oracle_dict, rel_l2

pytestmark = pytest.mark.gpu


def _single(ctx, spec, f, dtype, cone_layout="rows"):
    layout = cone_layout if spec.cone else "pair"
    plan = Plan(ctx, spec)
    packed = ctx.pack(f, dtype, layout)
    vol = torch.empty((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32, device="cuda")
    plan.backproject_fbp(packed, dtype, layout, vol)
    return plan, packed, layout, vol


@pytest.mark.parametrize("cone_layout", ["rows", "quad"])
@pytest.mark.parametrize("kernel", ["auto", "direct"])
@pytest.mark.parametrize("dtype", ["f32", "f16"])
def test_batched_bp_equals_single_jobs(ctx, dtype, kernel, cone_layout):
    """The cone jobs run the LDS-staged tile (auto) or the direct one, in the
    batch and alone alike; with FP32 taps every batched volume also matches
    the FP32 reference-order oracle (rel-L2 <= 1e-5, the same voxels masked)."""
    specs = [fan_case(nx=48, nv=60, nu=80, rows=3), cone_case(nx=40, nz=21, nv=50),
             fan_case(nx=33, nv=45, nu=64, rows=1, scan_deg=200.0), cone_case(nx=56, nz=37, nv=64, top_row=True)]
    rng = np.random.default_rng(0)
    singles, jobs, fhost = [], [], []
    ctx.set_bp_kernel(kernel)
    try:
        for s in specs:
            fhost.append(rng.standard_normal((s.n_v, s.n_views, s.n_u)).astype(np.float32))
            f = torch.from_numpy(fhost[-1]).cuda()
            plan, packed, layout, vol = _single(ctx, s, f, dtype, cone_layout)
            if s.cone:
                assert ctx.bp_last_kernel() == ("lds" if kernel == "auto" else "direct")
            singles.append(vol.clone())
            out = torch.full_like(vol, 7.0)
            jobs.append((plan, packed, dtype, layout, out))
        ctx.backproject_batch(jobs)
        torch.cuda.synchronize()
    finally:
        ctx.set_bp_kernel("auto")
    for want, job in zip(singles, jobs):
        assert torch.equal(job[4], want)
    if dtype == "f32":
        for s, f, job in zip(specs, fhost, jobs):
            got = job[4].cpu().numpy()
            ref = recon.backproject_fbp(f, oracle_dict(s), np.float32)
            assert np.array_equal(got == 0, ref == 0)
            assert rel_l2(got, ref) <= 1e-5


def test_batched_bp_rejects_mixed_dtypes(ctx):
    from reconstruction_program_taichi_amd._abi import MctError

    s = fan_case(nx=16, nv=20, nu=24, rows=1)
    f = torch.rand((1, 20, 24), device="cuda")
    plan, p32, layout, vol = _single(ctx, s, f, "f32")
    p16 = ctx.pack(f, "f16", layout)
    with pytest.raises(MctError):
        ctx.backproject_batch([(plan, p32, "f32", layout, vol), (plan, p16, "f16", layout, vol)])


def _small_mix(n=14, seed=11):
    return serve.make_mix(n, qps=400.0, seed=seed, scale=0.12, prototypes_per_kind=1)


def test_served_results_equal_plain_results():
    jobs = _small_mix()
    plain = [j.fresh() for j in jobs]
    srv = serve.Server(hbm_budget=8 << 30)
    try:
        st = srv.run(jobs)
    finally:
        srv.close()
    sp = serve.run_plain(plain)
    assert st["jobs"] == len(jobs) and st["rejected"] == 0
    kinds = set()
    for a, b in zip(jobs, plain):
        kinds.add(a.kind)
        assert a.result is not None and b.result is not None
        assert torch.equal(a.result, b.result), f"job {a.jid} ({a.kind}) differs"
        assert bool(torch.isfinite(a.result).all())
    assert kinds == {"fbp", "fdk", "cgls"}
    print(f"served {st}\nplain {sp}")


def test_admission_waits_for_budget_and_rejects_oversized():
    jobs = _small_mix(n=8, seed=12)
    big = max(j.nbytes() for j in jobs)
    # room for the largest job only: jobs run one after another, none rejected
    srv = serve.Server(hbm_budget=big)
    try:
        st = srv.run(jobs)
    finally:
        srv.close()
    assert st["rejected"] == 0 and all(j.result is not None for j in jobs)
    # a budget below the largest job rejects exactly the jobs that cannot fit
    jobs2 = _small_mix(n=8, seed=12)
    limit = sorted(j.nbytes() for j in jobs2)[len(jobs2) // 2]
    srv = serve.Server(hbm_budget=limit)
    try:
        st = srv.run(jobs2)
    finally:
        srv.close()
    too_big = [j for j in jobs2 if j.nbytes() > limit]
    assert st["rejected"] == len(too_big) > 0
    assert all(j.result is None for j in too_big)
    assert all(j.result is not None for j in jobs2 if j.nbytes() <= limit)


def _cgls_jobs(n, seed=21):
    """n CGLS jobs of one geometry with different sinograms (the prototype's
    scaled by 1 + 0.1 i), arriving together."""
    from dataclasses import replace

    proto = [j for j in serve.make_mix(4 * n + 8, qps=400.0, seed=seed, scale=0.12,
                                       kinds=("cgls512",), prototypes_per_kind=1)][0]
    jobs = []
    for i in range(n):
        host = (proto.sgm_host * (1.0 + 0.1 * i)).pin_memory()
        jobs.append(replace(proto.fresh(), jid=i, sgm_host=host, arrival=0.0))
    return jobs


@pytest.mark.parametrize("n", [3, 5, 9, 16])
def test_stacked_cgls_equals_single_runs(n):
    """n CGLS jobs run as one stacked problem give each job's own iterates
    bit for bit (slices independent, per-job scalars, the same ray split as
    one slice: with 5 slices the forward projector's row bands are padded,
    and the split still follows the rays; from 8 slices on the
    back-projector reads slice-chunk records, 9: a partly filled chunk)."""
    jobs = _cgls_jobs(n)
    ctx = Context(0)
    dev = torch.device("cuda", 0)
    single = []
    for j in jobs:
        plan = Plan(ctx, j.spec)
        x = torch.empty(j.vol_shape(), dtype=torch.float32, device=dev)
        serve.cgls_run(ctx, plan, j, j.sgm_host.to(dev), x)
        single.append(x.clone())
        plan.close()
    plan = Plan(ctx, serve.cgls_stack_spec(jobs[0].spec, n))
    sgm = torch.cat([j.sgm_host.to(dev) for j in jobs])
    x = torch.empty((n,) + tuple(jobs[0].vol_shape()[1:]), dtype=torch.float32, device=dev)
    serve.cgls_run_stack(ctx, plan, jobs, sgm, x)
    torch.cuda.synchronize()
    for i in range(n):
        assert torch.equal(x[i:i + 1], single[i]), f"job {i} differs"
    assert not torch.equal(single[0], single[1])
    plan.close()
    ctx.close()


def test_server_stacks_cgls_jobs():
    """A burst of same-geometry CGLS jobs: the server runs them as stacks,
    and every result equals the plain one-job-at-a-time result."""
    jobs = _cgls_jobs(6, seed=22)
    plain = [j.fresh() for j in jobs]
    srv = serve.Server(hbm_budget=8 << 30)
    try:
        st = srv.run(jobs)
        stacks = list(srv.cgls_stacks)
    finally:
        srv.close()
    serve.run_plain(plain)
    assert st["jobs"] == 6 and sum(stacks) == 6 and max(stacks) > 1, stacks
    for a, b in zip(jobs, plain):
        assert torch.equal(a.result, b.result), f"job {a.jid} differs"


def test_jobs_differing_in_one_geometry_field_do_not_share_plans():
    """Two FDK jobs that differ only in the image centre along z, and two FBP
    jobs that differ only in the short-scan flag: the server keys plans on
    the whole geometry, so each result equals its own plain run (and the two
    of a pair differ)."""
    from dataclasses import replace

    mix = serve.make_mix(24, qps=400.0, seed=13, scale=0.12, kinds=("fdk256", "fbp256"),
                         prototypes_per_kind=1)
    from reconstruction_program_taichi_amd import geometry as geo

    fdk = next(j for j in mix if j.kind == "fdk")
    fbp = next(j for j in mix if j.kind == "fbp")
    # a 450-degree scan: the short-scan flag switches the redundancy ramp on
    scan = 2.5 * geo.PI
    fan = replace(fbp.spec, total_scan_angle=scan, angles=geo.angle_array(fbp.spec.n_views, scan, 0.0))
    jobs = [replace(fdk.fresh(), jid=0, arrival=0.0),
            replace(fdk.fresh(), jid=1, arrival=0.0, spec=replace(fdk.spec, cz=0.75 * fdk.spec.vox_h)),
            replace(fbp.fresh(), jid=2, arrival=0.0, spec=fan),
            replace(fbp.fresh(), jid=3, arrival=0.0, spec=replace(fan, short_scan=True))]
    plain = [j.fresh() for j in jobs]
    srv = serve.Server(hbm_budget=8 << 30)
    try:
        srv.run(jobs)
    finally:
        srv.close()
    serve.run_plain(plain)
    for a, b in zip(jobs, plain):
        assert torch.equal(a.result, b.result), f"job {a.jid} differs from its plain run"
    assert not torch.equal(jobs[0].result, jobs[1].result)
    assert not torch.equal(jobs[2].result, jobs[3].result)
