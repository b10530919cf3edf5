"""Graph capture of the batched back-projection and of the FDK pipeline.

The library's launch paths queue only kernels once their tables and tap
caches exist, so they can be captured into a HIP graph (torch.cuda.CUDAGraph)
and replayed.  Each test replays several times, with eager work and
synchronisations between the replays, poisons the outputs before every
replay and checks every replay against the eager result, bit for bit.
"""
import numpy as np
import pytest
import torch

from oracle import recon
from reconstruction_program_taichi_amd.device import Plan
from tests._cases import cone_case, fan_case, rel_l2

pytestmark = pytest.mark.gpu


def _capture(ctx, fn):
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    try:
        with torch.cuda.graph(g):
            ctx.bind_stream()  # the capture stream
            fn()
    finally:
        ctx.bind_stream(torch.cuda.current_stream())
    return g


def test_batched_backprojection_graph_replays(ctx):
    specs = [fan_case(nx=48, nv=60, nu=80, rows=3), cone_case(nx=40, nz=21, nv=50),
             cone_case(nx=56, nz=37, nv=64, top_row=True), fan_case(nx=33, nv=45, nu=64, rows=1)]
    rng = np.random.default_rng(5)
    jobs = []
    for s in specs:
        layout = "rows" if s.cone else "pair"
        f = torch.from_numpy(rng.standard_normal((s.n_v, s.n_views, s.n_u)).astype(np.float32)).cuda()
        plan = Plan(ctx, s)
        packed = ctx.pack(f, "f32", layout)
        vol = torch.empty((s.img_dim_z, s.img_dim, s.img_dim), dtype=torch.float32, device="cuda")
        jobs.append((plan, packed, "f32", layout, vol))
    batch = ctx.prepare_batch(jobs)
    batch.launch()
    torch.cuda.synchronize()
    want = [j[4].clone() for j in jobs]
    # the prepared batch equals the one-shot batched launch
    outs = [torch.empty_like(w) for w in want]
    ctx.backproject_batch([(p, f, d, l, o) for (p, f, d, l, _), o in zip(jobs, outs)])
    torch.cuda.synchronize()
    for w, o in zip(want, outs):
        assert torch.equal(w, o)
    g = _capture(ctx, batch.launch)
    for rep in range(3):
        for j in jobs:
            j[4].fill_(float("nan"))
        junk = torch.randn(1 << 16, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        for w, j in zip(want, jobs):
            assert torch.equal(j[4], w), f"replay {rep}"
    batch.close()


@pytest.mark.parametrize("method", ["bf16x3", "f32"])
def test_fdk_pipeline_graph_replays(ctx, method):
    spec = cone_case(nu=80, nrows=32, nz=30)
    rng = np.random.default_rng(7)
    raw = torch.from_numpy(rng.uniform(0, 1, (spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)).cuda()
    taps = recon.ramlak_taps(spec.n_u, 0.6, np.float32)
    plan = Plan(ctx, spec)
    work = torch.empty_like(raw)
    filt = torch.empty_like(raw)
    nbytes = ctx.pack(raw, "f32", "rows").numel()
    packed = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    vol = torch.empty((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32, device="cuda")

    def pipeline():
        plan.weight(work)
        ctx.filter_rows(work, filt, taps, 0.6, method)
        ctx.pack_into(filt, "f32", "rows", packed)
        plan.backproject_fbp(packed, "f32", "rows", vol)

    work.copy_(raw)
    pipeline()  # eager: builds the tap caches and the back-projector's tables
    torch.cuda.synchronize()
    want = vol.clone()
    assert ctx.bp_last_kernel() == "lds"
    g = _capture(ctx, pipeline)
    for rep in range(3):
        vol.fill_(float("nan"))
        work.copy_(raw)  # the weighting runs in place
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(vol, want), f"replay {rep}"
    # and the eager result is the oracle's
    w = recon.weight_sinogram(raw.cpu().numpy(), spec.u, spec.v, spec.angles, spec.sdd,
                              spec.total_scan_angle, False, dtype=np.float32)
    fr = recon.filter_rows(w, taps, 0.6, np.float32)
    ref = recon.backproject_fbp(fr, _oracle(spec), np.float32)
    assert rel_l2(want.cpu().numpy(), ref) <= 1e-4


def _oracle(spec):
    from tests._cases import oracle_dict

    return oracle_dict(spec)


@pytest.mark.parametrize("method", ["bf16x3", "f32"])
def test_frozen_context_keeps_a_captured_pipeline_valid(method):
    """While the context is frozen, calls that would grow its scratch or
    rewrite its cached taps raise MctError (the captured graph addresses
    them); same-size calls run, and the graph keeps replaying the eager
    result.  Unfrozen, the larger calls run again."""
    from reconstruction_program_taichi_amd._abi import MctError
    from reconstruction_program_taichi_amd.device import Context

    ctx = Context(0)
    spec = cone_case(nu=80, nrows=32, nz=30)
    rng = np.random.default_rng(9)
    raw = torch.from_numpy(rng.uniform(0, 1, (spec.n_v, spec.n_views, spec.n_u)).astype(np.float32)).cuda()
    taps = recon.ramlak_taps(spec.n_u, 0.6, np.float32)
    plan = Plan(ctx, spec)
    work, filt = torch.empty_like(raw), torch.empty_like(raw)
    packed = torch.empty(ctx.pack(raw, "f32", "rows").numel(), dtype=torch.uint8, device="cuda")
    vol = torch.empty((spec.img_dim_z, spec.img_dim, spec.img_dim), dtype=torch.float32, device="cuda")

    def pipeline():
        plan.weight(work)
        ctx.filter_rows(work, filt, taps, 0.6, method)
        ctx.pack_into(filt, "f32", "rows", packed)
        plan.backproject_fbp(packed, "f32", "rows", vol)

    work.copy_(raw)
    pipeline()
    torch.cuda.synchronize()
    want = vol.clone()
    g = _capture(ctx, pipeline)
    big = torch.rand((4 * spec.n_v, spec.n_views, spec.n_u), device="cuda")
    big_out = torch.empty_like(big)
    fp_out = torch.empty((spec.n_v, spec.n_views, spec.n_u), device="cuda")
    ctx.freeze()
    try:
        work.copy_(raw)
        pipeline()  # the captured sizes and taps: runs
        torch.cuda.synchronize()
        assert torch.equal(vol, want)
        with pytest.raises(MctError, match="frozen"):
            ctx.filter_rows(work, filt, taps * 2, 0.6, method)  # other taps
        if method != "f32":
            with pytest.raises(MctError, match="frozen"):
                ctx.filter_rows(big, big_out, taps, 0.6, method)  # larger input planes
        with pytest.raises(MctError, match="frozen"):
            plan.forward_project(vol, fp_out)  # no forward-projector scratch yet
        for rep in range(3):
            vol.fill_(float("nan"))
            work.copy_(raw)
            junk = torch.randn(1 << 16, device="cuda").sum()
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert float(junk) == float(junk)
            assert torch.equal(vol, want), f"replay {rep}"
    finally:
        ctx.freeze(False)
    del g
    ctx.filter_rows(big, big_out, taps * 2, 0.6, method)
    plan.forward_project(vol, fp_out)
    torch.cuda.synchronize()
    plan.close()
    ctx.close()
