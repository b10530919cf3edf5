"""Serving loop: FBP / FDK / CGLS-TV jobs of mixed geometry, continuous
batching, HBM-budget admission, p50/p99 latency.

The reference reconstructs one file per call with a fresh Taichi runtime
(run_mgfbp.py:18-47); a stream of requests there is a queue of such calls.
Here a long-lived server keeps plans (geometry tables), filter taps and
iterate state on the device and overlaps the work of different requests:

  * a job is admitted when its buffers fit the HBM budget, then its sinogram
    is copied host -> device on the copy stream while earlier jobs compute;
    uploads go in arrival order (small_bytes can let small uploads go
    first, h2d_window can bound the bytes in flight);
  * every scheduling step collects the FBP/FDK jobs whose sinogram has
    arrived, weights / filters / packs each, back-projects each cone volume
    in its own launch (its tiles fill the device) and all fan slices in ONE
    persistent launch (mct_backproject_batch), on compute stream A;
  * CGLS-TV jobs run their iterations on their own compute stream (a pool of
    several, round robin, is available), overlapping A; jobs that arrive
    while it is busy wait and then go as stacks of same-geometry jobs, one
    forward and one back-projection per iteration for the whole stack;
  * results go device -> host on the copy stream; buffers return to the
    budget when the copy has finished.

Each job's result is bit-identical alone (run_plain) and in the mix: the
batched launch runs the same tile code as the single-job kernels, a CGLS
stack treats its slices independently with per-job scalars, and the CG
reductions are deterministic.
"""
from __future__ import annotations

import math
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field, fields, replace

import numpy as np
import torch

from . import geometry as geo
from .device import Context, Plan


@dataclass
class Job:
    """One request: a fan FBP, cone FDK or CGLS-TV reconstruction."""

    jid: int
    kind: str  # "fbp" | "fdk" | "cgls"
    spec: geo.GeometrySpec  # back-projection geometry (FBP/FDK) or IR geometry (CGLS)
    sgm_host: torch.Tensor  # pinned float32 (rows, views, n_u)
    taps: np.ndarray | None = None  # ramp taps (FBP/FDK)
    det_w: float = 1.0
    iters: int = 10  # CGLS iterations
    lam: float = 1e-3  # CGLS Lambda * pixel ratio is derived below
    beta_tv: float = 1e-6
    storage: str = "f32"
    filter_method: str = "bf16x3"
    arrival: float = 0.0
    # filled by the server
    t_admit: float = math.nan
    t_done: float = math.nan
    result: torch.Tensor | None = None
    _dev: dict = field(default_factory=dict, repr=False, init=False)
    _events: dict = field(default_factory=dict, repr=False, init=False)

    def fresh(self) -> "Job":
        """The same request, not yet served (for running it again)."""
        return replace(self, t_admit=math.nan, t_done=math.nan, result=None)

    def vol_shape(self):
        return (self.spec.img_dim_z, self.spec.img_dim, self.spec.img_dim)

    def nbytes(self) -> int:
        """Device bytes the job holds while admitted."""
        s = self.spec
        sg = s.n_v * s.n_views * s.n_u * 4
        vol = s.img_dim_z * s.img_dim * s.img_dim * 4
        if self.kind == "cgls":
            return sg * 3 + vol * 7
        tap = 4 if self.storage == "f32" else 2
        packed = (s.n_v + 2) * s.n_views * (s.n_u - 1) * 2 * tap
        return sg * 2 + packed + vol


def geometry_key(spec: geo.GeometrySpec) -> tuple:
    """Every field of the geometry (arrays by dtype, shape and bytes): two
    jobs share a plan, or a stacked CGLS run, only when their geometry is
    the same in every respect the kernels read."""
    key = []
    for f in fields(spec):
        if f.name.startswith("_"):
            continue
        v = getattr(spec, f.name)
        if isinstance(v, np.ndarray):
            v = (v.dtype.str, v.shape, v.tobytes())
        key.append((f.name, v))
    return tuple(key)


class PlanCache:
    """LRU cache of device plans keyed by geometry (per context)."""

    def __init__(self, ctx: Context, capacity: int = 64):
        self.ctx = ctx
        self.capacity = capacity
        self.plans: OrderedDict = OrderedDict()

    def get(self, spec: geo.GeometrySpec) -> Plan:
        k = geometry_key(spec)
        p = self.plans.get(k)
        if p is None:
            p = Plan(self.ctx, spec)
            self.plans[k] = p
            if len(self.plans) > self.capacity:
                _, old = self.plans.popitem(last=False)
                old.close()
        else:
            self.plans.move_to_end(k)
        return p


# ------------------------------------------------------------------ device work
def fbp_stages(ctx: Context, plan: Plan, job: Job, sgm: torch.Tensor):
    """Weighting, ramp filter and repack of one FBP/FDK job on ctx's stream
    (the MFMA filters weigh in their input pass); returns (packed taps,
    layout)."""
    filt = torch.empty_like(sgm)
    if job.filter_method == "f32":
        plan.weight(sgm)
        ctx.filter_rows(sgm, filt, job.taps, job.det_w, job.filter_method)
    else:
        plan.weight_filter(sgm, filt, job.taps, job.det_w, job.filter_method)
    layout = "rows" if job.spec.cone else "pair"
    return ctx.pack(filt, job.storage, layout), layout


def cgls_run(ctx: Context, plan: Plan, job: Job, sgm: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """job.iters CGLS-TV iterations from x = 0 with unit TV weights
    (TikhonovSol, run_mgfbp_iterative_recon_IRN_CGLS.py:188-232), all on
    ctx's stream with FP64 scalars on the device."""
    s = job.spec
    dev = sgm.device
    vshape = job.vol_shape()
    B = {k: torch.empty(vshape, dtype=torch.float32, device=dev) for k in ("bpb", "d", "r", "q", "t", "w")}
    fp = torch.empty_like(sgm)
    mask = torch.empty_like(sgm)
    sc = torch.zeros(4, dtype=torch.float64, device=dev)
    c1 = job.lam * (s.n_v * s.n_views * s.n_u) / (s.img_dim_z * s.img_dim * s.img_dim)
    plan.ray_mask(mask)
    plan.backproject_ir(sgm, mask, B["bpb"])
    x.zero_()
    B["w"].fill_(1.0)

    def apply_A(v, out):
        plan.forward_project(v, fp, mask, 1)
        plan.backproject_ir(fp, None, B["t"])
        ctx.dtwd(v, B["w"], c1, B["t"], out)

    apply_A(x, B["q"])
    ctx.axpby(1.0, B["bpb"], -1.0, B["q"], B["d"])
    B["r"].copy_(B["d"])
    for _ in range(job.iters):
        apply_A(B["d"], B["q"])
        ctx.dot(B["r"], B["r"], sc[0:1])
        ctx.dot(B["d"], B["q"], sc[1:2])
        ctx.cg_update(x, B["r"], B["d"], B["q"], sc)
        ctx.cg_direction(B["d"], B["r"], sc)
    return x


def cgls_stack_spec(spec: geo.GeometrySpec, n: int) -> geo.GeometrySpec:
    """The IR geometry of n fan CGLS jobs of one geometry stacked as the
    slices of one problem (one detector row per slice)."""
    return replace(spec, img_dim_z=n, v=geo.det_pos_array(n, n, 0, 0), n_v_full=n)


def cgls_run_stack(ctx: Context, plan: Plan, jobs: list[Job], sgm: torch.Tensor,
                   x: torch.Tensor) -> torch.Tensor:
    """cgls_run for n fan jobs of one geometry (plan: cgls_stack_spec; sgm
    (n, views, n_u); x (n, N, N)): one forward and one back-projection per
    iteration for all of them, the TV term slice by slice (no coupling
    between jobs), and the inner products, step scalars and updates per job.
    Each job's iterates equal its own cgls_run bit for bit: every kernel
    treats the slices independently, and the forward projector splits each
    ray's steps as it would for one slice (split target scaled by n)."""
    n = len(jobs)
    job = jobs[0]
    s = job.spec
    dev = sgm.device
    vshape = (n, s.img_dim, s.img_dim)
    B = {k: torch.empty(vshape, dtype=torch.float32, device=dev) for k in ("bpb", "d", "r", "q", "t", "w")}
    fp = torch.empty_like(sgm)
    mask = torch.empty_like(sgm)
    sc = torch.zeros((n, 4), dtype=torch.float64, device=dev)
    c1 = job.lam * (s.n_v * s.n_views * s.n_u) / (s.img_dim_z * s.img_dim * s.img_dim)
    ctx.set_fp_split(256 * n)
    try:
        plan.ray_mask(mask)
        plan.backproject_ir(sgm, mask, B["bpb"])
        x.zero_()
        B["w"].fill_(1.0)

        def apply_A(v, out):
            plan.forward_project(v, fp, mask, 1)
            plan.backproject_ir(fp, None, B["t"])
            ctx.dtwd(v, B["w"], c1, B["t"], out, couple_z=False)

        apply_A(x, B["q"])
        ctx.axpby(1.0, B["bpb"], -1.0, B["q"], B["d"])
        B["r"].copy_(B["d"])
        for _ in range(job.iters):
            apply_A(B["d"], B["q"])
            for i in range(n):
                ctx.dot(B["r"][i], B["r"][i], sc[i, 0:1])
                ctx.dot(B["d"][i], B["q"][i], sc[i, 1:2])
                ctx.cg_update(x[i], B["r"][i], B["d"][i], B["q"][i], sc[i])
                ctx.cg_direction(B["d"][i], B["r"][i], sc[i])
    finally:
        ctx.set_fp_split(0)
    return x


def cgls_stack_key(job: Job) -> tuple:
    """Jobs that can share a stacked CGLS run."""
    return (geometry_key(job.spec), job.iters, job.lam)


def run_plain(jobs: list[Job], device: int = 0) -> dict:
    """One job at a time, synchronously, in arrival order: fresh plan,
    allocation, taps upload and host <-> device copies paid per job (the
    reference's per-call pattern); sinograms come from and results go to
    pinned host memory, as in the server.  Returns latency/throughput
    statistics."""
    ctx = Context(device)
    dev = torch.device("cuda", device)
    phases = {"setup": 0.0, "h2d": 0.0, "compute": 0.0, "d2h": 0.0}

    def mark(name, t):
        torch.cuda.synchronize(dev)
        now = time.perf_counter()
        phases[name] += now - t
        return now

    t0 = time.perf_counter()
    for job in sorted(jobs, key=lambda j: j.arrival):
        now = time.perf_counter() - t0
        if now < job.arrival:
            time.sleep(job.arrival - now)
        job.t_admit = time.perf_counter() - t0
        t = time.perf_counter()
        plan = Plan(ctx, job.spec)
        t = mark("setup", t)
        sgm = job.sgm_host.to(dev)
        t = mark("h2d", t)
        if job.kind == "cgls":
            x = torch.empty(job.vol_shape(), dtype=torch.float32, device=dev)
            cgls_run(ctx, plan, job, sgm, x)
        else:
            packed, layout = fbp_stages(ctx, plan, job, sgm)
            x = torch.empty(job.vol_shape(), dtype=torch.float32, device=dev)
            plan.backproject_fbp(packed, job.storage, layout, x)
        t = mark("compute", t)
        # into pinned host memory, as the server's results
        host = torch.empty(job.vol_shape(), dtype=torch.float32, pin_memory=True)
        host.copy_(x)
        job.result = host
        mark("d2h", t)
        plan.close()
        job.t_done = time.perf_counter() - t0
    ctx.close()
    out = stats(jobs)
    out["phase_s"] = {k: round(v, 4) for k, v in phases.items()}
    return out


class Server:
    """Continuous-batching server on one device (see the module comment)."""

    def __init__(self, device: int = 0, hbm_budget: int = 64 << 30, max_batch: int = 64,
                 cgls_streams: int = 1, h2d_window: int = 1 << 40, small_bytes: int = 0,
                 max_cgls_stack: int = 16):
        self.device = device
        self.dev = torch.device("cuda", device)
        self.s_copy = torch.cuda.Stream(self.dev)  # host <-> device copies
        self.s_a = torch.cuda.Stream(self.dev)
        self.ctx_a = Context(device)
        self.ctx_a.bind_stream(self.s_a)
        self.plans_a = PlanCache(self.ctx_a)
        # CGLS jobs run on their own stream(s) (one context and plan cache
        # each), beside stream A.  One stream measured best on the mixed burst
        # (210 jobs/s against 140-200 with two or three): more CGLS streams
        # take device time from the FDK work that bounds the makespan.
        self.s_b = [torch.cuda.Stream(self.dev) for _ in range(max(1, cgls_streams))]
        self.ctx_b = [Context(device) for _ in self.s_b]
        for c, s in zip(self.ctx_b, self.s_b):
            c.bind_stream(s)
        self.plans_b = [PlanCache(c) for c in self.ctx_b]
        self._next_b = 0
        # CGLS jobs that are loaded wait per stream until it is idle, then go
        # as stacks of same-geometry jobs (one forward and back-projection
        # per iteration for the whole stack).  At most 63 jobs: a fan problem
        # of 64 or more slices takes the one-ray-per-wavefront projector,
        # whose sums are grouped differently from a single slice's.
        self.max_cgls_stack = min(63, max(1, int(max_cgls_stack)))
        self._cgls_wait: list[list[Job]] = [[] for _ in self.s_b]
        self._cgls_busy: list = [None for _ in self.s_b]
        self.cgls_stacks: list[int] = []  # size of every CGLS run launched
        self.budget = int(hbm_budget)
        self.used = 0
        self.max_batch = max_batch
        self.h2d_window = int(h2d_window)  # bytes of sinogram uploads in flight
        # uploads up to this size may go ahead of larger ones (0: arrival
        # order, which measured best on the mixed burst: 210 jobs/s against
        # 190-200 with 16-64 MB)
        self.small_bytes = int(small_bytes)
        self.rejected: list[Job] = []

    def _admit(self, job: Job, now: float) -> bool:
        need = job.nbytes()
        if need > self.budget:
            self.rejected.append(job)  # can never fit: reject instead of blocking the queue
            job.t_done = now
            return True
        if self.used + need > self.budget:
            return False
        self.used += need
        job.t_admit = now
        if job.kind == "cgls":
            job._dev["lane"] = self._next_b
            self._next_b = (self._next_b + 1) % len(self.s_b)
        with torch.cuda.stream(self.s_copy):
            sgm = job.sgm_host.to(self.dev, non_blocking=True)
            # used on a compute stream later: keep the allocator from recycling
            # the block before that use has finished
            sgm.record_stream(self.s_b[job._dev["lane"]] if job.kind == "cgls" else self.s_a)
            job._dev["sgm"] = sgm
            ev = torch.cuda.Event()
            ev.record(self.s_copy)
        job._events["loaded"] = ev
        return True

    def _launch_fbp_batch(self, batch: list[Job]):
        by_dtype: dict[str, list] = {}
        s = self.s_a
        with torch.cuda.stream(s):
            for job in batch:
                s.wait_event(job._events["loaded"])
                sgm = job._dev["sgm"]
                plan = self.plans_a.get(job.spec)
                packed, layout = fbp_stages(self.ctx_a, plan, job, sgm)
                vol = torch.empty(job.vol_shape(), dtype=torch.float32, device=self.dev)
                job._dev.update(packed=packed, vol=vol)
                if job.spec.cone:
                    # a cone volume's tiles fill the device on their own; its
                    # single launch runs the LDS kernel without the batch
                    # kernel's per-item job lookup and register budget
                    plan.backproject_fbp(packed, job.storage, layout, vol)
                else:
                    by_dtype.setdefault(job.storage, []).append((plan, packed, job.storage, layout, vol))
            for group in by_dtype.values():
                for i in range(0, len(group), self.max_batch):
                    self.ctx_a.backproject_batch(group[i:i + self.max_batch])
            done = torch.cuda.Event()
            done.record(s)
        self._copy_back(batch, done)

    def _launch_cgls(self, jobs: list[Job], lane: int) -> torch.cuda.Event:
        """One stacked CGLS run of same-geometry jobs on CGLS stream `lane`."""
        s = self.s_b[lane]
        n = len(jobs)
        self.cgls_stacks.append(n)
        with torch.cuda.stream(s):
            for job in jobs:
                s.wait_event(job._events["loaded"])
            spec = jobs[0].spec
            plan = self.plans_b[lane].get(cgls_stack_spec(spec, n) if n > 1 else spec)
            sgm = torch.cat([j._dev["sgm"] for j in jobs]) if n > 1 else jobs[0]._dev["sgm"]
            x = torch.empty((n, spec.img_dim, spec.img_dim), dtype=torch.float32, device=self.dev)
            if n > 1:
                cgls_run_stack(self.ctx_b[lane], plan, jobs, sgm, x)
            else:
                cgls_run(self.ctx_b[lane], plan, jobs[0], sgm, x)
            for i, job in enumerate(jobs):
                job._dev["vol"] = x[i:i + 1]
            done = torch.cuda.Event()
            done.record(s)
        self._copy_back(jobs, done)
        return done

    def _pump_cgls(self):
        """Starts the waiting CGLS jobs of every idle CGLS stream, grouped
        into stacks of one geometry."""
        for lane, wait in enumerate(self._cgls_wait):
            busy = self._cgls_busy[lane]
            if not wait or (busy is not None and not busy.query()):
                continue
            groups: dict = {}
            for job in wait:
                groups.setdefault(cgls_stack_key(job), []).append(job)
            wait.clear()
            for group in groups.values():
                for i in range(0, len(group), self.max_cgls_stack):
                    self._cgls_busy[lane] = self._launch_cgls(group[i:i + self.max_cgls_stack], lane)

    def _copy_back(self, jobs: list[Job], computed: torch.cuda.Event):
        with torch.cuda.stream(self.s_copy):
            self.s_copy.wait_event(computed)
            for job in jobs:
                host = torch.empty(job.vol_shape(), dtype=torch.float32, pin_memory=True)
                job._dev["vol"].record_stream(self.s_copy)
                host.copy_(job._dev["vol"], non_blocking=True)
                job.result = host
                ev = torch.cuda.Event()
                ev.record(self.s_copy)
                job._events["done"] = ev

    def run(self, jobs: list[Job]) -> dict:
        """Serves the jobs at their arrival times (seconds from start)."""
        order = deque(sorted(jobs, key=lambda j: j.arrival))
        pending: deque[Job] = deque()
        loading: list[Job] = []
        running: list[Job] = []
        t0 = time.perf_counter()
        while order or pending or loading or running:
            now = time.perf_counter() - t0
            arrived = False
            while order and order[0].arrival <= now:
                pending.append(order.popleft())
                arrived = True
            if arrived and self.small_bytes > 0 and len(pending) > 1:
                # small uploads (fan slices, CGLS jobs) go ahead of large ones,
                # arrival order within each class: they cost the large jobs
                # next to nothing and start the device (and the CGLS stream)
                # while the large sinograms are in flight
                pending = deque(sorted(pending, key=lambda j: j.sgm_host.numel() * 4 > self.small_bytes))
            # Uploads are issued in arrival order while less than h2d_window
            # bytes are in flight (unbounded by default: bounding it delayed
            # the large jobs more than it helped the small ones).
            inflight = sum(j.sgm_host.numel() * 4 for j in loading)
            while pending and (not loading or inflight + pending[0].sgm_host.numel() * 4 <= self.h2d_window) \
                    and self._admit(pending[0], now):
                job = pending.popleft()
                inflight += job.sgm_host.numel() * 4
                if not math.isnan(job.t_done):
                    continue  # rejected
                loading.append(job)
            ready = [j for j in loading if j._events["loaded"].query()]
            if ready:
                loading = [j for j in loading if j not in ready]
                fb = [j for j in ready if j.kind != "cgls"]
                if fb:
                    self._launch_fbp_batch(fb)
                for j in ready:
                    if j.kind == "cgls":
                        self._cgls_wait[j._dev["lane"]].append(j)
                running.extend(ready)
            self._pump_cgls()
            finished = [j for j in running if "done" in j._events and j._events["done"].query()]
            for j in finished:
                j.t_done = time.perf_counter() - t0
                self.used -= j.nbytes()
                j._dev.clear()
                j._events.clear()
            if finished:
                running = [j for j in running if j not in finished]
            if not (ready or finished):
                wait = (order[0].arrival - now) if (order and not pending and not loading and not running) else 0.0
                time.sleep(min(max(wait, 0.0), 0.002) if wait > 0 else 0.0002)
        torch.cuda.synchronize(self.dev)
        out = stats([j for j in jobs if j not in self.rejected])
        out["rejected"] = len(self.rejected)
        return out

    def close(self):
        for s in (self.s_copy, self.s_a, *self.s_b):
            s.synchronize()
        for pc in (self.plans_a, *self.plans_b):
            for p in pc.plans.values():
                p.close()
            pc.plans.clear()
        self.ctx_a.close()
        for c in self.ctx_b:
            c.close()


def stats(jobs: list[Job]) -> dict:
    lat = np.array([j.t_done - j.arrival for j in jobs])
    t_end = max(j.t_done for j in jobs)
    t_start = min(j.arrival for j in jobs)
    return {"jobs": len(jobs), "jobs_per_s": len(jobs) / max(t_end - t_start, 1e-9),
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "makespan_s": float(t_end - t_start)}


# ------------------------------------------------------------------ synthetic mix
def make_mix(n_jobs: int, qps: float, seed: int = 3, scale: float = 1.0, device: int = 0,
             storage: str = "f32", kinds=("fbp256", "fbp512", "fbp1024", "fdk256", "fdk512", "cgls512"),
             prototypes_per_kind: int = 2) -> list[Job]:
    """Poisson arrivals at qps; kinds drawn uniformly.  Geometry prototypes
    (seeded) share one analytic ellipse/ellipsoid sinogram each, to bound
    host memory.  scale < 1 shrinks every size (tests)."""
    from . import phantom

    rng = np.random.default_rng(seed)
    dev = torch.device("cuda", device)
    protos: dict = {}

    def proto(kind, k):
        key = (kind, k)
        if key in protos:
            return protos[key]
        prng = np.random.default_rng(seed * 1000 + 17 * len(protos) + k)
        fov = 153.6  # mm across the image
        if kind.startswith("fbp") or kind.startswith("cgls"):
            n = max(int(int(kind.lstrip("fbpcgls")) * scale), 16)
            if kind.startswith("fbp"):
                views = max(int(prng.integers(360, 1441) * scale), 24)
                nu = max(int(prng.integers(400, 1601) * scale), 24)
            else:
                views, nu = max(int(240 * scale), 24), max(int(1.6 * n), 24)
            pix = fov / n
            sid, sdd = 500.0, 1000.0
            w = fov * 1.42 * (sdd / sid) * 1.05 / nu  # the detector covers the image diagonal
            ang = geo.angle_array(views, 2 * geo.PI, 0.0)
            u = geo.det_pos_array(nu, nu, -w, 0.0)
            v = geo.det_pos_array(1, 1, 0, 0)
            spec = geo.GeometrySpec(img_dim=n, img_dim_z=1, pix=pix, vox_h=0.0, n_views=views, sid=sid,
                                    sdd=sdd, total_scan_angle=2 * geo.PI, angles=ang, u=u, v=v,
                                    n_v_full=1)
            ell = phantom.random_ellipsoids(10, seed=int(prng.integers(1 << 30)), r_fov=0.4 * fov)
            sg = phantom.project(ell, ang, u, v, sid, sdd, device=dev)
        else:
            n = max(int(int(kind[3:]) * scale), 16)
            views = max(int(prng.integers(360, 1081) * scale), 24)
            nu = max(min(1024, 2 * n), 24)
            nr = max(min(256, n // 2), 8)
            pix = vh = fov / n
            sid, sdd = 750.0, 1250.0
            w = fov * 1.42 * (sdd / sid) * 1.05 / nu
            h = fov * 0.8 * (sdd / sid) / nr  # the cone covers 80% of the volume height
            ang = geo.angle_array(views, 2 * geo.PI, 0.0)
            u = geo.det_pos_array(nu, nu, -w, 0.0)
            v = geo.det_pos_array(nr, nr, h, 0.0)
            spec = geo.GeometrySpec(img_dim=n, img_dim_z=n, pix=pix, vox_h=vh, n_views=views, sid=sid,
                                    sdd=sdd, total_scan_angle=2 * geo.PI, angles=ang, u=u, v=v, cone=True,
                                    n_v_full=nr, det_h=h)
            ell = phantom.random_ellipsoids(20, seed=int(prng.integers(1 << 30)), r_fov=0.4 * fov,
                                            z_half=0.3 * fov)
            sg = phantom.project(ell, ang, u, v, sid, sdd, device=dev)
        host = torch.empty(sg.shape, dtype=torch.float32, pin_memory=True)
        host.copy_(sg)
        taps = geo.hamming_taps(spec.n_u, abs(float(spec.u[1] - spec.u[0])), 1.0, spec.sdd)
        protos[key] = (spec, host, taps)
        return protos[key]

    jobs = []
    t = 0.0
    for i in range(n_jobs):
        t += rng.exponential(1.0 / qps)
        kind = kinds[int(rng.integers(len(kinds)))]
        spec, host, taps = proto(kind, int(rng.integers(prototypes_per_kind)))
        w = abs(float(spec.u[1] - spec.u[0]))
        jk = "cgls" if kind.startswith("cgls") else ("fdk" if kind.startswith("fdk") else "fbp")
        jobs.append(Job(jid=i, kind=jk, spec=spec, sgm_host=host, taps=taps, det_w=w, arrival=t,
                        storage=storage, iters=10))
    torch.cuda.synchronize(dev)
    return jobs
