"""
Benchmark: differentiable rasterisation fwd+bwd on one MI355X.

Workload (one "step"): the noisy 82k-triangle icosphere (subdivision 6,
40,962 vertices, radius 0.8, N(0, 0.01) vertex noise, seed 0) rendered from 64
Fibonacci-sphere cameras at distance 3.2, 512x512, 8 per-vertex float32
attributes U(0,1) (seed 1):

    pos  = clip_space(verts, VP)                  [64, V, 4]
    rast = rasterize(pos, tri)                    [64, 512, 512, 4]
    col  = interpolate(attr, rast, tri)           [64, 512, 512, 8]
    y    = antialias(col, rast, pos, tri)
    L    = sum((y - target)^2),  target U(0,1) (seed 2)
    L.backward()   -> d/d verts, d/d attr

"batched": the whole step captured once in a HIP graph and replayed (one launch
per operator for all 64 views).  "plain": a Python loop over the 64 views, each
an eager single-view fwd+bwd (diffrp's one-session-per-view usage).
Inputs are resident in HBM before the timed region.

Prints ONE JSON line.  ``--gpus`` other than 1 exits: this build measures one GPU.
"""
import argparse
import hashlib
import faulthandler
import json
import os
import platform
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from diffrp_amd import autotune, nvdr  # noqa: E402
from diffrp_amd.loss import l2_loss  # noqa: E402
from diffrp_amd.camera import PerspectiveCamera, clip_space  # noqa: E402
from diffrp_amd.geometry import fibonacci_directions, noisy_icosphere  # noqa: E402

METRIC = 'fwd+bwd Mpixel/s @512² 82k tris, batched-64-view vs plain; % HBM roofline'
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(n_views, H, W, V, F, C, layers=1):
    """Bytes each operator must move once (fwd+bwd), per view-layer:
    (128 + 28 C) N + (96 + 8 C) V + 72 F, plus 16 N per extra peel layer."""
    N = H * W
    per_view = (128 + 28 * C) * N + (96 + 8 * C) * V + 72 * F + 16 * N * (layers - 1)
    return n_views * per_view


# per-operator algorithmic bytes for one launch over n views (fwd or bwd half)
def op_bytes(op, n, N, V, F, C):
    if op == 'interpolate_rasterize_bwd':  # the two backward operators' bytes: the fused pass replaces both calls
        return op_bytes('interpolate_bwd', n, N, V, F, C) + op_bytes('rasterize_bwd', n, N, V, F, C)
    return {
        'rasterize_fwd': n * (16 * N + 16 * V + 12 * F),
        'rasterize_bwd': n * (32 * N + 32 * V + 12 * F),
        'interpolate_fwd': n * (16 * N + 4 * C * N + 12 * F) + 4 * C * V,
        'interpolate_bwd': n * (4 * C * N + 16 * N + 16 * N + 12 * F) + 4 * C * V,
        'antialias_fwd': n * (4 * C * N + 16 * N + 4 * C * N + 16 * V + 12 * F),
        'antialias_bwd': n * (4 * C * N + 4 * C * N + 16 * N + 4 * C * N + 16 * V + 12 * F + 16 * V),
    }[op]


class Scene:
    def __init__(self, dev, views, res, subdiv, C, sigma=0.01, radius=0.8, cpu_only=False):
        verts, tri = noisy_icosphere(subdiv, radius, sigma, seed=0)
        self.V, self.F = verts.shape[0], tri.shape[0]
        self.H = self.W = res
        self.C = C
        self.n = views
        g1 = torch.Generator().manual_seed(1)
        g2 = torch.Generator().manual_seed(2)
        attr = torch.rand(self.V, C, generator=g1)
        target = torch.rand(views, res, res, C, generator=g2)
        VP = []
        for d in fibonacci_directions(views):
            cam = PerspectiveCamera.looking_at_origin_from(d * 3.2, h=res, w=res)
            VP.append(cam.P('cpu') @ cam.V('cpu'))
        self.verts_cpu, self.tri_cpu, self.attr_cpu, self.VP_cpu = verts, tri, attr, torch.stack(VP)
        self.target_cpu = target
        if cpu_only:
            return
        self.verts = verts.to(dev).requires_grad_(True)
        self.attr = attr.to(dev).requires_grad_(True)
        self.tri = tri.to(dev)
        self.VP = torch.stack(VP).to(dev)
        self.target = target.to(dev)
        # the antialias per-vertex / per-triangle state is prepared beside the rasterizer
        self.ctx = nvdr.RasterizeCudaContext(dev, prepare_antialias=True)
        self.topo = nvdr.antialias_construct_topology_hash(self.tri)

    def step_batched(self):
        pos = clip_space(self.verts, self.VP)
        rast, _ = nvdr.rasterize(self.ctx, pos, self.tri, (self.H, self.W), grad_db=False)
        col, _ = nvdr.interpolate(self.attr, rast, self.tri)
        y = nvdr.antialias(col, rast, pos, self.tri, topology_hash=self.topo)
        loss = l2_loss(y, self.target)
        loss.backward()
        # the step's buffers, detached (a captured step's buffers hold each replay's values; keeping
        # the tensors themselves would keep this step's autograd graph, and the AccumulateGrad
        # nodes of the leaves on the stream that made them, alive into the next step or capture)
        self.last = (pos.detach(), rast.detach(), col.detach(), y.detach())
        return loss

    def step_plain(self):
        total = 0.0
        for v in range(self.n):
            pos = clip_space(self.verts, self.VP[v])
            rast, _ = nvdr.rasterize(self.ctx, pos, self.tri, (self.H, self.W), grad_db=False)
            col, _ = nvdr.interpolate(self.attr, rast, self.tri)
            y = nvdr.antialias(col, rast, pos, self.tri)
            loss = l2_loss(y, self.target[v:v + 1])
            loss.backward()
            total = total + loss.detach()
        return total

    def zero_grad(self):
        self.verts.grad = None
        self.attr.grad = None


def time_region(fn, steps, warmup):
    """Wall-clock mean of ``steps`` calls after ``warmup`` untimed ones, synchronised on both sides."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def time_steps(fn, steps, warmup):
    """``warmup`` untimed calls, then exactly ``steps`` timed ones between two
    device synchronisations, each bracketed by HIP events on the current
    stream.  Returns (median event time, wall-clock mean) in seconds."""
    for _ in range(warmup):
        fn()
    st = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(st)
        fn()
        e1.record(st)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return float(np.median([e0.elapsed_time(e1) for e0, e1 in evs])) * 1e-3, wall


def time_events(fn, iters):
    """Average duration of fn's GPU work on the current stream, HIP events."""
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(iters):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / iters * 1e-3


def per_op_times(S, iters=10):
    """Time each operator alone (forward and backward launches) over the 64-view batch."""
    from diffrp_amd import _lib
    from diffrp_amd._lib import ptr, stream_of
    L = _lib.lib()
    with torch.no_grad():
        pos = clip_space(S.verts, S.VP).contiguous()
        rast, _ = nvdr.rasterize(S.ctx, pos, S.tri, (S.H, S.W), grad_db=False)
        col, _ = nvdr.interpolate(S.attr, rast, S.tri)
    n, H, W, C, V, F = S.n, S.H, S.W, S.C, S.V, S.F
    boxes = rast._drp_boxes[0]
    (tw, th), cap = S.ctx.config(n, H, W, F)
    rws = L.drp_rasterize_workspace_size(n, H, W, n * F, tw, th, cap)
    rws_t = torch.empty(rws, dtype=torch.uint8, device=pos.device)
    rout = torch.empty_like(rast)
    boxes2 = torch.empty_like(boxes)
    attr = S.attr.detach()
    out = torch.empty_like(col)
    gout = torch.randn_like(col)
    gattr = torch.zeros_like(attr)
    grast = torch.empty_like(rast)
    gpos = torch.zeros_like(pos)
    gcol = torch.empty_like(col)
    aws = L.drp_antialias_workspace_size(n, H, W, V, F, 1)
    aws_t = torch.empty(aws, dtype=torch.uint8, device=pos.device)
    opp = S.topo.opp
    off, inc = nvdr.incidence(S.tri, V)
    vpg = nvdr._views_per_group(n, F)
    iws = L.drp_interpolate_bwd_workspace_size(n, H, W, F, C, 0, vpg)
    iws_t = torch.empty(iws, dtype=torch.uint8, device=pos.device)
    pws = L.drp_rasterize_bwd_workspace_size(n, H, W, F)
    pws_t = torch.empty(pws, dtype=torch.uint8, device=pos.device)
    fws = L.drp_interpolate_rasterize_bwd_workspace_size(n, H, W, F, C, vpg)
    fws_t = torch.empty(fws, dtype=torch.uint8, device=pos.device)
    s = stream_of(pos)
    fns = {
        'rasterize_fwd': lambda: L.drp_rasterize_fwd(s, ptr(pos), n, V, 1, ptr(S.tri), F, None, H, W, None, ptr(rout),
                                                     None, ptr(rws_t), rws, tw, th, cap, n * F, F, ptr(boxes2)),
        'interpolate_fwd': lambda: L.drp_interpolate_fwd(s, ptr(attr), n, V, 0, C, ptr(rast), H, W, ptr(S.tri), F,
                                                         None, None, 0, ptr(out), None),
        'antialias_fwd': lambda: L.drp_antialias_fwd(s, ptr(col), n, H, W, C, ptr(rast), ptr(pos), V, 1, ptr(S.tri),
                                                     F, ptr(opp), ptr(out), ptr(aws_t), aws),
        'antialias_bwd': lambda: L.drp_antialias_bwd(s, ptr(col), n, H, W, C, ptr(rast), ptr(pos), V, 1, ptr(S.tri),
                                                     F, ptr(opp), ptr(gout), ptr(gcol), ptr(gpos), 1.0, 1,
                                                     ptr(aws_t), aws),
        'interpolate_bwd': lambda: L.drp_interpolate_bwd(s, ptr(attr), n, V, 0, C, ptr(rast), H, W, ptr(S.tri), F,
                                                         None, None, 0, ptr(gout), None, ptr(gattr), ptr(grast),
                                                         None, ptr(boxes), vpg, ptr(off), ptr(inc), ptr(iws_t), iws),
        'rasterize_bwd': lambda: L.drp_rasterize_bwd(s, ptr(pos), n, V, 1, ptr(S.tri), F, None, H, W, ptr(rast),
                                                     ptr(grast), None, ptr(gpos), ptr(boxes), ptr(off), ptr(inc),
                                                     ptr(pws_t), pws),
        # what the step runs for the two backward passes above: one pass over the pixels (as nvdr calls
        # it, without writing grad_rast)
        'interpolate_rasterize_bwd': lambda: L.drp_interpolate_rasterize_bwd(
            s, ptr(attr), n, V, C, ptr(rast), H, W, ptr(S.tri), F, ptr(pos), 1, ptr(gout), ptr(gattr), None,
            ptr(gpos), ptr(boxes), vpg, ptr(off), ptr(inc), ptr(fws_t), fws),
    }
    res = {}
    for k, f in fns.items():
        t = time_events(f, iters)
        res[k] = {'ms': t * 1e3, 'bytes': op_bytes(k, n, H * W, V, F, C)}
        res[k]['gbs'] = res[k]['bytes'] / t / 1e9
    return res


# the operator calls of the captured step (interpolate and rasterize backward run as one pass there)
STEP_OPS = ('rasterize_fwd', 'interpolate_fwd', 'antialias_fwd', 'antialias_bwd', 'interpolate_rasterize_bwd')


# kernels each operator call launches (the unit of the per-operator timings)
OP_KERNELS = {
    'rasterize_fwd': ['k_zero_i32', 'k_setup_bin', 'k_fine', 'k_fine_split'],
    'interpolate_fwd': ['k_interp_fwd'],
    'antialias_fwd': ['k_aa_project', 'k_aa_orient', 'k_aa_silhouette', 'k_aa_pairs', 'k_aa_gather'],
    'antialias_bwd': ['k_aa_pairs_bwd', 'k_aa_gather_bwd'],
    'interpolate_bwd': ['k_tile_reduce<1,C>', 'k_tile_gather<1,C>', 'k_gather_attr'],
    'rasterize_bwd': ['k_tile_reduce<0>', 'k_tile_gather<0>', 'k_gather_pos'],
    'interpolate_rasterize_bwd': ['k_corner_slots_clear', 'k_corner_slots', 'k_tile_reduce<2,C>',
                                  'k_tile_gather<2,C>', 'k_gather_attr', 'k_gather_pos'],
}

TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles', 'traffic_64view_512.txt')


def sources_digest():
    """SHA-256 over the HIP sources and the C-ABI header (file names and contents):
    the build a counter summary was measured on."""
    root = os.path.dirname(os.path.abspath(__file__))
    csrc = os.path.join(root, 'diffrp_amd', 'csrc')
    h = hashlib.sha256()
    for path in sorted(os.path.join(csrc, f) for f in os.listdir(csrc)) + [os.path.join(root, 'include', 'drp.h')]:
        h.update(os.path.basename(path).encode())
        with open(path, 'rb') as f:
            h.update(f.read())
    return h.hexdigest()


def measured_traffic(op, S, subdiv):
    """HBM bytes per call of ``op`` from the committed counter summary (FETCH_SIZE and
    WRITE_SIZE passes, calibrated on a 256 MiB copy), when it covers this workload and
    was measured on these sources (else None: a summary of other kernels says nothing)."""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    want = f'workload views={S.n} res={S.H}x{S.W} subdiv={subdiv} channels={S.C}'
    if want not in lines or f'sources sha256={sources_digest()}' not in lines:
        return None
    for ln in lines:
        f = ln.split()
        if len(f) == 4 and f[0] == 'op' and f[1] == op and f[2] == 'bytes':
            return int(f[3])
    return None


def host_threads():
    """Host threads this process may use: OMP_NUM_THREADS when the environment
    sets it (the CPU share of a shared host), else every logical CPU."""
    n = os.cpu_count() or 1
    try:
        return max(1, min(n, int(os.environ.get('OMP_NUM_THREADS', n))))
    except ValueError:
        return n


def cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def _cpu_step(oracle, pos, tri, opp, attr, target, H, W, backward):
    """One view through the oracle: rasterize, interpolate, antialias, L2 (and its backward)."""
    rast = oracle.rasterize(pos, tri, (H, W))
    col = oracle.interpolate(attr, rast, tri)
    y = oracle.antialias(col, rast, pos, tri, opp)
    if not backward:
        return y
    gy = 2.0 * (y - target)
    gcol, gpos_aa = oracle.antialias_bwd(col, rast, pos, tri, gy, opp)
    gattr, grast = oracle.interpolate_bwd(attr, rast, tri, gcol)
    gpos = oracle.rasterize_bwd(pos, tri, rast, grast) + gpos_aa
    return gpos, gattr


def cpu_baseline(S, views):
    """The CPU oracle (the C restatement, OpenMP over the host threads this
    process may use, plus the torch-CPU clip transform and loss) on the same workload: ``views`` of the
    64 views fwd+bwd (the headline), and the single-view configurations C1
    (README-size mesh at 256^2, forward) and C2 (the s=6 mesh at 1024^2,
    fwd+bwd)."""
    import oracle
    oracle.build()
    cores = host_threads()
    oracle.set_threads(cores)
    torch.set_num_threads(cores)
    tri = S.tri_cpu.numpy()
    opp = oracle.topology(tri)
    attr = S.attr_cpu.numpy()
    t0 = time.perf_counter()
    for v in range(views):
        pos = clip_space(S.verts_cpu, S.VP_cpu[v]).numpy()
        gpos, gattr = _cpu_step(oracle, pos, tri, opp, attr, S.target_cpu[v:v + 1].numpy(), S.H, S.W, True)
        _ = float(gpos.sum()) + float(gattr.sum())
    dt = time.perf_counter() - t0
    # C1: forward of the README-size mesh at 256^2, 3 channels
    S1 = Scene('cpu', 1, 256, 3, 3, sigma=0.0, cpu_only=True)
    p1 = clip_space(S1.verts_cpu, S1.VP_cpu[0]).numpy()
    t1 = time.perf_counter()
    for _ in range(5):
        _cpu_step(oracle, p1, S1.tri_cpu.numpy(), None, S1.attr_cpu.numpy(), None, 256, 256, False)
    t1 = (time.perf_counter() - t1) / 5
    # C2: the s=6 mesh at 1024^2, 8 channels, fwd+bwd
    S2 = Scene('cpu', 1, 1024, 6, 8, cpu_only=True)
    p2 = clip_space(S2.verts_cpu, S2.VP_cpu[0]).numpy()
    t2 = time.perf_counter()
    _cpu_step(oracle, p2, S2.tri_cpu.numpy(), opp, S2.attr_cpu.numpy(), S2.target_cpu.numpy(), 1024, 1024, True)
    t2 = time.perf_counter() - t2
    return {
        'value': views * S.H * S.W / dt / 1e6,
        'unit': 'Mpixel/s',
        'cores': cores,
        'kind': 'port',
        'sample': f'{views} of the 64 views at {S.H}x{S.W}, fwd+bwd, C oracle with OpenMP on {cores} threads '
                  f'+ torch-CPU ({dt:.1f} s; {cpu_model()})',
        'cpu_model': cpu_model(),
        'C1_fwd_ms': round(t1 * 1e3, 2),
        'C2_fwd_bwd_mpix_s': round(1024 * 1024 / t2 / 1e6, 3),
    }


def other_configs(dev):
    """Single-view configurations (reported beside the headline, not the bench value):
    C1 = README-size mesh (icosphere s=3, r=0.8) at 256^2, forward only, 3 channels;
    C2 = the noisy s=6 mesh at 1024^2, forward + backward, 8 channels.  Eager calls
    (the diffrp usage shape) and, for C2, the HIP-graph replay of the same step."""
    res = {}
    S1 = Scene(dev, 1, 256, 3, 3, sigma=0.0)
    with torch.no_grad():
        def fwd():
            pos = clip_space(S1.verts, S1.VP)
            rast, _ = nvdr.rasterize(S1.ctx, pos, S1.tri, (256, 256), grad_db=False)
            col, _ = nvdr.interpolate(S1.attr, rast, S1.tri)
            return nvdr.antialias(col, rast, pos, S1.tri, topology_hash=S1.topo)
        t1 = time_region(fwd, 20, 3)
    res['C1'] = {'workload': 'icosphere s=3 (642 verts, 1280 faces), 256x256, 3 channels, forward, eager',
                 'ms': round(t1 * 1e3, 4), 'mpix_s': round(256 * 256 / t1 / 1e6, 1)}
    S2 = Scene(dev, 1, 1024, 6, 8)

    def step():
        S2.zero_grad()
        S2.step_batched()
    t2 = time_region(step, 10, 3)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            step()
    torch.cuda.current_stream().wait_stream(side)
    S2.zero_grad()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        S2.step_batched()
    t2g = time_region(g.replay, 10, 3)
    res['C2'] = {'workload': 'noisy icosphere s=6 (40962 verts, 81920 faces), 1024x1024, 8 channels, fwd+bwd',
                 'eager_ms': round(t2 * 1e3, 4), 'graph_ms': round(t2g * 1e3, 4),
                 'mpix_s': round(1024 * 1024 / t2g / 1e6, 1)}
    del S2, g
    progress('C1, C2 timed')
    res['C4'] = uneven_config(dev)
    progress('C4 timed')
    res['C5'] = layered_config(dev)
    progress('C5 timed')
    return res


def capture(step):
    """``step`` captured in a HIP graph after two warm-up runs on a side stream."""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            step()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        step()
    return g


def uneven_draw(seed=3, n=32, C=4, sigma=0.02):
    """The uneven batch: ``n`` icospheres of subdivision uniform in 3..8 with
    N(0, sigma) vertex noise, each at a resolution drawn from {256, 512, 1024,
    2048}^2, seen from Fibonacci directions at distance 3.2 (draws from ``seed``)."""
    rng = np.random.default_rng(seed)
    subdivs = rng.integers(3, 9, n)
    sizes = rng.choice([256, 512, 1024, 2048], n)
    dirs = fibonacci_directions(n)
    pos, tri, res = [], [], []
    for i, (sd, sz) in enumerate(zip(subdivs, sizes)):
        v, f = noisy_icosphere(int(sd), 0.8, sigma, seed=int(rng.integers(1 << 30)))
        cam = PerspectiveCamera.looking_at_origin_from(np.asarray(dirs[i]) * 3.2, h=int(sz), w=int(sz))
        pos.append(clip_space(v, cam.P('cpu') @ cam.V('cpu')).contiguous())
        tri.append(f.int().contiguous())
        res.append((int(sz), int(sz)))
    return pos, tri, res


def uneven_config(dev, steps=10):
    """C4: the 32-mesh uneven batch (seed 3, C=4), tuned, fwd+bwd of rasterize,
    interpolate, antialias and an L2 loss as one captured HIP graph of batched
    launches, against a loop of per-item eager steps."""
    from diffrp_amd import batch
    pos, tri, res = uneven_draw()
    ctx = nvdr.RasterizeCudaContext(dev)
    tune = batch.tune_batch(ctx, [p.to(dev) for p in pos], [t.to(dev) for t in tri], res)
    ub = batch.UnevenBatch(ctx, [t.to(dev) for t in tri], [p.shape[0] for p in pos], res)
    g = torch.Generator().manual_seed(6)
    P = torch.cat(pos).to(dev).contiguous().requires_grad_(True)
    A = torch.rand(P.shape[0], 4, generator=g).to(dev).requires_grad_(True)
    T = torch.rand(ub.n_pix, 4, generator=g).to(dev)

    def step():
        P.grad = None
        A.grad = None
        rast, boxes = ub.rasterize(P)
        out = ub.antialias(ub.interpolate(A, rast, boxes), rast, P)
        l2_loss(out, T).backward()
    graph = capture(step)
    t, wall = time_steps(graph.replay, steps, 2)
    items = [(p.shape[0], t_.shape[0], h * w) for p, t_, (h, w) in zip(pos, tri, res)]
    npx = sum(n for _, _, n in items)
    nf = sum(f for _, f, _ in items)
    alg = sum(240 * n + 128 * v + 72 * f for v, f, n in items)
    # the per-item loop: one eager drop-in step per mesh
    pcs = [p.to(dev).requires_grad_(True) for p in pos]
    tcs = [t_.to(dev) for t_ in tri]
    acs = [A.detach()[int(ub.plan.vbase[i]):int(ub.plan.vbase[i]) + pos[i].shape[0]].clone().requires_grad_(True)
           for i in range(len(pos))]
    tgs = ub.split(T)

    def loop():
        for p, t_, a, tg, r in zip(pcs, tcs, acs, tgs, res):
            rast, _ = nvdr.rasterize(ctx, p, t_, r, grad_db=False)
            y = nvdr.antialias(nvdr.interpolate(a, rast, t_)[0], rast, p, t_)
            l2_loss(y, tg).backward()
    t_loop = time_region(loop, 3, 1)
    return {'workload': f'{len(pos)} noisy icospheres s=3..8 ({nf} faces), 256^2..2048^2 ({npx} pixels), '
                        f'4 channels, fwd+bwd, batched launches in one HIP graph',
            'ms': round(t * 1e3, 4), 'wall_ms': round(wall * 1e3, 4), 'mpix_s': round(npx / t / 1e6, 1),
            'mtri_s': round(nf / t / 1e6, 1), 'algorithmic_bytes': alg,
            'roofline_frac': round(alg / t / 1e9 / HBM_PEAK_GBS, 4),
            'per_item_loop_ms': round(t_loop * 1e3, 3), 'batched_vs_loop': round(t_loop / t, 2),
            'autotune': {'tile': list(tune['tile']), 'candidates_ms': tune['candidates_ms']}}


def layered_config(dev):
    """C5: icosphere s=9 (5,242,880 faces, N(0, 0.005) noise), 2048^2, 4 peeled
    layers composited, 16 channels, antialias, fwd+bwd of an L2 loss: with each
    layer recomputed in the backward pass, and with the layers' buffers kept."""
    from diffrp_amd import layered
    H = W = 2048
    layers, C = 4, 16
    v, f = noisy_icosphere(9, 0.8, 0.005, seed=0)
    cam = PerspectiveCamera.looking_at_origin_from([0.6, 0.5, 2.0], h=H, w=W)
    g = torch.Generator().manual_seed(9)
    attr = torch.rand(v.shape[0], C, generator=g).to(dev).requires_grad_(True)
    target = torch.rand(1, H, W, 3, generator=g).to(dev)
    verts = v.to(dev).requires_grad_(True)
    tri = f.int().to(dev)
    VP = (cam.P('cpu') @ cam.V('cpu')).to(dev)
    ctx = nvdr.RasterizeCudaContext(dev)
    light = torch.tensor([0.3, 0.5, 0.8], device=dev)

    def shade(gb, rast):
        n = torch.nn.functional.normalize(gb[..., 6:9] - 0.5, dim=-1, eps=1e-6)
        lit = 0.5 + 0.5 * (n * light).sum(-1, keepdim=True)
        rgb = gb[..., 9:12] * lit + 0.05 * (gb[..., 0:2].sum(-1, keepdim=True) + gb[..., 2:6].sum(-1, keepdim=True)
                                            + gb[..., 13:16].sum(-1, keepdim=True))
        return torch.cat([rgb, 0.3 + 0.6 * gb[..., 12:13]], -1)

    out = {}
    for mode, recompute in (('recompute', True), ('stored', False)):
        def step():
            verts.grad = None
            attr.grad = None
            pos = clip_space(verts, VP)
            img = layered.render_layers(ctx, pos, tri, attr, (H, W), layers, shade, recompute=recompute)
            l2_loss(img, target).backward()
        step()
        torch.cuda.synchronize()
        torch.cuda.reset_peak_memory_stats(dev)
        base = torch.cuda.memory_allocated(dev)
        t, wall = time_steps(step, 3, 1)
        peak = torch.cuda.max_memory_allocated(dev) - base
        out[mode] = {'ms': round(t * 1e3, 3), 'mpix_s': round(layers * H * W / t / 1e6, 1),
                     'peak_hbm_gib': round(peak / 2 ** 30, 3)}
    N, V, F = H * W, v.shape[0], f.shape[0]
    alg = layers * (48 * N + 48 * V + 24 * F) + layers * ((48 + 8 * C) * N + 8 * C * V + 24 * F) + \
        ((32 + 20 * 3) * N + 48 * V + 24 * F) + (layers - 1) * 16 * N
    out['workload'] = (f'icosphere s=9 ({F} faces), 2048x2048, {layers} peeled layers, {C} channels, '
                       f'antialias, fwd+bwd, eager')
    out['algorithmic_bytes'] = alg
    out['roofline_frac_recompute'] = round(alg / (out['recompute']['ms'] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    return out


def progress(msg):
    """One line per finished stage on stderr (stdout carries only the result line)."""
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def main():
    faulthandler.enable()
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--views', type=int, default=64)
    ap.add_argument('--res', type=int, default=512)
    ap.add_argument('--subdiv', type=int, default=6)
    ap.add_argument('--channels', type=int, default=8)
    ap.add_argument('--plain-steps', type=int, default=3)
    ap.add_argument('--cpu-views', type=int, default=64)
    ap.add_argument('--no-plain', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-ops', action='store_true')
    ap.add_argument('--no-tune', action='store_true')
    ap.add_argument('--no-configs', action='store_true', help='skip the single-view C1 / C2 timings')
    ap.add_argument('--pmc-calibration', action='store_true',
                    help='first copy 256 MiB device-to-device three times: the reference for counter scaling')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU: the rasterisation path has no multi-GPU mode in this build '
              '(it would shard by view); run with --gpus 1', file=sys.stderr)
        sys.exit(2)
    dev = torch.device('cuda:0')
    if args.pmc_calibration:
        a = torch.ones(64 * 2**20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    S = Scene(dev, args.views, args.res, args.subdiv, args.channels)
    tune = None
    if not args.no_tune:
        with torch.no_grad():
            pos_all = clip_space(S.verts, S.VP)
            tune = autotune.tune_rasterize(S.ctx, pos_all, S.tri, (S.H, S.W))
            autotune.tune_rasterize(S.ctx, pos_all[0], S.tri, (S.H, S.W))  # the plain loop's single view
        progress(f'autotuned: {tune["best"]["tile"]}, capacity {tune["best"]["cap"]}')

    # eager reference step, then capture the whole fwd+bwd in a HIP graph
    S.zero_grad()
    S.step_batched()
    torch.cuda.synchronize()
    ref_gv, ref_ga = S.verts.grad.clone(), S.attr.grad.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            S.zero_grad()
            S.step_batched()
    torch.cuda.current_stream().wait_stream(side)
    S.zero_grad()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        S.step_batched()
    # replays must reproduce the eager gradients (checked on several replays)
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        for a, b, name in ((S.verts.grad, ref_gv, 'verts'), (S.attr.grad, ref_ga, 'attr')):
            err = (a - b).abs().max().item() / max(b.abs().max().item(), 1e-30)
            if err > 1e-4:
                raise RuntimeError(f'graph replay gradient mismatch on {name}: {err:.3e}')

    progress('captured step replays match the eager gradients')
    t_batched, t_wall = time_steps(graph.replay, args.steps, args.warmup)
    progress(f'batched: {t_batched * 1e3:.4f} ms per step')
    pix = S.n * S.H * S.W
    value = pix / t_batched / 1e6

    plain = None
    if not args.no_plain:
        def plain_step():
            S.zero_grad()
            S.step_plain()
        t_plain = time_region(plain_step, args.plain_steps, 1)
        plain = {'mpix_s': pix / t_plain / 1e6, 'ms_per_step': t_plain * 1e3, 'steps': args.plain_steps}

    alg = algorithmic_bytes(S.n, S.H, S.W, S.V, S.F, S.C)
    progress('plain loop timed')
    ops = None if args.no_ops else per_op_times(S)
    progress('operators timed')
    roof = None
    if ops:
        dom = max(STEP_OPS, key=lambda k: ops[k]['ms'])
        roof = {'bound': 'hbm', 'kernel': dom, 'kernels': OP_KERNELS[dom], 'achieved': round(ops[dom]['gbs'], 1),
                'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': round(ops[dom]['gbs'] / HBM_PEAK_GBS, 4),
                'traffic': measured_traffic(dom, S, args.subdiv),
                'bytes_per_launch': ops[dom]['bytes'], 'ms_per_launch': round(ops[dom]['ms'], 4)}
    cpu = None if args.no_cpu else cpu_baseline(S, args.cpu_views)
    progress('cpu baseline timed')
    configs = None if args.no_configs else other_configs(dev)

    out = {
        'metric': METRIC,
        'value': round(value, 2),
        'unit': 'Mpixel/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(t_batched * 1e3, 4),
        'timing': f'median of {args.steps} HIP-event step times (graph replays); wall-clock mean '
                  f'{t_wall * 1e3:.4f} ms',
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': f'{S.n}-view fwd+bwd, noisy icosphere s={args.subdiv} ({S.V} verts, {S.F} faces), '
                               f'{S.H}x{S.W}, {S.C} attribute channels, batched HIP-graph replay',
                   'views': S.n, 'resolution': [S.H, S.W], 'faces': S.F, 'verts': S.V, 'channels': S.C},
        'mtri_s': round(S.n * S.F / t_batched / 1e6, 2),
        'step_roofline': {'algorithmic_bytes': alg, 'achieved_gbs': round(alg / t_batched / 1e9, 1),
                          'frac': round(alg / t_batched / 1e9 / HBM_PEAK_GBS, 4)},
        'plain': plain,
        'batched_vs_plain': round(value / plain['mpix_s'], 2) if plain else None,
        'ops': {k: {'ms': round(v['ms'], 4), 'gbs': round(v['gbs'], 1)} for k, v in ops.items()} if ops else None,
        'roofline': roof,
        'cpu_baseline': cpu,
        'other_configs': configs,
        'autotune': None if tune is None else {
            'tile': list(tune['best']['tile']), 'bin_capacity': tune['best']['cap'],
            'candidates_ms': {k: round(v['ms'], 4) for k, v in tune['all'].items()}},
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()
