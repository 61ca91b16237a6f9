"""Benchmark: one 512^3 complex-to-complex forward 3D FFT per step on one MI355X.

Workload (the headline configuration): a 512^3 complex64 Gaussian random
field with P(k) ~ k^-3 (seed 2) plus 1e-3 uniform noise, block-decomposed over
16 virtual ranks with dims [1,4,4] (local blocks 512x128x128), transformed in
place by the fused FP32 plan: three Stockham passes (z, x, y) over the block
layout.  Inputs are resident in HBM before the timed region starts.

Prints one JSON line with ms per transform, GFLOP/s (5 N log2 N), the HBM
roofline of the dominant pass, the rel-L2 / max-abs error against the FP64
path, and a CPU baseline (the C restatement of the reference dataflow).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3d-ffts-mpi_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_grf(a2afft, torch, shape, seed, slope=-3.0, noise=1e-3):
    """The GRF input on the device (complex128 global C-order grid): the
    package's generator, tested against the host recipe at 32^3 and 64^3."""
    return a2afft.inputs.gaussian_random_field_device(shape, seed, slope, noise)


def host_cores():
    """(cores the process may run on, CPU model): the affinity mask, capped by
    the cgroup CPU quota when one is set (cpu.max), so every thread counted
    gets a core of its own."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(period)))
    except Exception:
        pass
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except Exception:
        pass
    return (min(aff, quota) if quota else aff), aff, quota, model


def library_gpu_fftn(a2afft, torch, Ng, precision, rounds=3, iters=5):
    """The same grid size through the GPU's vendor FFT library (rocFFT, via
    torch.fft.fftn / ifftn, out of place) and through this library's plain-grid
    plan (natural C order, no virtual ranks, in place), on one natural-order
    Ng^3 array: ms per transform, timed as forward + inverse pairs (so values
    stay finite) and halved, median of `rounds` rounds of `iters` pairs."""
    dt = torch.complex64 if precision == "fp32" else torch.complex128
    x = torch.empty((Ng, Ng, Ng), dtype=dt, device="cuda")
    a2afft.fill_random(x, 7, "uniform")
    y = torch.empty_like(x)
    plan = a2afft.GridDfft((Ng, Ng, Ng), precision)

    def pair_ms(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        res = []
        for _ in range(rounds):
            e0.record()
            for _ in range(iters):
                fn()
            e1.record()
            torch.cuda.synchronize()
            res.append(e0.elapsed_time(e1) / iters / 2)
        return sorted(res)[len(res) // 2]

    def roc():
        torch.fft.fftn(x, out=y)
        torch.fft.ifftn(y, out=x)

    def ours():
        plan.forward(x)
        plan.inverse(x)

    roc_ms = pair_ms(roc)
    ours_ms = pair_ms(ours)
    plan.destroy()
    del x, y
    return {"what": f"one natural-order {Ng}^3 {'complex64' if precision == 'fp32' else 'complex128'} grid, "
                    "ms per transform (forward + inverse pairs, halved, median of 3)",
            "rocfft_fftn_ms": round(roc_ms, 4), "a2afft_grid_plan_ms": round(ours_ms, 4)}


def cpu_baseline(Ng, P, dims, threads, seed, samples=3):
    """The oracle (C restatement of the reference stage order, FP64) on full
    transforms of the same grid shape, OpenMP over `threads` cores: the median
    of `samples` timed forwards (ms), every sample's time, and the threads used."""
    import ctypes
    import numpy as np
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    O.set_threads(threads)
    rc, d = O.decompose(Ng, P, dims)
    assert rc == 0
    x0 = O.fill_uniform(seed, Ng ** 3).reshape(P, -1)
    x = np.empty_like(x0)
    S = np.zeros_like(x0)
    times = []
    for _ in range(samples):
        x[...] = x0
        t0 = time.perf_counter()
        st = O.lib().orc_forward(ctypes.byref(d), x.ctypes.data_as(ctypes.c_void_p), S.ctypes.data_as(ctypes.c_void_p))
        times.append((time.perf_counter() - t0) * 1e3)
        assert st == 0
    assert np.isfinite(x).all()
    return float(np.median(times)), times, O.get_threads()


def library_cpu_fftn(Ng, threads, samples=3):
    """scipy's pocketfft fftn of an Ng^3 complex128 grid on `threads` workers:
    the median of `samples` forwards (ms), or None without scipy."""
    try:
        import numpy as np
        import scipy.fft as sfft
    except Exception:
        return None
    a = np.random.default_rng(0).standard_normal((Ng, Ng, Ng)).astype(np.complex128)
    times = []
    for _ in range(samples):
        t0 = time.perf_counter()
        y = sfft.fftn(a, workers=threads)
        times.append((time.perf_counter() - t0) * 1e3)
        del y
    return float(np.median(times))


def _traffic(key):
    """Measured HBM bytes per launch (profiles/traffic.json), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f).get(key)
    except Exception:
        return None


class Rotation:
    """`count` device copies of `src`, refilled outside every timed region, so
    each timed transform runs on the input itself: a forward multiplies values
    by up to N, and repeated in-place forwards of one buffer overflow float32
    within a few calls."""

    def __init__(self, torch, src, count):
        self.torch, self.src = torch, src
        self.bufs = [torch.empty_like(src) for _ in range(max(1, count))]
        self.refill()

    def refill(self):
        for b in self.bufs:
            b.copy_(self.src)
        self.torch.cuda.synchronize()

    def __getitem__(self, i):
        return self.bufs[i % len(self.bufs)]

    def finite(self):
        t = self.torch
        return all(bool(t.isfinite(t.view_as_real(b)).all()) for b in self.bufs)


def rotation_count(src, steps, budget_bytes=24 << 30):
    """Buffers for `steps` timed calls within a memory budget (at least 1)."""
    per = src.numel() * src.element_size()
    return max(1, min(steps, budget_bytes // max(1, per)))


def _events_ms(torch, fn, rot, steps, warmup=1):
    """Average ms of fn(buf) over `steps` calls, HIP events on the launch
    stream; warm-up calls first, then every buffer of `rot` refilled from the
    input (untimed), then the timed calls on buffers i = 0, 1, ... in turn."""
    for i in range(max(1, warmup)):
        fn(rot[i])
    rot.refill()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for i in range(steps):
        fn(rot[i])
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def run_c3b(args, torch, a2afft):
    """One rank's share of a 2048^3 grid over 64 ranks, dims [1,8,8]: the
    block [2048][256][256] complex64 (GRF, seed 2, on that shape) through the
    x stage.  Staged as the reference: pack (x_a2a_to_x_pencils) -> batched
    2048-point FFT -> unpack (x_pencils_to_x_a2a), on the pencil layout the
    exchange delivers; fused: the FFT kernel alone on those pencils (with
    dims[0] = 1 pack and unpack are the identity), and the strided x pass
    straight on the block."""
    shape, P, dims, seed = (2048, 256, 256), 64, (1, 8, 8), 2
    lgs = shape
    nloc = shape[0] * shape[1] * shape[2]
    log("[bench] generating 2048x256x256 GRF share")
    g = make_grf(a2afft, torch, shape, seed)
    block = g.to(torch.complex64)
    del g
    # pencils as the exchange delivers them: fast_z_to_x ([lx][ly][lz] -> [lz][ly][lx])
    pencils = torch.empty_like(block)
    a2afft.launch_remap(6, block, pencils, P, nloc, lgs, dims, nbuffers=1)
    torch.cuda.synchronize()
    # FP64 references
    ref_p = pencils.to(torch.complex128)
    a2afft.forward_1d_fft(ref_p, 2048)
    work = pencils.clone()
    a2afft.forward_1d_fft(work, 2048)
    rel_p, max_p = a2afft.error_metrics(work, ref_p)
    # split-BF16 MFMA: the pencils are the z pass of a 256x256x2048 grid plan
    plan_s = a2afft.GridDfft((256, 256, 2048), "split_bf16")
    ws = pencils.clone()
    a2afft.execute_pass(plan_s, ws, 0, -1)
    rel_s, max_s = a2afft.error_metrics(ws, ref_p)
    plan_h = a2afft.GridDfft((256, 256, 2048), "split_fp16")
    ws.copy_(pencils)
    a2afft.execute_pass(plan_h, ws, 0, -1)
    rel_h, max_h = a2afft.error_metrics(ws, ref_p)
    del ref_p, ws
    plan = a2afft.GridDfft(shape, "fp32")
    ref_b = block.to(torch.complex128)
    plan64 = a2afft.GridDfft(shape, "fp64")
    a2afft.execute_pass(plan64, ref_b, 1, -1)
    wb = block.clone()
    a2afft.execute_pass(plan, wb, 1, -1)
    rel_b, max_b = a2afft.error_metrics(wb, ref_b)
    del ref_b, plan64
    if args.autotune:
        plan.autotune(wb, iters=3)
    log(plan.describe())
    tmp = torch.empty_like(pencils)

    def staged(buf):
        a2afft.launch_remap(2, buf, tmp, P, nloc, lgs, dims, nbuffers=1)
        a2afft.forward_1d_fft(tmp, 2048)
        a2afft.launch_remap(3, tmp, work, P, nloc, lgs, dims, nbuffers=1)

    nrot = rotation_count(block, args.steps, 16 << 30)
    rot = Rotation(torch, pencils, nrot)
    t_staged = _events_ms(torch, staged, rot, args.steps, args.warmup)
    t_pencil = _events_ms(torch, lambda buf: a2afft.forward_1d_fft(buf, 2048), rot, args.steps, args.warmup)
    t_split = _events_ms(torch, lambda buf: a2afft.execute_pass(plan_s, buf, 0, -1), rot, args.steps, args.warmup)
    t_split16 = _events_ms(torch, lambda buf: a2afft.execute_pass(plan_h, buf, 0, -1), rot, args.steps, args.warmup)
    finite = rot.finite()
    del rot
    rot = Rotation(torch, block, nrot)
    t_block = _events_ms(torch, lambda buf: a2afft.execute_pass(plan, buf, 1, -1), rot, args.steps, args.warmup)
    finite = finite and rot.finite()
    del rot
    best = min(t_pencil, t_block)
    bytes_fft = 2 * nloc * 8
    # the same access patterns without arithmetic: the contiguous 2048-point
    # pass (the z pass of a 256x256x2048 grid has the pencils' layout) and the
    # block's strided x pass
    pc_pencil = a2afft.GridDfft((256, 256, 2048), "fp32").pattern_copy(work, 0, 10)
    pc_block = plan.pattern_copy(wb, 1, 10)
    ach = bytes_fft / (t_pencil * 1e-3) / 1e9
    out = {
        "metric": "ms per x stage of one 2048x256x256 rank share (c2c, 2048-point lines)",
        "value": round(best, 4), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(best, 4), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "one rank's 2048x256x256 complex64 share of 2048^3 over dims [1,8,8], "
                               "GRF P(k)~k^-3 seed 2 + 1e-3 noise, x stage", "name": "c3b"},
        "ms": {"staged_pack_fft_unpack": round(t_staged, 4), "fused_pencil_fft": round(t_pencil, 4),
               "fused_block_x_pass": round(t_block, 4), "split_bf16_pencil_fft": round(t_split, 4),
               "split_fp16_pencil_fft": round(t_split16, 4)},
        "split_bf16_pencils": {"ms": round(t_split, 4), "rel_l2_vs_fp64": rel_s, "max_abs_vs_fp64": max_s,
                               "hbm_gbs": round(bytes_fft / (t_split * 1e-3) / 1e9, 1)},
        "split_fp16_pencils": {"ms": round(t_split16, 4), "rel_l2_vs_fp64": rel_h, "max_abs_vs_fp64": max_h,
                               "hbm_gbs": round(bytes_fft / (t_split16 * 1e-3) / 1e9, 1)},
        "gflops": round(5.0 * nloc * math.log2(2048) / (best * 1e-3) / 1e9, 1),
        "pattern_copy_gbs": {"pencil_fft": round(bytes_fft / (pc_pencil * 1e-3) / 1e9, 1),
                             "block_x_pass": round(bytes_fft / (pc_block * 1e-3) / 1e9, 1)},
        "vs_pattern_copy": {"pencil_fft": round(pc_pencil / t_pencil, 3), "block_x_pass": round(pc_block / t_block, 3)},
        "rel_l2_vs_fp64": {"pencils": rel_p, "block": rel_b},
        "max_abs_vs_fp64": {"pencils": max_p, "block": max_b},
        "roofline": {"bound": "hbm", "kernel": "stockham 2048 contiguous", "bytes_per_launch": bytes_fft,
                     "launch_ms": round(t_pencil, 4), "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": _traffic("c3b_pencil_2048")},
        "finite": finite,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def run_grid(args, torch, a2afft):
    """One natural-order nx x ny x nz grid (no virtual ranks) through the
    plain-grid plan, in place: uniform Philox input (seed 7).  Reports ms per
    forward (rotating copies of the input, refilled outside the timed region),
    per-pass ms, rel-L2 and max-abs against rocFFT's FP64 fftn of the same
    input (grids up to 2^28 points), and rocFFT's own forward time on the
    grid's precision (out of place, where two copies fit)."""
    shape = tuple(int(v) for v in args.shape.lower().split("x"))
    if len(shape) != 3:
        raise SystemExit("--shape NXxNYxNZ")
    prec = args.precision or "fp32"
    if prec not in ("fp32", "fp64"):
        raise SystemExit("--config grid runs fp32 or fp64")
    cdt = torch.complex64 if prec == "fp32" else torch.complex128
    esize = 8 if prec == "fp32" else 16
    n = shape[0] * shape[1] * shape[2]
    log(f"[bench] grid {shape} {prec}: {n * esize / 2**30:.2f} GiB")
    x = torch.empty(shape, dtype=cdt, device="cuda")
    a2afft.fill_random(x, 7, "uniform")
    plan = a2afft.GridDfft(shape, prec)
    if args.autotune:
        work = torch.empty_like(x)
        work.copy_(x)
        plan.autotune(work, iters=3)
        del work
    log(plan.describe())
    rel = mx = None
    if n <= (1 << 28):
        y = x.clone()
        plan.forward(y)
        ref = torch.fft.fftn(x.to(torch.complex128))
        d = y.to(torch.complex128) - ref
        rel = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(ref))
        mx = float(d.abs().max())
        del y, ref, d
    nrot = rotation_count(x, args.steps, 16 << 30)
    rot = Rotation(torch, x, nrot)
    # At most 3 timed forwards land on one copy between refills: values grow
    # by at most N per forward, so 3 stay finite in float32 even at 2^33
    # points (a 64 GiB grid gets one copy, so 3 timed forwards).
    steps = min(args.steps, 3 * nrot)
    warm = min(args.warmup, 3 * nrot)
    t_fwd = _events_ms(torch, lambda buf: plan.forward(buf), rot, steps, warm)
    t_pass = [_events_ms(torch, lambda buf, p=p: a2afft.execute_pass(plan, buf, p, -1), rot, steps, 1)
              for p in range(3)]
    finite = rot.finite()
    del rot
    t_roc = None
    free, _ = torch.cuda.mem_get_info()
    if free > 3 * n * esize:
        out_buf = torch.empty_like(x)
        torch.fft.fftn(x, out=out_buf)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            torch.fft.fftn(x, out=out_buf)
        e1.record()
        torch.cuda.synchronize()
        t_roc = e0.elapsed_time(e1) / args.steps
        del out_buf
    bytes_pass = 2 * n * esize
    dom = max(range(3), key=lambda p: t_pass[p])
    ach = bytes_pass / (t_pass[dom] * 1e-3) / 1e9
    chosen = [ln for ln in plan.describe().splitlines() if ln.startswith("chosen")]
    out = {
        "metric": f"ms per {shape[0]}x{shape[1]}x{shape[2]} c2c 3D FFT (natural order)",
        "value": round(t_fwd, 4), "unit": "ms", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": round(t_fwd, 4), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32" if prec == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": f"one natural-order {shape[0]}x{shape[1]}x{shape[2]} "
                               f"{'complex64' if prec == 'fp32' else 'complex128'} grid, uniform Philox seed 7, "
                               "plain-grid plan, forward", "name": "grid", "shape": list(shape), "precision": prec,
                   "autotune": bool(args.autotune)},
        "kernels": chosen,
        "pass_ms": {"z": round(t_pass[0], 4), "x": round(t_pass[1], 4), "y": round(t_pass[2], 4)},
        "pass_gbs": {k: round(bytes_pass / (t * 1e-3) / 1e9, 1) for k, t in zip("zxy", t_pass)},
        "hbm_gbs_whole": round(3 * bytes_pass / (t_fwd * 1e-3) / 1e9, 1),
        "roofline_frac_whole": round(3 * bytes_pass / (t_fwd * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "gflops": round(5.0 * n * math.log2(n) / (t_fwd * 1e-3) / 1e9, 1),
        "rel_l2_vs_fp64": rel, "max_abs_vs_fp64": mx, "finite": finite,
        "library_gpu": {"what": "rocFFT through torch.fft.fftn, out of place, same precision, forward",
                        "rocfft_fftn_ms": None if t_roc is None else round(t_roc, 4)},
        "roofline": {"bound": "hbm", "kernel": f"{'zxy'[dom]} pass", "bytes_per_launch": bytes_pass,
                     "launch_ms": round(t_pass[dom], 4), "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": None,
    }
    plan.destroy()
    print(json.dumps(out), flush=True)


C5_SHAPES = [(32, 32, 32), (64, 64, 64), (128, 128, 128), (256, 256, 256), (128, 256, 64), (256, 64, 512)]


def c5_shapes(seed=4, count=64):
    """64 grid shapes drawn with Philox (seed 4) from the six candidate shapes."""
    from a2afft import inputs
    u = (inputs.uniform(seed, count).real + 1.0) / 2.0
    return [C5_SHAPES[min(5, int(v * 6))] for v in u]


def run_c5(args, torch, a2afft):
    """64 independent complex64 grids of uneven sizes, forward 3D FFT of all of
    them in one persistent launch (and its hipGraph), against 192 separate
    single-grid launches; aggregate GFLOP/s and per-grid rel-L2 vs FP64."""
    shapes = c5_shapes()
    b = a2afft.Batch(shapes)
    x = b.empty()
    a2afft.fill_random(x, 4, "uniform")
    torch.cuda.synchronize()
    # per-grid error vs the FP64 single-grid plan on the same (rounded) input
    y = x.clone()
    b.forward(y)
    assert b.status() == 0
    errs = []
    for i, s in enumerate(shapes):
        lo, hi = b.offsets[i], b.offsets[i + 1]
        ref = x[lo:hi].to(torch.complex128).view(*s).contiguous()
        a2afft.GridDfft(s, "fp64").forward(ref)
        errs.append(a2afft.error_metrics(y[lo:hi], ref.reshape(-1))[0])
        del ref
    # the library's autotune: workgroup counts, queue layouts, takes and phase
    # orders (every setting computes the same bits)
    b.autotune(y, iters=3)
    for line in b.describe().splitlines():
        log(f"[bench] batch {line}")
    sched = b.schedule()
    del y
    rot = Rotation(torch, x, rotation_count(x, args.steps))
    t_eager = _events_ms(torch, lambda buf: b.forward(buf), rot, args.steps, args.warmup)
    assert b.status() == 0
    # the batch keeps one captured graph per buffer and direction
    t_graph = _events_ms(torch, lambda buf: b.graph_launch(buf, -1), rot, args.steps, args.warmup)
    assert b.status() == 0
    plans = {s: a2afft.GridDfft(s, "fp32") for s in set(shapes)}

    def separate(buf):
        for v, s in zip(b.views(buf), shapes):
            plans[s].forward(v)

    t_sep = _events_ms(torch, separate, rot, args.steps, args.warmup)
    finite = rot.finite()
    del rot
    # the same 64 grids, each decomposed over 8 virtual ranks, dims [2,2,2]:
    # the persistent launch addresses the rank blocks directly (the remap
    # folded into the passes); the input is the same field in block layout
    dims = [(2, 2, 2)] * len(shapes)
    br = a2afft.Batch(shapes, dims=dims)
    xr = torch.empty_like(x)
    for i, s in enumerate(shapes):
        lo, hi = b.offsets[i], b.offsets[i + 1]
        xr[lo:hi] = a2afft.inputs.global_to_blocks(x[lo:hi].view(*s), dims[i]).reshape(-1)
    yr = xr.clone()
    br.forward(yr)
    assert br.status() == 0
    errs_r = []
    for i, s in enumerate(shapes):
        lo, hi = b.offsets[i], b.offsets[i + 1]
        ref = x[lo:hi].to(torch.complex128).view(*s).contiguous()
        a2afft.GridDfft(s, "fp64").forward(ref)
        got = a2afft.inputs.blocks_to_global(yr[lo:hi].view(8, -1), dims[i], s).reshape(-1)
        errs_r.append(a2afft.error_metrics(got.contiguous(), ref.reshape(-1))[0])
        del ref, got
    br.autotune(yr, iters=3)
    del yr
    rot = Rotation(torch, xr, rotation_count(xr, args.steps))
    t_ranks = _events_ms(torch, lambda buf: br.forward(buf), rot, args.steps, args.warmup)
    assert br.status() == 0
    finite = finite and rot.finite()
    del rot, xr
    npts = sum(s[0] * s[1] * s[2] for s in shapes)
    flops = sum(5.0 * (s[0] * s[1] * s[2]) * math.log2(s[0] * s[1] * s[2]) for s in shapes)
    bytes_alg = 6 * npts * 8
    best = min(t_eager, t_graph)
    ach = bytes_alg / (best * 1e-3) / 1e9
    counts = {f"{s[0]}x{s[1]}x{s[2]}": shapes.count(s) for s in C5_SHAPES}
    out = {
        "metric": "aggregate GFLOP/s, 64 uneven c2c 3D FFTs in one persistent launch",
        "value": round(flops / (best * 1e-3) / 1e9, 1), "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(best, 4), "higher_is_better": True, "scaling": "strong",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 complex64 grids drawn with seed 4 from 32^3, 64^3, 128^3, 256^3, 128x256x64, "
                               "256x64x512; U[-1,1) values", "name": "c5", "counts": counts,
                   "points": npts, "schedule": {k: sched[k] for k in ("workgroups", "xcd_queues", "chunk", "group_points",
                                                                  "skewed")}},
        "ms": {"persistent_eager": round(t_eager, 4), "persistent_graph": round(t_graph, 4),
               "separate_launches_192": round(t_sep, 4), "persistent_rank_blocks_dims222": round(t_ranks, 4)},
        "rel_l2_vs_fp64": {"max": max(errs), "mean": sum(errs) / len(errs)},
        "rank_blocks_rel_l2_vs_fp64": {"max": max(errs_r), "mean": sum(errs_r) / len(errs_r)},
        "roofline": {"bound": "hbm", "kernel": "batch_kernel (persistent)", "bytes_per_launch": bytes_alg,
                     "launch_ms": round(best, 4), "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": _traffic("c5_batch")},
        "finite": finite,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def precision_variants(a2afft, torch, Ng, P, dims, seed, steps, warmup=30, rounds=3):
    """The same Ng^3 GRF (complex64, block layout over dims) through the FP32,
    split-BF16 and split-FP16 plans: ms per forward (events over `steps`
    forwards on rotating copies of the input, the three variants interleaved
    over `rounds` rounds, median per variant) and rel-L2 / max-abs against the
    FP64 plan on the same rounded input."""
    import statistics
    g = make_grf(a2afft, torch, Ng, seed)
    x = a2afft.inputs.global_to_blocks(g.to(torch.complex64), dims).contiguous()
    del g
    dist64 = a2afft.Distribution(Ng, P, dims=dims, dtype="c128")
    ref = x.to(torch.complex128)
    a2afft.Dfft(dist64, precision="fp64", exec="fused").forward(ref)
    dist = a2afft.Distribution(Ng, P, dims=dims, dtype="c64")
    precs = ("fp32", "split_bf16", "split_fp16")
    plans = {p: a2afft.Dfft(dist, precision=p, exec="fused") for p in precs}
    out = {}
    data = torch.empty_like(x)
    for prec in precs:
        data.copy_(x)
        plans[prec].forward(data)
        rel, mx = a2afft.error_metrics(data, ref)
        out[prec] = {"rel_l2_vs_fp64": rel, "max_abs_vs_fp64": mx}
    del ref, data
    rot = Rotation(torch, x, rotation_count(x, steps))
    times = {p: [] for p in precs}
    for r in range(rounds):
        for prec in precs:
            times[prec].append(_events_ms(torch, lambda buf: plans[prec].forward(buf), rot, steps,
                                          warmup if r == 0 else max(1, warmup // 5)))
    finite = rot.finite()
    del rot, x
    for prec in precs:
        ms = statistics.median(times[prec])
        out[prec].update({"ms": round(ms, 4), "ms_rounds": [round(t, 4) for t in times[prec]],
                          "hbm_gbs_whole": round(6 * Ng ** 3 * 8 / (ms * 1e-3) / 1e9, 1)})
        log(f"[bench] {prec}: {ms:.4f} ms (rounds {times[prec]}) rel-L2 {out[prec]['rel_l2_vs_fp64']:.3e}")
    out["finite"] = finite
    torch.cuda.synchronize()
    return out


def run_c4(args, torch, a2afft):
    """512^3 complex64 GRF (seed 3) over dims [1,4,4]: split-BF16 MFMA radix-16
    passes beside plain FP32 and split-FP16, error vs the FP64 plan per variant."""
    Ng, P, dims, seed = args.ng or 512, 16, (1, 4, 4), 3
    v = precision_variants(a2afft, torch, Ng, P, dims, seed, args.steps, args.warmup)
    bf = v["split_bf16"]
    out = {
        "metric": "ms per 512^3 c2c forward, split-BF16 MFMA radix-16 vs plain FP32 (rel-L2 vs FP64)",
        "value": bf["ms"], "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": bf["ms"], "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "bf16x2 (MFMA) + f32", "data": "synthetic",
        "config": {"workload": f"{Ng}^3 c2c forward, complex64, {P} virtual ranks dims {list(dims)}, "
                               f"GRF P(k)~k^-3 seed {seed} + 1e-3 noise", "name": "c4"},
        "variants": v,
        "split_bf16_speedup_vs_fp32": round(v["fp32"]["ms"] / bf["ms"], 4),
        "split_bf16_within_1e-5": bf["rel_l2_vs_fp64"] <= 1e-5,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # 30 untimed forwards: after the FP64 accuracy check, 5 left the timed
    # loop 1.5 % slower than the steady state
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--config", default="c3", choices=["c3", "c2", "c3b", "c4", "c5", "grid"],
                    help="c3: 512^3 complex64 over dims [1,4,4], GRF seed 2 (the headline); "
                         "c2: 256^3 complex128 over dims [1,2,4], N(0,1) seed 1; "
                         "c3b: one rank's 2048x256x256 complex64 share of 2048^3 over dims [1,8,8], x stage; "
                         "c4: 512^3 GRF seed 3, FP32 vs split-BF16 vs split-FP16; "
                         "c5: 64 uneven complex64 grids (seed 4) in one persistent launch; "
                         "grid: one natural-order grid of --shape (any supported extents, up to 2^40 points)")
    ap.add_argument("--no-variants", action="store_true",
                    help="c3: skip the FP32 / split-BF16 / split-FP16 comparison on the seed-3 field")
    ap.add_argument("--precision", default=None, choices=["fp32", "fp64", "split_bf16", "split_fp16"])
    ap.add_argument("--ng", type=int, default=None)
    ap.add_argument("--shape", default="480x480x480", help="--config grid: NXxNYxNZ")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--autotune", action="store_true", help="time every kernel variant per pass and keep the fastest")
    ap.add_argument("--chunk", type=int, default=-1, help="z columns per x/y chunk (-1: plan default / autotuned)")
    ap.add_argument("--kernel", action="append", default=[],
                    help="force a variant: pass:T:E:flags (pass 0 z, 1 x, 2 y; flags 1 XCD, 2 pad, 4 "
                         "non-temporal, 8 z-tiles outermost); repeatable")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one GPU and the multi-GPU exchange is "
              "design-only; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import a2afft

    if args.config == "c3b":
        return run_c3b(args, torch, a2afft)
    if args.config == "c5":
        return run_c5(args, torch, a2afft)
    if args.config == "grid":
        return run_grid(args, torch, a2afft)
    if args.config == "c4":
        return run_c4(args, torch, a2afft)
    if args.config == "c2":
        Ng, P, dims, seed, field = args.ng or 256, 8, (1, 2, 4), 1, "normal"
        args.precision = args.precision or "fp64"
    else:
        Ng, P, dims, seed, field = args.ng or 512, 16, (1, 4, 4), 2, "grf"
        args.precision = args.precision or "fp32"
    dtype = "c128" if args.precision == "fp64" else "c64"
    esize = 8 if dtype == "c64" else 16
    N = Ng ** 3

    log(f"[bench] generating {Ng}^3 {field} input on the device")
    if field == "grf":
        g = make_grf(a2afft, torch, Ng, seed)
    else:
        g = torch.empty(Ng ** 3, dtype=torch.complex128, device="cuda")
        a2afft.fill_random(g, seed, "normal")
        g = g.reshape(Ng, Ng, Ng)
    x = a2afft.inputs.global_to_blocks(g.to(torch.complex64 if dtype == "c64" else torch.complex128), dims)
    del g
    x = x.contiguous()
    torch.cuda.synchronize()

    dist = a2afft.Distribution(Ng, P, dims=dims, dtype=dtype)
    plan = a2afft.Dfft(dist, precision=args.precision, exec="fused")
    plan.make_plans(None)
    for kspec in args.kernel:
        pi, T, E, flags = (int(v) for v in kspec.split(":"))
        plan.set_kernel(pi, T, E, flags)
    if args.chunk >= 0:
        plan.set_chunking(args.chunk)
    if args.autotune:
        log("[bench] autotuning kernel variants")
        work = x.clone()
        plan.autotune(work, iters=5)
        del work
    log(plan.describe())

    # ---- accuracy vs the FP64 path on the same (already rounded) input -----
    log("[bench] accuracy check vs FP64")
    data = x.clone()
    plan.forward(data)
    dist64 = a2afft.Distribution(Ng, P, dims=dims, dtype="c128")
    plan64 = a2afft.Dfft(dist64, precision="fp64", exec="fused")
    plan64.make_plans(None)
    ref = x.to(torch.complex128)
    plan64.forward(ref)
    rel_l2, max_abs = a2afft.error_metrics(data, ref)
    del ref
    torch.cuda.synchronize()

    # ---- timing ----------------------------------------------------------
    # Every timed forward runs on its own copy of the input (refilled between
    # timed loops, outside them), so no step transforms an already transformed
    # (and, after a few in-place forwards, overflowed) field.
    log(f"[bench] warmup {args.warmup}, timed {args.steps}")
    rot = Rotation(torch, x, rotation_count(x, args.steps))
    for i in range(args.warmup):
        plan.forward(rot[i])
    rot.refill()
    # timed region: K whole forwards, back to back, bracketed by synchronize
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        plan.forward(rot[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    finite = rot.finite()
    # the timed forwards computed the same bits as the accuracy check's
    same_bits = bool(torch.equal(torch.view_as_real(rot[0]), torch.view_as_real(data)))
    del data

    # per-kernel durations: the same K forwards launch by launch, with HIP
    # events between consecutive launches on the launch stream
    rot.refill()
    nl = plan.num_launches()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(nl + 1)] for _ in range(args.steps)]
    kinds = [0] * nl
    torch.cuda.synchronize()
    for s in range(args.steps):
        buf = rot[s]
        ev[s][0].record()
        for i in range(nl):
            kinds[i] = plan.execute_launch(buf, i, -1)
            ev[s][i + 1].record()
    torch.cuda.synchronize()
    pass_ms = [0.0, 0.0, 0.0]
    launch_ms = [0.0, 0.0, 0.0]
    nper = [kinds.count(p) for p in range(3)]
    for s in range(args.steps):
        for i in range(nl):
            pass_ms[kinds[i]] += ev[s][i].elapsed_time(ev[s][i + 1]) / args.steps
    ev_ms = sum(ev[s][0].elapsed_time(ev[s][nl]) for s in range(args.steps)) / args.steps
    # an event between two launches delays the second by a fixed amount:
    # measure it as the difference to the same forwards with events at the
    # ends only, and take it off every launch
    rot.refill()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for i in range(args.steps):
        plan.forward(rot[i])
    e1.record()
    torch.cuda.synchronize()
    ends_ms = e0.elapsed_time(e1) / args.steps
    gap_ms = max(0.0, (ev_ms - ends_ms) / nl)
    for p in range(3):
        launch_ms[p] = pass_ms[p] / max(1, nper[p]) - gap_ms
        pass_ms[p] = launch_ms[p] * nper[p]
    finite = finite and rot.finite()
    data = rot[0]
    del rot

    # measured device-copy bandwidth over the same buffer (the library's copy
    # kernel, fast_copy_fft): the practical ceiling of one read + one write pass
    copy_dst = torch.empty_like(data)
    for _ in range(3):
        a2afft.fast_copy_fft(copy_dst, data)
    c0, c1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    c0.record()
    for _ in range(10):
        a2afft.fast_copy_fft(copy_dst, data)
    c1.record()
    torch.cuda.synchronize()
    copy_ms = c0.elapsed_time(c1) / 10
    del copy_dst

    # each pass's access pattern alone: a copy kernel with the pass kernel's
    # addresses, tile order, cache policy and LDS footprint (no arithmetic)
    pattern_ms = None
    if args.precision in ("fp32", "fp64"):
        pattern_ms = [plan.pattern_copy(data, p, 10) for p in range(3)]

    bytes_pass = 2 * N * esize
    names = ["z", "x", "y"]
    dom = max(range(3), key=lambda p: pass_ms[p])
    # algorithmic bytes of one launch of the dominant kernel: the launch's
    # share of one read + one write of the grid
    bytes_launch = bytes_pass / max(1, nper[dom])
    achieved = bytes_launch / (launch_ms[dom] * 1e-3) / 1e9
    gflops = 5.0 * N * math.log2(N) / (ms_per_step * 1e-3) / 1e9
    whole_gbs = 3 * bytes_pass / (ms_per_step * 1e-3) / 1e9

    traffic = _traffic(f"{Ng}_{args.precision}_{names[dom]}")

    cpu = None
    if not args.no_cpu:
        threads, aff, quota, model = host_cores()
        log(f"[bench] CPU baseline on {threads} threads (affinity {aff}, cgroup quota {quota}), {model}")
        cpu_ms, samples, used = cpu_baseline(Ng, P, dims, threads, seed)
        cpu = {"value": round(cpu_ms, 1), "unit": "ms", "cores": used, "kind": "port",
               "sample": f"median of {len(samples)} full {Ng}^3 forwards, complex128, {P} virtual ranks dims "
                         f"{list(dims)}, reference stage order (oracle C restatement, OpenMP)",
               "samples_ms": [round(t, 1) for t in samples], "affinity_cores": aff, "cgroup_quota_cores": quota,
               "cpu_model": model}
        lib_ms = library_cpu_fftn(Ng, threads)
        if lib_ms is not None:
            cpu["library_fftn"] = {"value": round(lib_ms, 1), "unit": "ms", "cores": threads,
                                   "what": f"scipy.fft.fftn (pocketfft) of one {Ng}^3 complex128 grid, "
                                           "median of 3"}

    out = {
        "metric": "ms/512^3 c2c 3D FFT + % HBM roofline" if args.config == "c3" else f"ms/{Ng}^3 c2c 3D FFT",
        "value": round(ms_per_step, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "fp64": "f64", "split_bf16": "bf16x2 (MFMA) + f32",
                  "split_fp16": "fp16x2 (MFMA) + f32"}[args.precision],
        "data": "synthetic",
        "config": {"workload": f"{Ng}^3 c2c forward 3D FFT, {'complex64' if dtype == 'c64' else 'complex128'}, "
                               f"{P} virtual ranks dims {list(dims)}, "
                               + (f"GRF P(k)~k^-3 seed {seed} + 1e-3 noise" if field == "grf"
                                  else f"N(0,1) re/im seed {seed}"),
                   "name": args.config, "precision": args.precision, "exec": "fused", "Ng": Ng, "ranks": P,
                   "dims": list(dims),
                   "autotune": bool(args.autotune)},
        "gflops": round(gflops, 1),
        "hbm_gbs_whole": round(whole_gbs, 1),
        "roofline_frac_whole": round(whole_gbs / HBM_PEAK_GBS, 4),
        "pass_ms": {names[p]: round(pass_ms[p], 4) for p in range(3)},
        "pass_gbs": {names[p]: round(bytes_pass / (pass_ms[p] * 1e-3) / 1e9, 1) for p in range(3)},
        "launches": {names[p]: nper[p] for p in range(3)},
        "copy_gbs": round(bytes_pass / (copy_ms * 1e-3) / 1e9, 1),
        "pattern_copy_gbs": ({names[p]: round(bytes_pass / (pattern_ms[p] * 1e-3) / 1e9, 1) for p in range(3)}
                             if pattern_ms else None),
        "pass_vs_pattern_copy": ({names[p]: round(pattern_ms[p] / pass_ms[p], 3) for p in range(3)}
                                 if pattern_ms else None),
        "event_ms_per_step": round(ends_ms, 4),
        "event_gap_ms": round(gap_ms, 5),
        "rel_l2_vs_fp64": rel_l2,
        "max_abs_vs_fp64": max_abs,
        "finite": finite,
        "timed_equals_checked": same_bits,
        "rotating_inputs": True,
        "roofline": {"bound": "hbm", "kernel": f"stockham {names[dom]} pass", "bytes_per_launch": int(bytes_launch),
                     "launch_ms": round(launch_ms[dom], 4), "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "cpu_baseline": cpu,
    }
    if args.config == "c3" and args.precision in ("fp32", "fp64"):
        out["library_gpu"] = library_gpu_fftn(a2afft, torch, Ng, args.precision)
    if args.config == "c3" and not args.no_variants and args.ng in (None, 512):
        log("[bench] precision variants on the seed-3 field")
        out["precision_variants_seed3"] = precision_variants(a2afft, torch, Ng, P, dims, 3, args.steps, args.warmup)
    print(json.dumps(out), flush=True)
    if not finite:
        log("[bench] the timed transforms produced non-finite values")
        sys.exit(3)


if __name__ == "__main__":
    main()
