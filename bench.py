"""Benchmark: fwd+bwd frames/s at 1M splats, 512x512.

Headline (`value`): the reference's frame on the config-1 workload -- 1M surfels
on 64 random ellipsoids, one 512x512 view (focal 614.4), precision 0.01,
smoothing 0 -- through the drop-in renderer: affine + linear transform,
sequential Poisson sampling of k = 40 splats per pixel (LBVH, Hermite FGT,
capacity descents, keyed draws), EWA weights, per-pixel z-sort and
front-to-back alpha compositing, albedo interpolation ("None" shading), L2 loss
(sum) against a target rendered from centres perturbed by N(0, 0.01), and the
backward of every op.  Its roofline line is the sampler pair
(k_sample_walk + k_sample_draw), the frame's dominant kernels.

Extras on the same scene: the Poisson-rate path of the metric's comparison --
Poisson-rate sampling with inclusion probabilities (rate 10, 1/pi weights),
EWA weights, albedo interpolation without occlusion (the Horvitz-Thompson
estimate, unbiased), L2, fwd+bwd -- with the per-tile learned rates against the
plain ones at equal gradient variance; and the batched-view box room.

Prints one JSON line (rank 0, N = 1 GPU).  Kernel and per-op device times are
taken with HIP events on the stream the work launches on (torch's current
stream).  HBM traffic per launch comes from the rocprofv3 counter summaries
committed under profiles/.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "unbiased-differentiable-splatting_amd")
for _p in (PKG, ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "fwd+bwd frames/s at 1M splats 512² (learned vs plain rates, equal grad variance)"
HBM_PEAK_GBS = 8000.0
RATE, KCAP = 10.0, 40
PROFILES = os.path.join(ROOT, "profiles")


def measured_traffic(summary):
    """HBM bytes per launch from a counter summary under profiles/ (its
    'traffic_bytes_per_launch:' line: rocprofv3 FETCH_SIZE + WRITE_SIZE of the
    kernel, each scaled by the same counter on a 1 GiB device copy, bench.py
    --calibrate-copy), or None when the summary is missing."""
    try:
        with open(os.path.join(PROFILES, summary)) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


TRAFFIC_SUMMARY = {"k_sample": "hbm_traffic_k_sample_1M_512x512_k40.md",
                   "k_fmm_eval": "hbm_traffic_k_fmm_eval_1M_512x512.md",
                   "k_poisson": "hbm_traffic_k_poisson_1M_512x512.md"}


def build_problem(dev, n_override=None):
    import numpy as np
    import torch
    from splatting_amd import scenes
    from splatting_amd.render import Splatting, Options

    sc = scenes.config(1, n_override)
    M = torch.from_numpy(sc.view).to(dev)[None].contiguous()
    P = torch.from_numpy(sc.positions).to(dev)[None].contiguous().requires_grad_(True)
    Nn = torch.from_numpy(sc.normals).to(dev)[None].contiguous().requires_grad_(True)
    S = torch.from_numpy(sc.stddevs).to(dev)[None].contiguous().requires_grad_(True)
    A = torch.from_numpy(sc.albedo).to(dev)[None].contiguous().requires_grad_(True)
    amb = torch.zeros(1, 3, device=dev)
    lamp_d = torch.tensor([[[0.0, 0.0, 1.0]]], device=dev)
    lamp_i = torch.ones(1, 1, 3, device=dev)
    opts = Options(numSamples=40, precision=0.01, shadingMode="None", device=dev)
    r = Splatting(opts, None, sc.width, sc.height, sc.focal, 0.1, 50.0, 0.0)
    r.keepIndicesOnHost = False
    with torch.no_grad():
        tgt_pos = torch.from_numpy(scenes.perturbed_target_positions(sc)).to(dev)[None].contiguous()
        r.seed = 10 ** 6
        target, _, _ = r(None, None, tgt_pos, Nn.detach(), S.detach(), A.detach(), amb, lamp_d, lamp_i,
                         transformMatrix=M)
        r.seed = None
    return sc, r, (P, Nn, S, A), (amb, lamp_d, lamp_i, M), target


class OpTimer:
    """Wraps the op-surface functions and records HIP events around each call."""

    def __init__(self):
        import torch
        from cudaExtensions import sampling, splatting, filtering, interpolation, transform
        self.torch = torch
        self.events = {}
        self.active = False
        self._wrap(sampling, "sample", "sample")
        self._wrap(splatting, "forward", "splat_fwd")
        self._wrap(splatting, "backward", "splat_bwd")
        self._wrap(filtering, "forward", "filter_fwd")
        self._wrap(filtering, "backward", "filter_bwd")
        self._wrap(interpolation, "forward", "interp_fwd")
        self._wrap(interpolation, "backward", "interp_bwd")
        for nm in ("affineForward", "linearForward", "affineBackward", "linearBackward"):
            self._wrap(transform, nm, nm)

    def _wrap(self, mod, name, key):
        fn = getattr(mod, name)
        timer = self

        def wrapped(*a, **kw):
            if not timer.active:
                return fn(*a, **kw)
            e0 = timer.torch.cuda.Event(enable_timing=True)
            e1 = timer.torch.cuda.Event(enable_timing=True)
            e0.record()
            out = fn(*a, **kw)
            e1.record()
            timer.events.setdefault(key, []).append((e0, e1))
            return out
        setattr(mod, name, wrapped)

    def summary(self, steps):
        out = {}
        for k, evs in self.events.items():
            tot = sum(a.elapsed_time(b) for a, b in evs)
            out[k] = tot / steps
        return out


def algorithmic_bytes(N, P, k):
    """Per-op algorithmic HBM bytes of one fwd+bwd frame (each op-surface tensor counted once)."""
    return {
        "affineForward": 24 * N, "linearForward": 24 * N + 0,
        "sample": 32 * N + 4 * P * k,
        "splat_fwd": 32 * N + 8 * P * k,
        "filter_fwd": 24 * N + 12 * P * k,
        "interp_fwd": 12 * N + 8 * P * k + 12 * P,
        "interp_bwd": 24 * N + 12 * P * k + 12 * P,
        "filter_bwd": 24 * N + 16 * P * k,
        "splat_bwd": 64 * N + 8 * P * k,
        "affineBackward": 36 * N, "linearBackward": 36 * N,
    }


def _reference_frame_on_host(sc, rows, k, threads):
    """The reference's frame by the oracle (C, OpenMP, `threads` threads): the full
    tree build, then sampling and EWA / compositing / interpolation fwd+bwd (FP64)
    on `rows` pixel rows through the image centre.  Returns (tree seconds, seconds
    of the per-row part)."""
    import numpy as np
    from oracle import oracle as O
    O.set_threads(threads)
    p, n = sc.camera_space()
    nb = np.where(n[:, 2:3] > 0, -n, n).astype(np.float32)
    W, H, f = sc.width, sc.height, sc.focal
    i0 = W // 2 - rows // 2
    t0 = time.time()
    idx, t_tree = O.sample_view_rows(W, H, f, p, nb, sc.stddevs, k, i0, i0 + rows, 0.0, 0.01, seed=3)
    t_sample = time.time() - t0 - t_tree
    idx4 = idx[None]
    p4, n4, s4, a4 = p[None].astype(np.float64), n[None].astype(np.float64), sc.stddevs[None].astype(np.float64), \
        sc.albedo[None].astype(np.float64)
    t1 = time.time()
    w = O.splat_forward(W, H, f, idx4, p4, n4, s4)
    dw = O.filter_forward(idx4, w, W, H, f, p[None], n[None])
    img = O.interp_forward(idx4, dw, a4)
    g = 2.0 * img
    gW, gA = O.interp_backward(g, idx4, dw, a4)
    gw = O.filter_backward(idx4, w, gW, W, H, f, p[None], n[None])
    O.splat_backward(W, H, f, gw, idx4, p4, n4, s4)
    return t_tree, t_sample + (time.time() - t1)


def cpu_model():
    """The host CPU's model name (lscpu, else /proc/cpuinfo)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sc, k=40, threads=None):
    """Reference semantics on the host cores: the oracle's whole frame (all 512
    rows) on the box's thread share, the same frame on one thread (16 rows, the
    per-row part scaled to the image), and the FP64 direct Gaussian sums (the
    density the Hermite FGT approximates) at 4,096 random pixels over every culled
    splat."""
    import math
    import numpy as np
    from oracle import oracle as O
    threads = threads or min(16, os.cpu_count() or 1)
    W = sc.width
    tt, tr = _reference_frame_on_host(sc, W, k, threads)
    rows1 = 16
    tt1, tr1 = _reference_frame_on_host(sc, rows1, k, 1)
    frame1 = tt1 + tr1 * W / float(rows1)
    p, n = sc.camera_space()
    nb = np.where(n[:, 2:3] > 0, -n, n).astype(np.float32)
    O.set_threads(threads)
    T = O.Tree(sc.width, sc.height, sc.focal, p, nb, sc.stddevs, 0.0, 0.01)
    h = float(T.h)
    leaves = np.ascontiguousarray(T.cen[T.nc - 1:], np.float32)
    rng = np.random.default_rng(4096)
    q = np.stack([rng.integers(0, sc.width, 4096), rng.integers(0, sc.height, 4096)], 1).astype(np.float64)
    c = 1.0 / (2.0 * math.pi * math.sqrt(h))
    t0 = time.time()
    O.direct_sums(leaves, h, c, q)
    t_ds = time.time() - t0
    O.set_threads(1)
    t0 = time.time()
    O.direct_sums(leaves, h, c, q[:256])
    t_ds1 = (time.time() - t0) * 16.0
    O.set_threads(threads)
    return {"value": 1.0 / (tt + tr), "unit": "frames/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": "the whole frame on %d threads: oracle tree build over all %d splats (%.2fs), then sequential "
                      "Poisson sampling (k = %d) and EWA / compositing / interpolation fwd+bwd in FP64 over all "
                      "%dx%d pixels (%.2fs)" % (threads, sc.n, tt, k, sc.width, sc.height, tr),
            "single_thread": {"value": 1.0 / frame1, "unit": "frames/s", "cores": 1,
                              "sample": "the same frame on 1 thread: tree build (%.2fs) + %d of %d pixel rows "
                                        "(%.2fs) scaled x%g" % (tt1, rows1, W, tr1, W / float(rows1))},
            "direct_sums": {"pixels": 4096, "splats": int(leaves.shape[0]),
                            "gaussians": 4096 * int(leaves.shape[0]), "seconds": t_ds, "cores": threads,
                            "single_thread_seconds": t_ds1,
                            "sample": "FP64 exp sums at 4096 random pixels on %d threads; the 1-thread time is 256 "
                                      "pixels scaled x16" % threads}}


def timed_steps(fn, steps, kernel=None):
    """Runs `steps` calls of fn between a device synchronisation on both sides;
    returns (ms per step, summed ms and launches of the profiled sampling kernel)."""
    import torch
    from cudaExtensions import sampling
    torch.cuda.synchronize()
    if kernel:
        sampling.profile_sampler_begin(kernel)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.time()
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    wall = time.time() - t0
    k_ms, k_n = sampling.profile_sampler_end() if kernel else (0.0, 0)
    return max(e0.elapsed_time(e1), wall * 1e3) / steps, k_ms, k_n


def poisson_paths(dev, fr, seeds=64, probe_steps=10, compare=True):
    """The plain-rate step (band proposal at RATE) and the alternatives that reach
    its gradient variance: the footprint proposal at the rate its measured kappa
    gives, and, when the trained rate network needs fewer samples than the plain
    rates, the learned-rate step.  Each alternative with kappa < 1 is checked at
    its rate and timed against the plain step; the fastest is chosen.  Returns
    (chosen step, its name, the comparison record)."""
    import torch
    from cudaExtensions import sampling
    from splatting_amd import learned as L
    counter = [0]

    def plain_step():
        counter[0] += 1
        fr.step(1000 + counter[0])

    rec = {"estimator": "Horvitz-Thompson (sum of min(w, 1) / pi times albedo, no occlusion), L2 loss, raw adjoints",
           "plain_rate": RATE, "kcap": KCAP}
    steps = {"plain": plain_step}
    notes = []
    if not compare:
        rec["chosen"] = "plain (proposals and learned rates not compared)"
        return plain_step, "plain", rec
    # footprint proposal (the wide band plus each splat's own EWA footprint)
    cmpf = L.compare_proposals(fr, list(range(30_000, 30_000 + 2 * seeds)), groups=16)
    rec["footprint"] = cmpf
    if cmpf["kappa"] < 1.0:
        # the rate of equal variance: bracketed from the kappa estimate (the
        # variance falls faster than 1 / rate over this range: heavy tails at low
        # rates), then bisected in log rate; each value is the median of 8 seed
        # groups of 8
        kf = cmpf["kcap_footprint"]
        target_v = cmpf["band_grad_variance"]
        tried = []

        def var_at(r, it):
            fr.proposal = "footprint"
            try:
                v, _ = L.gradient_variance(fr, list(range(60_000 + 1000 * it, 60_000 + 1000 * it + seeds)), rate=r,
                                           kcap=kf, groups=8)
            finally:
                fr.proposal = "bands"
            tried.append([r, v])
            return v

        r0 = cmpf["footprint_rate_equal_samples"] * cmpf["kappa"]
        lo = hi = None
        r, it = r0, 0
        while (lo is None or hi is None) and it < 6:
            if var_at(r, it) > target_v:
                lo = r
                r = r * 2.0 if hi is None else (r * hi) ** 0.5
            else:
                hi = r
                r = r * 0.5 if lo is None else (lo * r) ** 0.5
            it += 1
        if lo is not None and hi is not None:
            for _ in range(2):
                mid = (lo * hi) ** 0.5
                if var_at(mid, it) > target_v:
                    lo = mid
                else:
                    hi = mid
                it += 1
        cmpf["rate_search"] = tried
        if hi is None:
            notes.append("the footprint proposal did not reach the band variance at the rates tried")
        else:
            rf = hi                                       # the lowest rate measured at or below the target
            cmpf.update({"rate_equal_variance": rf, "grad_variance_at_rate": [v for r_, v in tried if r_ == rf][-1]})

            def footprint_step():
                counter[0] += 1
                fr.proposal = "footprint"
                try:
                    fr.step(7000 + counter[0], rate=rf, kcap=kf)
                finally:
                    fr.proposal = "bands"

            steps["footprint"] = footprint_step
    else:
        notes.append("footprint needs %.2fx the samples" % cmpf["kappa"])
    # learned per-tile rates
    if not os.path.exists(L.DEFAULT_WEIGHTS):
        notes.append("no trained rate network")
    else:
        net = L.RateNet.load(L.DEFAULT_WEIGHTS, dev)
        cmp = L.compare_rates(fr, net, list(range(20_000, 20_000 + seeds)))
        ts = cmp.pop("tile_scale")
        rec.update(cmp)
        kappa = cmp["kappa"]
        if kappa >= 1.0:
            notes.append("the learned rates need %.2fx the samples" % kappa)
        else:
            kcl = cmp["kcap_learned"]

            def learned_step():
                counter[0] += 1
                _, _, _, g = fr.step(5000 + counter[0], tile_scale=ts, rate=RATE * kappa, kcap=kcl)
                P, Nc, nb = fr.camera()
                f2 = sampling.tileFeatures(fr.W, fr.H, fr.f, P, nb, fr.S, 0.0, fr.prec, fr.dens, g)
                ts.copy_(net.tile_scale(f2[0], L.tile_counts(fr.last_ids)[0])[None])

            v_check, _ = L.gradient_variance(fr, list(range(40_000, 40_000 + seeds)), tile_scale=ts,
                                             rate=RATE * kappa, kcap=kcl, groups=4)
            rec.update({"learned_rate": RATE * kappa, "learned_grad_variance_at_rate": v_check})
            steps["learned"] = learned_step
    print("bench: footprint kappa %.3f %s, samples per pixel %.2f (bands %.2f)" % (
        cmpf["kappa"], cmpf["kappa_ci95"], cmpf["samples_per_pixel_footprint"], cmpf["samples_per_pixel_bands"]),
        file=sys.stderr, flush=True)
    probe = {name: timed_steps(st, probe_steps)[0] for name, st in steps.items()}
    rec["probe_ms"] = probe
    chosen = min(probe, key=probe.get)
    rec["chosen"] = chosen + ("" if not notes else " (%s)" % "; ".join(notes))
    return steps[chosen], chosen, rec


def reference_path(dev, args, sc_n):
    """The reference's method on the same scene: sequential Poisson sampling of
    k = 40 splats per pixel, EWA weights, z-sort / compositing and interpolation
    through the drop-in renderer, fwd+bwd; the sampler kernel timed with HIP events."""
    import torch
    sc, r, params, cam, target = build_problem(dev, args.splats)
    P, Nn, S, A = params
    amb, lamp_d, lamp_i, M = cam
    timer = OpTimer()

    def step():
        for t in params:
            t.grad = None
        image, _, _ = r(None, None, P, Nn, S, A, amb, lamp_d, lamp_i, transformMatrix=M)
        loss = ((image - target) ** 2).sum()
        loss.backward()
        return loss

    for _ in range(args.warmup):
        step()
    # the frame rate from plain steps; the per-op events and the sampler's kernel
    # events from separate instrumented steps (they add their own launches)
    ms, _, _ = timed_steps(step, args.steps)
    timer.active = True
    _, ks_ms, ks_n = timed_steps(step, args.steps, "sample")
    timer.active = False
    # the reference's renderer also copies the index tensor to the host every frame
    # (render.py:296); the frame above keeps it on the device, this one copies it
    r.keepIndicesOnHost = True
    ms_copy, _, _ = timed_steps(step, max(args.steps // 2, 1))
    r.keepIndicesOnHost = False
    ops = timer.summary(args.steps)
    Np, Pp, k = sc.n, sc.width * sc.height, 40
    nbytes = algorithmic_bytes(Np, Pp, k)
    ks_avg = ks_ms / max(ks_n, 1)
    achieved = nbytes["sample"] / (ks_avg * 1e-3) / 1e9
    return {"method": "sequential Poisson sampling, k=40 per pixel, z-sort/compositing (reference semantics)",
            "fps": 1000.0 / ms, "ms_per_step": ms, "ms_per_step_with_index_copy_to_host": ms_copy,
            "roofline": {"bound": "hbm", "kernel": "k_sample (k_sample_walk + k_sample_draw)", "achieved": achieved,
                         "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": measured_traffic(TRAFFIC_SUMMARY["k_sample"]),
                         "kernel_ms": ks_avg, "kernel_launches": ks_n, "algorithmic_bytes": nbytes["sample"]},
            "ops_ms": {k_: round(v, 4) for k_, v in sorted(ops.items())}}


def batched_views(dev, steps, warmup, n_splats=2_000_000, views=16):
    """Batched-view workload: the 8 x 6 x 3 m box room (2M surfels on the walls and
    on 200 boxes, std-devs U[0.01, 0.03]) seen by 16 cameras at 640x480 with uneven
    culled counts; one fwd+bwd Poisson-rate frame over all views per step, with
    every stage one launch over the batch, launch shapes and the Hermite order of
    the densities from the autotuner, the whole frame replayed as one captured HIP
    graph (the Philox key advances inside the graph)."""
    import torch
    from cudaExtensions import sampling
    from splatting_amd import autotune, learned as L, scenes
    sc = scenes.room_scene(n_splats, views)
    fr = L.PoissonFrame(sc, dev, rate=RATE, kcap=KCAP)
    fr.target = torch.zeros((fr.V, fr.W, fr.H, 3), dtype=torch.float32, device=dev)
    P, Nc, nb = fr.camera()
    in_front = (P[..., 2] < 0).sum(1).tolist()

    def sample(batched=True):
        sampling.samplePoisson(fr.W, fr.H, fr.f, P, nb, None, fr.S, KCAP, 0.0, fr.prec, seed=1, rate=RATE,
                               densities=fr.dens, batched=batched)

    saved = {k: sampling.get_tile_config(k) for k in autotune.KERNELS}
    tuning = autotune.tune(sample)
    tuning["degree"] = autotune.tune_degree(sample)
    seq_ms, _, _ = timed_steps(lambda: sample(False), 3)
    bat_ms, _, _ = timed_steps(sample, 3)
    for _ in range(warmup):
        fr.step(7, with_loss=False)
    eager_ms, _, _ = timed_steps(lambda: fr.step(11, with_loss=False), steps)
    key = torch.tensor([1 << 20], dtype=torch.int64, device=dev)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fr.step(0, seed_tensor=key, with_loss=False)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        key.add_(1)
        fr.step(0, seed_tensor=key, with_loss=False)
    for _ in range(warmup):
        graph.replay()
    graph_ms, _, _ = timed_steps(graph.replay, steps)
    sampling.check_errors()
    for k, c in saved.items():
        sampling.set_tile_config(k, *c)
    sampling.set_fixed_degree(6)
    best = min(eager_ms, graph_ms)
    return {"workload": "box room 8x6x3 m, %d surfels on walls + 200 boxes, %d cameras at 640x480, "
                        "Poisson-rate sampling at rate %g, fwd+bwd over all views per step" % (n_splats, views, RATE),
            "views": views, "width": fr.W, "height": fr.H, "splats": n_splats, "splats_in_front_per_view": in_front,
            "frames_per_s": 1000.0 / best, "views_per_s": views * 1000.0 / best,
            "graph_ms_per_step": graph_ms, "eager_ms_per_step": eager_ms,
            "sampling_ms": {"one_launch_per_pass": bat_ms, "view_by_view": seq_ms},
            "autotune": tuning,
            "memory_gb": torch.cuda.max_memory_allocated() / 1e9}


def poisson_extra(dev, args, sc):
    """The Poisson-rate path of the metric's comparison on the same scene: plain
    against learned rates at equal gradient variance, the chosen step timed, and
    the rooflines of its two per-pixel passes."""
    from splatting_amd import learned as L
    target = L.make_target(sc, dev, RATE, KCAP)
    fr = L.PoissonFrame(sc, dev, rate=RATE, kcap=KCAP, target=target)
    print("bench: proposals and rates at equal gradient variance", file=sys.stderr, flush=True)
    step, chosen, rates = poisson_paths(dev, fr, compare=not args.no_rates)
    for _ in range(args.warmup):
        step()
    print("bench: Poisson-rate frames (%s rates)" % chosen, file=sys.stderr, flush=True)
    ms, _, _ = timed_steps(step, args.steps)
    _, fe_ms, fe_n = timed_steps(step, max(args.steps // 2, 1), "fmm_eval")
    _, po_ms, po_n = timed_steps(step, max(args.steps // 2, 1), "poisson")
    # Algorithmic bytes per launch, n = splats in view, P = pixels:
    #   k_fmm_eval (both bands): per band the 80 B children record of each of the n - 1
    #     internal nodes and one 4 B density per pixel; the expansion coefficients it
    #     reads (196 B per accepted node, a count only the device knows) are not counted,
    #     so this is a lower bound;
    #   k_poisson: the 32 B walk record of each of the 2n - 1 nodes, 8 B position and
    #     4 B id per leaf, two 4 B densities per pixel and 8 B (index, 1/pi) per slot.
    n_vis = int((L.splat_tiles(fr) >= 0).sum().item())
    P = sc.width * sc.height
    spp = float((fr.last_ids >= 0).sum().item()) / (P * fr.V)
    fr.check_errors()
    kern = {
        "k_fmm_eval": {"kernel_ms": fe_ms / max(fe_n, 1), "kernel_launches": fe_n,
                       "algorithmic_bytes": 2 * (80 * (n_vis - 1) + 4 * P)},
        "k_poisson": {"kernel_ms": po_ms / max(po_n, 1), "kernel_launches": po_n,
                      "algorithmic_bytes": 32 * (2 * n_vis - 1) + 12 * n_vis + 8 * P + 8 * P * KCAP},
    }
    for name, k in kern.items():
        k["achieved"] = k["algorithmic_bytes"] / (k["kernel_ms"] * 1e-3) / 1e9
        k["frac"] = k["achieved"] / HBM_PEAK_GBS
        k["traffic"] = measured_traffic(TRAFFIC_SUMMARY[name])
    return {"workload": "the same scene, Poisson-rate sampling at rate %g (%.2f samples per pixel, 1/pi weights), EWA "
                        "weights, albedo interpolation without occlusion (Horvitz-Thompson, unbiased), L2, fwd+bwd; "
                        "the fastest of the band proposal at rate %g, the footprint proposal and the learned rates "
                        "at the band proposal's gradient variance: %s" % (RATE, spp, RATE, chosen),
            "fps": 1000.0 / ms, "ms_per_step": ms, "samples_per_pixel": spp, "splats_in_view": n_vis,
            "compositing": False, "rates": rates, "kernels": kern}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--splats", type=int, default=None, help="override the splat count (testing only)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-rates", action="store_true",
                    help="time the band proposal at the plain rate without the footprint / learned comparisons")
    ap.add_argument("--no-poisson", action="store_true", help="skip the Poisson-rate path")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-view (box room) workload")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a device copy of known size (1 GiB read + 1 GiB written per launch), "
                         "to scale HBM traffic counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        src.fill_(1.0)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4, "launches": 3}))
        return
    from splatting_amd import scenes
    print("bench: reference frame", file=sys.stderr, flush=True)
    sc = scenes.config(1, args.splats)
    ref = reference_path(dev, args, sc.n)
    roof = ref.pop("roofline")
    line = {
        "metric": METRIC, "value": ref["fps"], "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ref["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "1M surfels on 64 ellipsoids, one 512x512 view, the reference's frame: sequential "
                               "Poisson sampling of k = 40 splats per pixel, EWA weights, per-pixel z-sort and "
                               "front-to-back alpha compositing, albedo interpolation, L2 loss, fwd+bwd",
                   "splats": sc.n, "width": sc.width, "height": sc.height, "samples_per_pixel": 40,
                   "compositing": True, "parallelism": "none (1 GPU)"},
        "roofline": roof,
        "ops_ms": ref.pop("ops_ms"),
        "ops_ms_note": "HIP events around each op call in separate instrumented steps: an op's span includes any "
                       "time the device waits for the host to enqueue the op's launches",
        "index_copy_to_host": {"ms_per_step": ref.pop("ms_per_step_with_index_copy_to_host"),
                               "note": "the same frame with the renderer's per-frame copy of the (W, H, 40) index "
                                       "tensor to the host, as the reference's render.py:296 does; value keeps it "
                                       "on the device"},
    }
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(sc)
    if not args.no_poisson:
        line["poisson_path"] = poisson_extra(dev, args, sc)
    if not args.no_batched:
        print("bench: batched views", file=sys.stderr, flush=True)
        line["batched_views"] = batched_views(dev, max(args.steps // 2, 3), 2)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
