"""Benchmark: one fwd+bwd frame of Poisson-sampled EWA splatting per step.

Workload: 1M surfels on 64 random ellipsoids, one 512x512 view (focal 614.4),
k = 40 samples per pixel, precision 0.01, smoothing 0, shading mode "None"
(albedo interpolation), L2 loss (sum) against a target rendered from centres
perturbed by N(0, 0.01).  One step = affine + linear transform, per-pixel
sampling (LBVH + Hermite FGT + sequential Poisson), EWA weights, z-sort /
compositing, interpolation, loss, and the backward of all of them.

Prints one JSON line (rank 0, N = 1 GPU).  Per-op device times are taken
with HIP events on the stream the ops launch on (torch's current stream).
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
# HBM bytes per launch of the sampler kernel at the 1M-splat 512x512 k=40 frame: rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE on k_sample, each scaled by the same counter on a 1 GiB
# device copy (bench.py --calibrate-copy); measured offline, not by this script
TRAFFIC_BYTES = 3.374e9


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


def cpu_baseline(sc, rows=None, k=40, threads=None):
    """Oracle (C, OpenMP) on the same frame: full tree build, then sampling +
    EWA/compositing/interpolation fwd+bwd on `rows` pixel rows (default: all;
    fewer rows are scaled to the full image)."""
    import numpy as np
    from oracle import oracle as O
    threads = threads or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    p, n = sc.camera_space()
    nb = np.where(n[:, 2:3] > 0, -n, n).astype(np.float32)
    W, H, f = sc.width, sc.height, sc.focal
    rows = rows or W
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
    t_ops = time.time() - t1
    frame = t_tree + (W / float(rows)) * (t_sample + t_ops)
    return {"value": 1.0 / frame, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": "oracle tree build over all %d splats (%.2fs) + sampling and EWA/compositing/"
                      "interpolation fwd+bwd on %d of %d pixel rows (%.2fs), per-row part scaled x%g"
                      % (sc.n, t_tree, rows, W, t_sample + t_ops, W / float(rows))}


def rates_comparison(dev, sc, steps, rate=10.0, kcap=40, seeds=32):
    """Poisson-rate sampling with 1/pi weights on the same scene: the plain FMM-derived
    rates against the per-tile learned rates at equal gradient variance.  The learned
    multipliers come from the trained rate network (splatting_amd/rate_net.npz) fed
    with the tile features of the previous frame; the rate that equalises the mean
    per-parameter gradient variance is rate * V_learned / V_plain (variance scales
    with 1 / rate), checked by a second measurement.  When the learned rates need
    more samples than the plain ones for the same variance, the plain rates are used
    (they are the network's fallback) and reported as the chosen path."""
    import os
    import torch
    from cudaExtensions import sampling
    from splatting_amd import learned as L

    def timed(fn, n):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.time()
        e0.record()
        for _ in range(n):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return max(e0.elapsed_time(e1), (time.time() - t0) * 1e3) / n

    target = L.make_target(sc, dev, rate, kcap)
    fr = L.PoissonFrame(sc, dev, rate=rate, kcap=kcap, target=target)
    counter = [0]

    def plain_step():
        counter[0] += 1
        fr.step(1000 + counter[0])

    plain_ms = timed(plain_step, steps)
    seed_list = list(range(20_000, 20_000 + seeds))
    v_plain, _ = L.gradient_variance(fr, seed_list)
    out = {"estimator": "Horvitz-Thompson, linear splatting, L2 loss, raw adjoints",
           "plain_rate": rate, "kcap": kcap, "plain_fps": 1000.0 / plain_ms, "plain_ms": plain_ms,
           "plain_grad_variance": v_plain, "variance_seeds": seeds}
    wpath = L.DEFAULT_WEIGHTS
    if not os.path.exists(wpath):
        out.update({"learned": None, "chosen": "plain (no trained rate network)",
                    "equal_variance_fps": 1000.0 / plain_ms})
        return out
    net = L.RateNet.load(wpath, dev)
    _, _, _, g = fr.step(99)
    P, Nc, nb = fr.camera()
    feats = sampling.tileFeatures(fr.W, fr.H, fr.f, P, nb, fr.S, 0.0, fr.prec, fr.dens, g)
    _, (_, _, ids, _, _) = fr.image(98)
    counts = L.tile_counts(ids)[0]
    ts = net.tile_scale(feats[0], counts)[None].contiguous()
    v_learned, _ = L.gradient_variance(fr, seed_list, tile_scale=ts, kcap=4 * kcap)
    kappa = v_learned / v_plain
    out.update({"learned_grad_variance_equal_budget": v_learned, "rate_factor_for_equal_variance": kappa,
                "tile_scale_range": [float(ts.min().item()), float(ts.max().item())]})
    if kappa < 1.0:
        v_check, _ = L.gradient_variance(fr, seed_list, tile_scale=ts, rate=rate * kappa, kcap=4 * kcap)

        def learned_step():
            counter[0] += 1
            _, _, _, gg = fr.step(5000 + counter[0], tile_scale=ts, rate=rate * kappa, kcap=4 * kcap)
            Pq, Ncq, nbq = fr.camera()
            f2 = sampling.tileFeatures(fr.W, fr.H, fr.f, Pq, nbq, fr.S, 0.0, fr.prec, fr.dens, gg)
            net.forward(f2.reshape(-1, 32))

        learned_ms = timed(learned_step, steps)
        out.update({"learned_fps": 1000.0 / learned_ms, "learned_grad_variance_at_rate": v_check})
        chosen = "learned" if learned_ms < plain_ms else "plain (fallback: learned slower)"
        out.update({"chosen": chosen, "equal_variance_fps": 1000.0 / min(plain_ms, learned_ms),
                    "speedup_vs_plain": plain_ms / min(plain_ms, learned_ms)})
    else:
        out.update({"learned_fps": None, "chosen": "plain (fallback: learned rates need %.2fx the samples)" % kappa,
                    "equal_variance_fps": 1000.0 / plain_ms, "speedup_vs_plain": 1.0})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--splats", type=int, default=None, help="override the splat count (testing only)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-rates", action="store_true", help="skip the learned-vs-plain rates comparison")
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
    print("bench: building workload", file=sys.stderr, flush=True)
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

    print("bench: warmup", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    timer.active = True
    from cudaExtensions import sampling
    sampling.profile_sampler_begin()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.time()
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    wall = time.time() - t0
    timer.active = False
    ks_ms, ks_n = sampling.profile_sampler_end()
    dev_ms = e0.elapsed_time(e1)
    ms_per_step = max(dev_ms, wall * 1e3) / args.steps
    fps = 1000.0 / ms_per_step
    ops = timer.summary(args.steps)
    Np, Pp, k = sc.n, sc.width * sc.height, 40
    nbytes = algorithmic_bytes(Np, Pp, k)
    # dominant kernel: the sampler kernel (one launch per view), timed with HIP events
    # around each launch on its own stream; its algorithmic bytes are the sample op's
    ks_avg = ks_ms / max(ks_n, 1)
    achieved = nbytes["sample"] / (ks_avg * 1e-3) / 1e9
    frame_bytes = 68 * Pp * k + 332 * Np + 24 * Pp
    line = {
        "metric": METRIC, "value": fps, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "1M surfels on 64 ellipsoids, one 512x512 view, k=40, precision 0.01, plain rates",
                   "splats": Np, "width": sc.width, "height": sc.height, "samples_per_pixel": k,
                   "parallelism": "none (1 GPU)"},
        "roofline": {"bound": "hbm", "kernel": "k_sample", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": TRAFFIC_BYTES,
                     "kernel_ms": ks_avg, "kernel_launches": ks_n, "algorithmic_bytes": nbytes["sample"],
                     "frame_frac": frame_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "ops_ms": {k_: round(v, 4) for k_, v in sorted(ops.items())},
    }
    if not args.no_rates:
        print("bench: learned vs plain rates", file=sys.stderr, flush=True)
        line["rates"] = rates_comparison(dev, sc, args.steps)
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(sc)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
