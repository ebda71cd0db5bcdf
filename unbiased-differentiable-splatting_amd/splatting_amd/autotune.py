"""Autotuner for the launch shapes of the per-pixel sampling passes (no
counterpart in the reference, whose kernels have fixed launch shapes).

The two passes that dominate Poisson-rate sampling -- the fixed-order density
evaluation ("fmm_eval", two launches per batch) and the thinning pass
("poisson") -- run one wavefront per pixel tile.  Their speed depends on the
tile shape (how coherent the 64 pixels' tree walks are), the number of
wavefronts per workgroup and whether workgroups are ordered so that each XCD
works on one band of the image; their results do not (each pixel's walk and
draws are its own), so any choice is safe and the tuner only times.

``tune(run)`` calls ``run()`` -- one call of the real workload, e.g.
``samplePoisson`` on the batch of views to be rendered -- under every
candidate shape, times each kernel's launches with HIP events on their stream,
keeps the fastest shape per kernel and leaves it set.  ``tune_degree(run)``
picks the order of the fixed-order Hermite expansions (4 .. 8) the same way,
timing whole calls: every order keeps the precision bound (a node expands only
where its truncation bound at that order meets it), lower orders expand fewer,
smaller nodes with cheaper series, so the fastest order depends on the
workload.  Unlike the shapes, the order changes the densities (within the
bound) and therefore the drawn samples.  Results can be cached in a JSON file
keyed by the workload (``workload_key``).
"""
import json
import os
import time

SHAPES = [(8, 8), (16, 4), (4, 16), (32, 2)]
WAVES = [1, 2, 4, 8]
XCD = [0, 1]
KERNELS = ("fmm_eval", "poisson")


def workload_key(width, height, views, splats, kcap):
    """Cache key: image size, views, splat count rounded to 2 significant digits, capacity."""
    s = float("%.2g" % max(int(splats), 1))
    return "%dx%d_v%d_n%d_k%d" % (width, height, views, int(s), kcap)


def _time(run, kernel, repeats):
    from cudaExtensions import sampling
    best = float("inf")
    for _ in range(repeats):
        sampling.profile_sampler_begin(kernel)
        run()
        ms, n = sampling.profile_sampler_end()
        if n:
            best = min(best, ms)
    return best


def tune(run, kernels=KERNELS, repeats=2, shapes=SHAPES, waves=WAVES, xcd=XCD, log=None):
    """Times every (tile shape, wavefronts per workgroup, block order) candidate for
    each kernel on `run` and sets the fastest.  Returns {kernel: {"config": (tileW,
    tileH, wavesPerBlock, xcdOrder), "ms": best summed kernel ms per call of run,
    "default_ms": the same for the 8 x 8, 4, 0 default, "candidates": n}}."""
    from cudaExtensions import sampling
    run()                                                 # allocator and code warm
    out = {}
    for kernel in kernels:
        table = []
        for (tw, th) in shapes:
            for wpb in waves:
                for xo in xcd:
                    sampling.set_tile_config(kernel, tw, th, wpb, xo)
                    table.append(((tw, th, wpb, xo), _time(run, kernel, repeats)))
        cfg, ms = min(table, key=lambda r: r[1])
        default = dict(table).get((8, 8, 4, 0))
        sampling.set_tile_config(kernel, *cfg)
        out[kernel] = {"config": list(cfg), "ms": ms, "default_ms": default, "candidates": len(table)}
        if log:
            log("autotune %s: %s %.3f ms (default %.3f ms, %d candidates)" % (kernel, cfg, ms, default or -1,
                                                                             len(table)))
    return out


DEGREES = (4, 5, 6, 7, 8)


def tune_degree(run, degrees=DEGREES, repeats=3, sync=None, log=None):
    """Times whole calls of ``run`` at every expansion order and sets the fastest.
    Returns {"degree": p, "ms": best ms per call, "table": {p: ms}}."""
    from cudaExtensions import sampling
    if sync is None:
        import torch
        sync = torch.cuda.synchronize
    table = {}
    for p in degrees:
        sampling.set_fixed_degree(p)
        run()
        sync()
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            run()
            sync()
            best = min(best, (time.perf_counter() - t0) * 1e3)
        table[p] = best
    p = min(table, key=table.get)
    sampling.set_fixed_degree(p)
    if log:
        log("autotune fixed degree: %d (%s)" % (p, ", ".join("%d: %.3f ms" % kv for kv in sorted(table.items()))))
    return {"degree": p, "ms": table[p], "table": table}


def apply(result):
    """Sets the shapes of a tune() result (or of a cached one)."""
    from cudaExtensions import sampling
    for kernel, r in result.items():
        if kernel == "degree":
            sampling.set_fixed_degree(r["degree"])
        else:
            sampling.set_tile_config(kernel, *r["config"])


def load(path, key):
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get(key)


def save(path, key, result):
    data = {}
    if os.path.exists(path):
        with open(path) as fh:
            data = json.load(fh)
    data[key] = result
    with open(path, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)


def tune_cached(run, path, key, log=None, **kw):
    """tune() unless `path` already holds a result for `key` (then it is applied)."""
    r = load(path, key)
    if r is not None:
        apply(r)
        return r
    r = tune(run, log=log, **kw)
    save(path, key, r)
    return r
