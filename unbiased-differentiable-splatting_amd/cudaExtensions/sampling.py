"""sampling -- FGT-guided per-pixel sequential Poisson sampling.

Mirrors FMMKdTreeSampling/sampling.cpp:48-183 (``sample``, ``covariance``).
The reference seeds cuRAND with time(NULL); here the uniforms come from
Philox4x32-10 keyed by (seed, view, pixel, launch, leaf ordinal).  ``seed``
is an optional keyword; without it a per-process counter advances on every
call (``manual_seed`` resets it), so consecutive frames draw fresh samples.
"""
import torch

from . import _lib

_seed_state = {"next": 0}
last_error = None   # device int32 flag of the last sample() call (non-zero: descent too deep)


def manual_seed(seed):
    _seed_state["next"] = int(seed)


def _next_seed():
    s = _seed_state["next"]
    _seed_state["next"] = s + 1
    return s


def _check_splats(points, normals, stdDevs):
    """points / normals (bn, N, 3), stdDevs (1 | bn, N, 2) (samplingCuda.cu:304-320)."""
    _lib.expect(points, "points", None, None, 3)
    bn, N = points.shape[0], points.shape[1]
    _lib.expect(normals, "normals", bn, N, 3)
    _lib.expect(stdDevs, "stdDevs", (1, bn), N, 2)


def _error_flag(errors, dev):
    """The device word the sampler ORs its error bits into: ``errors`` (an int32
    (1,) tensor the caller keeps, never cleared here) or a fresh zero."""
    if errors is None:
        return torch.zeros(1, dtype=torch.int32, device=dev)
    _lib.require_gpu(errors)
    if errors.dtype != torch.int32 or errors.numel() != 1 or not errors.is_contiguous() or errors.device != dev:
        raise ValueError("errors must be a contiguous int32 tensor of one element on %s" % dev)
    return errors


def raise_errors(e):
    """Raise for the error bits ``e`` of a sampling call (0: nothing)."""
    e = int(e)
    if e & 16:
        raise RuntimeError("sampling.samplePoisson: a tile's rate multiplier was negative or not finite "
                           "(nothing drawn there)")
    if e & 8:
        raise RuntimeError("sampling.samplePoisson: a pixel's tree walk exceeded its node bound")
    if e & 4:
        raise RuntimeError("sampling.samplePoisson: a pixel drew more than numSamples splats (raise "
                           "numSamples or lower the rate)")
    if e != 0:
        raise RuntimeError("sampling.sample: descent exceeded the 63-level trace bound")


def sample(width, height, focallength, points, normals, colors, stdDevs, numSamples, smoothing,
           precision=0.001, seed=None, errors=None):
    """Returns int32 indices (bn, width, height, numSamples); -1 marks an empty slot.
    ``errors``: optional int32 (1,) device tensor the call ORs its error bits into
    (checked by the caller, e.g. with raise_errors(errors.item()))."""
    global last_error
    _check_splats(points, normals, stdDevs)
    dev = _lib.require_gpu(points, normals, stdDevs)
    points = _lib.f32(points, "points")
    normals = _lib.f32(normals, "normals")
    stdDevs = _lib.f32(stdDevs, "stdDevs")
    bn, N, _ = points.shape
    W, H, k = int(width), int(height), int(numSamples)
    if k > 256:
        raise ValueError("numSamples > 256 is not supported")
    out = torch.empty((bn, W, H, k), dtype=torch.int32, device=dev)
    err = _error_flag(errors, dev)
    L = _lib.lib()
    wsb = L.ude_sample_workspace_bytes(N, W, H, k)
    ws = _lib.workspace(wsb, dev)
    s = _next_seed() if seed is None else int(seed)
    rc = L.ude_sample(bn, N, W, H, k, float(focallength), float(smoothing), float(precision),
                      _lib.ptr(points), _lib.ptr(normals), _lib.ptr(stdDevs), stdDevs.shape[0],
                      s & 0xFFFFFFFFFFFFFFFF, _lib.ptr(out), _lib.ptr(err), _lib.ptr(ws), ws.numel(),
                      _lib.stream(dev))
    _lib.check(rc, "sampling.sample")
    last_error = err
    return out


def covariance(width, height, focallength, points, normals, stdDevs, smoothing):
    """Per-splat screen-space covariance (bn, N, 2, 2) (samplingCuda.cu:323-361)."""
    _check_splats(points, normals, stdDevs)
    dev = _lib.require_gpu(points, normals, stdDevs)
    points = _lib.f32(points, "points")
    normals = _lib.f32(normals, "normals")
    stdDevs = _lib.f32(stdDevs, "stdDevs")
    bn, N, _ = points.shape
    out = torch.zeros((bn, N, 2, 2), dtype=torch.float32, device=dev)
    rc = _lib.lib().ude_covariance(bn, N, int(width), int(height), float(focallength), float(smoothing),
                                   _lib.ptr(points), _lib.ptr(normals), _lib.ptr(stdDevs), stdDevs.shape[0],
                                   _lib.ptr(out), _lib.stream(dev))
    _lib.check(rc, "sampling.covariance")
    return out


def approximateFMMWeights(width, height, focallength, points, normals, colors, stdDevs, smoothing, precision):
    """Fixed-order (degree 6) Hermite FGT density image (bn, width, height):
    sum_i exp(-|q - s_i|^2 / 2h^2) / (2 pi sqrt(h)) at every pixel q
    (sampling.cpp:136-175).  ``colors`` is accepted and unused, as in the reference."""
    _check_splats(points, normals, stdDevs)
    dev = _lib.require_gpu(points, normals, stdDevs)
    points = _lib.f32(points, "points")
    normals = _lib.f32(normals, "normals")
    stdDevs = _lib.f32(stdDevs, "stdDevs")
    bn, N, _ = points.shape
    W, H = int(width), int(height)
    out = torch.empty((bn, W, H), dtype=torch.float32, device=dev)
    L = _lib.lib()
    ws = _lib.workspace(L.ude_fmm_weights_workspace_bytes(N, W, H), dev)
    rc = L.ude_fmm_weights(bn, N, W, H, float(focallength), float(smoothing), float(precision),
                           _lib.ptr(points), _lib.ptr(normals), _lib.ptr(stdDevs), stdDevs.shape[0],
                           _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream(dev))
    _lib.check(rc, "sampling.approximateFMMWeights")
    return out


PROPOSALS = {"bands": 0, "footprint": 1}


def samplePoisson(width, height, focallength, points, normals, colors, stdDevs, numSamples, smoothing,
                  precision=0.001, seed=None, rate=None, tileScale=None, densities=None, seedTensor=None,
                  batched=True, errors=None, proposal="bands", footprintScale=None):
    """Poisson-rate sampling with inclusion probabilities (no reference counterpart).

    Every visible splat i is included at pixel q independently with the
    defensive-mixture probability
    pi_i = min(1, r/2 (g0_i / max(D0, c0) + g1_i / max(D1, c1))), r = rate times the
    tile's multiplier, g_b = c_b exp(-|q - s_i|^2 / 2 h_b^2) for the tree bandwidth
    (b = 0) and a wide bandwidth covering the visible splats (b = 1), D_b the
    fixed-order density of bandwidth b; nothing where D0 < precision.  Returns
    (indices int32 (bn, W, H, numSamples), -1 padded; invPi float32, 1 / pi_i per
    slot, 0 padded), so that sum_s weight_s * value_s * invPi_s is an unbiased
    estimate of the sum over all visible splats.  ``rate`` (expected samples per
    pixel, default numSamples / 4) and the optional per-16x16-tile multipliers
    ``tileScale`` (bn, ceil(W/16), ceil(H/16)) set the sampling budget; a pixel
    with more than numSamples inclusions sets an error flag checked by
    check_errors().  ``densities`` (optional float32 (bn, 2, W, H)) receives D0 and
    D1 (the input of tileFeatures).  ``seedTensor`` (optional int64 (1,) on the
    device) holds the Philox key instead of ``seed``; it is read when the kernel
    runs, so a captured graph draws fresh samples after the caller advances it.
    ``batched``: keep every view's tree resident and run each per-pixel pass as
    one launch over all views (bn times the workspace of one view), else view
    by view.  ``errors``: optional int32 (1,) device tensor the call ORs its
    error bits into (bit 4: a pixel had more than numSamples inclusions, the
    extra ones dropped; bit 8: a walk exceeded its node bound; bit 16: a tile's
    rate times its multiplier was negative or not finite, nothing drawn there); it is never
    cleared here, so one flag can collect a whole training run and be checked at
    a point that synchronises anyway (raise_errors(errors.item())).
    ``proposal``: "bands" (the mixture above) or "footprint": the wide band plus
    each splat's own footprint, pi_i = min(1, r (g1_i / (2 N1) + f_i / (2 kw N1)))
    with f_i = min(1, w_i, exp(-max(0, d - 2)^2 / 2h_1^2)) at distance d from the
    splat, w_i its EWA weight (FP32 screen covariance), N1 = max(D1, c1);
    kw = ``footprintScale``
    or, when it is None, the visible splats' mean footprint mass over h_1^1.5
    (each half then expects about r / 2 samples)."""
    global last_error
    _check_splats(points, normals, stdDevs)
    dev = _lib.require_gpu(points, normals, stdDevs)
    points = _lib.f32(points, "points")
    normals = _lib.f32(normals, "normals")
    stdDevs = _lib.f32(stdDevs, "stdDevs")
    bn, N, _ = points.shape
    W, H, k = int(width), int(height), int(numSamples)
    rate = float(k) / 4.0 if rate is None else float(rate)
    ts = None
    if tileScale is not None:
        _lib.require_gpu(tileScale)
        ts = _lib.f32(tileScale, "tileScale")
        if tuple(ts.shape) != (bn, (W + 15) // 16, (H + 15) // 16):
            raise ValueError("tileScale must have shape (bn, ceil(W/16), ceil(H/16))")
    ids = torch.empty((bn, W, H, k), dtype=torch.int32, device=dev)
    ipi = torch.empty((bn, W, H, k), dtype=torch.float32, device=dev)
    err = _error_flag(errors, dev)
    L = _lib.lib()
    ws = _lib.workspace(L.ude_poisson_workspace_bytes(bn if batched else 1, N, W, H), dev)
    s = _next_seed() if seed is None and seedTensor is None else int(seed or 0)
    if seedTensor is not None:
        _lib.require_gpu(seedTensor)
        if seedTensor.dtype != torch.int64 or seedTensor.numel() != 1:
            raise ValueError("seedTensor must be a single int64 on the device")
    if densities is not None:
        _lib.require_gpu(densities)
        if densities.dtype != torch.float32 or tuple(densities.shape) != (bn, 2, W, H) or \
                not densities.is_contiguous():
            raise ValueError("densities must be a contiguous float32 tensor of shape (bn, 2, W, H)")
    if proposal not in PROPOSALS:
        raise ValueError("proposal must be 'bands' or 'footprint'")
    fs = 0.0 if footprintScale is None else float(footprintScale)
    if proposal == "footprint" and footprintScale is not None and not (0.0 < fs < 3.0e38):
        raise ValueError("footprintScale must be a positive finite number")
    _lib.check(L.ude_set_poisson_proposal(PROPOSALS[proposal], fs), "sampling.samplePoisson")
    rc = L.ude_sample_poisson(bn, N, W, H, k, float(focallength), float(smoothing), float(precision), rate,
                              _lib.ptr(ts), _lib.ptr(points), _lib.ptr(normals), _lib.ptr(stdDevs),
                              stdDevs.shape[0], s & 0xFFFFFFFFFFFFFFFF, _lib.ptr(seedTensor), _lib.ptr(ids),
                              _lib.ptr(ipi),
                              _lib.ptr(densities), _lib.ptr(err), _lib.ptr(ws), ws.numel(), _lib.stream(dev))
    _lib.check(rc, "sampling.samplePoisson")
    last_error = err
    return ids, ipi


def tileFeatures(width, height, focallength, points, normals, stdDevs, smoothing, precision, densities,
                 imageGrad=None):
    """32 features per 16x16 pixel tile (bn, ceil(W/16), ceil(H/16), 32) for the rate
    network, from the splats, the densities returned by samplePoisson and
    (optionally) the image gradient (bn, W, H, 3)."""
    _check_splats(points, normals, stdDevs)
    dev = _lib.require_gpu(points, normals, stdDevs, densities)
    points = _lib.f32(points, "points")
    normals = _lib.f32(normals, "normals")
    stdDevs = _lib.f32(stdDevs, "stdDevs")
    densities = _lib.f32(densities, "densities")
    bn, N, _ = points.shape
    W, H = int(width), int(height)
    if tuple(densities.shape) != (bn, 2, W, H):
        raise ValueError("densities must have shape (bn, 2, W, H)")
    g = None
    if imageGrad is not None:
        _lib.require_gpu(imageGrad)
        g = _lib.f32(imageGrad, "imageGrad")
        if tuple(g.shape) != (bn, W, H, 3):
            raise ValueError("imageGrad must have shape (bn, W, H, 3)")
    out = torch.empty((bn, (W + 15) // 16, (H + 15) // 16, 32), dtype=torch.float32, device=dev)
    L = _lib.lib()
    ws = _lib.workspace(L.ude_tile_features_workspace_bytes(W, H), dev)
    rc = L.ude_tile_features(bn, N, W, H, float(focallength), float(smoothing), float(precision), _lib.ptr(points),
                             _lib.ptr(normals), _lib.ptr(stdDevs), stdDevs.shape[0], _lib.ptr(densities), _lib.ptr(g),
                             _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream(dev))
    _lib.check(rc, "sampling.tileFeatures")
    return out


def check_errors():
    """Synchronising check of the last sample() call's device error flag."""
    if last_error is not None:
        raise_errors(last_error.item())


_PROFILED = {"sample": 0, "fmm_eval": 1, "poisson": 2}


def profile_sampler_begin(kernel="sample"):
    """Start bracketing every launch of one sampling kernel with HIP events
    (benchmarks): "sample" (the FGT walk, descents and keyed draws of
    ``sample``), "fmm_eval" (fixed-order density evaluation) or "poisson" (the
    thinning pass of ``samplePoisson``)."""
    _lib.check(_lib.lib().ude_profile_sampler_begin(_PROFILED[kernel]), "profile_sampler_begin")


def set_fixed_degree(p):
    """Order of the fixed-order Hermite expansions behind approximateFMMWeights and
    samplePoisson's densities: 4 .. 8 (default 6).  Every order keeps the
    precision bound; lower orders expand fewer, smaller nodes."""
    _lib.check(_lib.lib().ude_set_fixed_degree(int(p)), "set_fixed_degree")


def get_fixed_degree():
    return int(_lib.lib().ude_get_fixed_degree())


def set_tile_config(kernel, tileW, tileH, wavesPerBlock, xcdOrder=0):
    """Launch shape of a per-pixel pass ("fmm_eval" or "poisson"): one wavefront per
    tileW x tileH pixel tile (tileW * tileH = 64), wavesPerBlock wavefronts per
    workgroup, xcdOrder = 1 for XCD-contiguous tile bands.  Results do not depend
    on it."""
    kk = {"fmm_eval": 0, "poisson": 1}[kernel]
    _lib.check(_lib.lib().ude_set_tile_config(kk, int(tileW), int(tileH), int(wavesPerBlock), int(xcdOrder)),
               "set_tile_config")


def get_tile_config(kernel):
    import ctypes
    kk = {"fmm_eval": 0, "poisson": 1}[kernel]
    out = (ctypes.c_int * 4)()
    _lib.check(_lib.lib().ude_get_tile_config(kk, ctypes.cast(out, ctypes.c_void_p)), "get_tile_config")
    return tuple(out)


def profile_sampler_end():
    """Stop; returns (summed kernel milliseconds, number of launches)."""
    import ctypes
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int(0)
    _lib.check(_lib.lib().ude_profile_sampler_end(ctypes.byref(ms), ctypes.byref(n)), "profile_sampler_end")
    return ms.value, n.value
