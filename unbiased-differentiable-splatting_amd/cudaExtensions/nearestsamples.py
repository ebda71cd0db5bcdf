"""nearestsamples -- per splat, its nearest splats sampled alongside it.

Mirrors FMMKdTreeSampling/nearestSamples.cpp:30-75 (``nearest``, bound as
cudaExtensions.nearestsamples).  float32 and float64 are accepted, as the
reference dispatches both (nearestSamplesCuda.cu:541).
"""
import torch

from . import _lib


def nearest(indices, weights, points, numNearest):
    """(1, N, numNearest) int32: for every splat n, the numNearest nearest distinct
    splats (by distance, ties by index, n itself excluded) among those sampled at
    the pixels whose surface splat -- the first slot of largest weight -- is n;
    -1 where there are fewer.  indices / weights (bn, W, H, k), points (1, N, 3)
    (one point set for all views, as the reference assumes)."""
    _lib.expect(indices, "indices", None, None, None, None)
    _lib.expect(weights, "weights", *indices.shape)
    _lib.expect(points, "points", 1, None, 3)
    K = int(numNearest)
    if K <= 0:
        raise ValueError("numNearest must be positive")
    dev = _lib.require_gpu(indices, weights, points)
    indices = _lib.i32(indices, "indices")
    dt, (weights, points) = _lib.reals([("weights", weights), ("points", points)])
    bn, W, H, k = indices.shape
    N = points.shape[1]
    out = torch.full((1, N, K), -1, dtype=torch.int32, device=dev)
    if N == 0 or indices.numel() == 0:
        return out
    L = _lib.lib()
    ws = _lib.workspace(L.ude_nearest_workspace_bytes(N, K), dev)
    rc = _lib.fn("ude_nearest_samples", dt)(bn, W, H, k, N, K, _lib.ptr(indices), _lib.ptr(weights),
                                            _lib.ptr(points), _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                                            _lib.stream(dev))
    _lib.check(rc, "nearestsamples.nearest")
    return out
