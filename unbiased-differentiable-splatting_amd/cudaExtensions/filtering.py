"""filtering -- per-pixel z-sort and front-to-back alpha compositing.

Mirrors FMMKdTreeSampling/zSorting.cpp:24-48.  float32, float64 and float16
weights / points are accepted, as the reference dispatches all three
(zSortingCuda.cu:100, 220); half tensors are composited in float32.
"""
import torch

from . import _lib


_ALLOWED = (torch.float32, torch.float64, torch.float16)


def check_shapes(indices, weights, width, height, points, normals):
    """indices / weights (bn, width, height, k), points / normals (1 | bn, N, 3)."""
    _lib.expect(indices, "indices", None, int(width), int(height), None)
    _lib.expect(weights, "weights", *indices.shape)
    _lib.expect(points, "points", (1, indices.shape[0]), None, 3)
    _lib.expect(normals, "normals", *points.shape)


def forward(indices, weights, width, height, focallength, points, normals):
    check_shapes(indices, weights, width, height, points, normals)
    dev = _lib.require_gpu(indices, weights, points, normals)
    indices = _lib.i32(indices, "indices")
    dt, (weights, points, normals) = _lib.reals([("weights", weights), ("points", points), ("normals", normals)],
                                                _ALLOWED)
    bn, W, H, k = indices.shape
    if k > 1024:
        raise ValueError("filtering supports at most 1024 samples per pixel")
    out = torch.empty(weights.shape, dtype=dt, device=dev)
    rc = _lib.fn("ude_filter_forward", dt)(bn, points.shape[1], int(width), int(height), k, float(focallength),
                                           _lib.ptr(indices), _lib.ptr(weights), _lib.ptr(points), _lib.ptr(normals),
                                           points.shape[0], _lib.ptr(out), _lib.stream(dev))
    _lib.check(rc, "filtering.forward")
    return out


def backward(indices, weights, gradIn, width, height, focallength, points, normals, exact=False):
    """Gradient w.r.t. the weights.  By default the reference's adjoint, which drops
    a slot's own term where |w| <= 1e-4 and its occlusion term where |1 - w| <= 1e-4
    (alphacompositingGrad.cuh:7-26); exact=True keeps every term (the exact
    derivative; float32 and float64)."""
    check_shapes(indices, weights, width, height, points, normals)
    _lib.expect(gradIn, "gradIn", *indices.shape)
    dev = _lib.require_gpu(indices, weights, gradIn, points, normals)
    indices = _lib.i32(indices, "indices")
    dt, (weights, gradIn, points, normals) = _lib.reals([("weights", weights), ("gradIn", gradIn),
                                                         ("points", points), ("normals", normals)], _ALLOWED)
    bn, W, H, k = indices.shape
    if k > 1024:
        raise ValueError("filtering supports at most 1024 samples per pixel")
    if exact and dt == torch.float16:
        raise TypeError("filtering.backward: exact=True takes float32 or float64 tensors")
    out = torch.empty(weights.shape, dtype=dt, device=dev)
    rc = _lib.fn("ude_filter_backward_exact" if exact else "ude_filter_backward", dt)(bn, points.shape[1], int(width), int(height), k, float(focallength),
                                            _lib.ptr(indices), _lib.ptr(weights), _lib.ptr(gradIn), _lib.ptr(points),
                                            _lib.ptr(normals), points.shape[0], _lib.ptr(out), _lib.stream(dev))
    _lib.check(rc, "filtering.backward")
    return out
