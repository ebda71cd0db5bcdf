"""filtering -- per-pixel z-sort and front-to-back alpha compositing.

Mirrors FMMKdTreeSampling/zSorting.cpp:24-48.
"""
import torch

from . import _lib


def forward(indices, weights, width, height, focallength, points, normals):
    dev = _lib.require_gpu(indices, weights, points, normals)
    indices = _lib.i32(indices, "indices")
    weights = _lib.f32(weights, "weights")
    points = _lib.f32(points, "points")
    normals = _lib.f32(normals, "normals")
    bn, W, H, k = indices.shape
    if k > 1024:
        raise ValueError("filtering supports at most 1024 samples per pixel")
    out = torch.empty(weights.shape, dtype=torch.float32, device=dev)
    rc = _lib.lib().ude_filter_forward(bn, points.shape[1], int(width), int(height), k, float(focallength),
                                       _lib.ptr(indices), _lib.ptr(weights), _lib.ptr(points), _lib.ptr(normals),
                                       points.shape[0], _lib.ptr(out), _lib.stream(dev))
    _lib.check(rc, "filtering.forward")
    return out


def backward(indices, weights, gradIn, width, height, focallength, points, normals):
    dev = _lib.require_gpu(indices, weights, gradIn, points, normals)
    indices = _lib.i32(indices, "indices")
    weights = _lib.f32(weights, "weights")
    gradIn = _lib.f32(gradIn, "gradIn")
    points = _lib.f32(points, "points")
    normals = _lib.f32(normals, "normals")
    bn, W, H, k = indices.shape
    if k > 1024:
        raise ValueError("filtering supports at most 1024 samples per pixel")
    out = torch.empty(weights.shape, dtype=torch.float32, device=dev)
    rc = _lib.lib().ude_filter_backward(bn, points.shape[1], int(width), int(height), k, float(focallength),
                                        _lib.ptr(indices), _lib.ptr(weights), _lib.ptr(gradIn), _lib.ptr(points),
                                        _lib.ptr(normals), points.shape[0], _lib.ptr(out), _lib.stream(dev))
    _lib.check(rc, "filtering.backward")
    return out
