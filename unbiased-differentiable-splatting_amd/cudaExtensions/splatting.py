"""splatting -- EWA weights per (pixel, sampled splat) and their adjoint.

Mirrors FMMKdTreeSampling/splatting.cpp:43-136.
"""
import torch

from . import _lib


def forward(width, height, focallength, indices, points, normals, stdDevs, smoothing):
    """weights (bn, W, H, k): min(w, 1) per slot, 0 where indices < 0."""
    dev = _lib.require_gpu(indices, points, normals, stdDevs)
    indices = _lib.i32(indices, "indices")
    points = _lib.f32(points, "points")
    normals = _lib.f32(normals, "normals")
    stdDevs = _lib.f32(stdDevs, "stdDevs")
    bn, W, H, k = indices.shape
    N = stdDevs.shape[1]
    out = torch.empty(indices.shape, dtype=torch.float32, device=dev)
    L = _lib.lib()
    ws = _lib.workspace(L.ude_splat_workspace_bytes(bn, N), dev)
    rc = L.ude_splat_forward(bn, N, W, H, k, float(focallength), float(smoothing), _lib.ptr(indices),
                             _lib.ptr(points), _lib.ptr(normals), _lib.ptr(stdDevs), stdDevs.shape[0],
                             _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream(dev))
    _lib.check(rc, "splatting.forward")
    return out


def backward(width, height, focallength, gradIn, indices, points, normals, stdDevs, smoothing):
    """[gradPoints, gradNormals, gradStdDevs] of sum(gradIn * weights)."""
    dev = _lib.require_gpu(gradIn, indices, points, normals, stdDevs)
    gradIn = _lib.f32(gradIn, "gradIn")
    indices = _lib.i32(indices, "indices")
    points = _lib.f32(points, "points")
    normals = _lib.f32(normals, "normals")
    stdDevs = _lib.f32(stdDevs, "stdDevs")
    bn, W, H, k = indices.shape
    N = stdDevs.shape[1]
    gp = torch.empty(points.shape, dtype=torch.float32, device=dev)
    gn = torch.empty(normals.shape, dtype=torch.float32, device=dev)
    gs = (torch.empty if stdDevs.shape[0] == bn else torch.zeros)(stdDevs.shape, dtype=torch.float32, device=dev)
    L = _lib.lib()
    ws = _lib.workspace(L.ude_splat_workspace_bytes(bn, N), dev)
    rc = L.ude_splat_backward(bn, N, W, H, k, float(focallength), float(smoothing), _lib.ptr(gradIn),
                              _lib.ptr(indices), _lib.ptr(points), _lib.ptr(normals), _lib.ptr(stdDevs),
                              stdDevs.shape[0], _lib.ptr(gp), _lib.ptr(gn), _lib.ptr(gs), _lib.ptr(ws), ws.numel(),
                              _lib.stream(dev))
    _lib.check(rc, "splatting.backward")
    return [gp, gn, gs]
