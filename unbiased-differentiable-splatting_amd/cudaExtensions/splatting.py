"""splatting -- EWA weights per (pixel, sampled splat) and their adjoint.

Mirrors FMMKdTreeSampling/splatting.cpp:43-136.  float32 and float64 tensors
are accepted, as the reference dispatches both (splattingCuda.cu:84, 214).
"""
import torch

from . import _lib


def check_shapes(width, height, indices, points, normals, stdDevs):
    """indices (bn, width, height, k), points / normals (bn, N, 3), stdDevs (1 | bn, N, 2)
    (the reference indexes points by view and stdDevs by view only when their batch is
    not 1, splattingCuda.cu:57-61)."""
    _lib.expect(indices, "indices", None, int(width), int(height), None)
    bn = indices.shape[0]
    _lib.expect(stdDevs, "stdDevs", (1, bn), None, 2)
    N = stdDevs.shape[1]
    _lib.expect(points, "points", bn, N, 3)
    _lib.expect(normals, "normals", bn, N, 3)


def _workspace(dtype, bn, N, dev):
    L = _lib.lib()
    nbytes = L.ude_splat_workspace_bytes_f64(bn, N) if dtype == torch.float64 else L.ude_splat_workspace_bytes(bn, N)
    return _lib.workspace(nbytes, dev)


def forward(width, height, focallength, indices, points, normals, stdDevs, smoothing):
    """weights (bn, W, H, k): min(w, 1) per slot, 0 where indices < 0."""
    check_shapes(width, height, indices, points, normals, stdDevs)
    dev = _lib.require_gpu(indices, points, normals, stdDevs)
    indices = _lib.i32(indices, "indices")
    dt, (points, normals, stdDevs) = _lib.reals([("points", points), ("normals", normals), ("stdDevs", stdDevs)])
    bn, W, H, k = indices.shape
    N = stdDevs.shape[1]
    out = torch.empty(indices.shape, dtype=dt, device=dev)
    ws = _workspace(dt, bn, N, dev)
    rc = _lib.fn("ude_splat_forward", dt)(bn, N, W, H, k, float(focallength), float(smoothing), _lib.ptr(indices),
                                          _lib.ptr(points), _lib.ptr(normals), _lib.ptr(stdDevs), stdDevs.shape[0],
                                          _lib.ptr(out), _lib.ptr(ws), ws.numel(), _lib.stream(dev))
    _lib.check(rc, "splatting.forward")
    return out


def backward(width, height, focallength, gradIn, indices, points, normals, stdDevs, smoothing, normalized=False):
    """[gradPoints, gradNormals, gradStdDevs] of sum(gradIn * weights).  With
    ``normalized`` (per-view std-devs only) the point and std-dev gradients come
    back post-processed as the reference's EWASplatting.backward does
    (conditionalNormalization, then NaN -> 0), in the same pass."""
    check_shapes(width, height, indices, points, normals, stdDevs)
    _lib.expect(gradIn, "gradIn", *indices.shape)
    dev = _lib.require_gpu(gradIn, indices, points, normals, stdDevs)
    indices = _lib.i32(indices, "indices")
    dt, (points, gradIn, normals, stdDevs) = _lib.reals([("points", points), ("gradIn", gradIn),
                                                         ("normals", normals), ("stdDevs", stdDevs)])
    bn, W, H, k = indices.shape
    N = stdDevs.shape[1]
    gp = torch.empty(points.shape, dtype=dt, device=dev)
    gn = torch.empty(normals.shape, dtype=dt, device=dev)
    gs = (torch.empty if stdDevs.shape[0] == bn else torch.zeros)(stdDevs.shape, dtype=dt, device=dev)
    ws = _workspace(dt, bn, N, dev)
    if normalized:
        if stdDevs.shape[0] != bn:
            raise ValueError("splatting.backward(normalized=True) needs per-view stdDevs (bn, N, 2)")
        rc = _lib.fn("ude_splat_backward_normalized", dt)(
            bn, N, W, H, k, float(focallength), float(smoothing), _lib.ptr(gradIn), _lib.ptr(indices),
            _lib.ptr(points), _lib.ptr(normals), _lib.ptr(stdDevs), _lib.ptr(gp), _lib.ptr(gn), _lib.ptr(gs),
            _lib.ptr(ws), ws.numel(), _lib.stream(dev))
        _lib.check(rc, "splatting.backward")
        return [gp, gn, gs]
    rc = _lib.fn("ude_splat_backward", dt)(bn, N, W, H, k, float(focallength), float(smoothing), _lib.ptr(gradIn),
                                           _lib.ptr(indices), _lib.ptr(points), _lib.ptr(normals), _lib.ptr(stdDevs),
                                           stdDevs.shape[0], _lib.ptr(gp), _lib.ptr(gn), _lib.ptr(gs), _lib.ptr(ws),
                                           ws.numel(), _lib.stream(dev))
    _lib.check(rc, "splatting.backward")
    return [gp, gn, gs]
