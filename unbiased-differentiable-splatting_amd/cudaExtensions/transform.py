"""transform -- per-point 4x4 affine / linear transforms and their adjoints.

Mirrors FMMKdTreeSampling/transform.cpp:34-94.  float32 and float64 are
accepted, as the reference dispatches both (transformCuda.cu:53, 135).
"""
import torch

from . import _lib


def check_shapes(transformation, vector):
    """transformation (bn, 4, 4), vector (1 | bn, N, 3)."""
    _lib.expect(transformation, "transformation", None, 4, 4)
    _lib.expect(vector, "vector", (1, transformation.shape[0]), None, 3)


def _forward(affine, transformation, vector):
    check_shapes(transformation, vector)
    dev = _lib.require_gpu(transformation, vector)
    dt, (M, v) = _lib.reals([("transformation", transformation), ("vector", vector)])
    bn = M.shape[0]
    vb, N, _ = v.shape
    out = torch.empty((bn, N, 3), dtype=dt, device=dev)
    rc = _lib.fn("ude_transform_forward", dt)(1 if affine else 0, bn, vb, N, _lib.ptr(M), _lib.ptr(v), _lib.ptr(out),
                                              _lib.stream(dev))
    _lib.check(rc, "transform.forward")
    return out


def _backward(affine, gradIn, transformation, vector):
    check_shapes(transformation, vector)
    _lib.expect(gradIn, "gradIn", transformation.shape[0], vector.shape[1], 3)
    dev = _lib.require_gpu(gradIn, transformation, vector)
    dt, (M, g, v) = _lib.reals([("transformation", transformation), ("gradIn", gradIn), ("vector", vector)])
    bn = M.shape[0]
    vb, N, _ = v.shape
    gM = torch.zeros(M.shape, dtype=dt, device=dev)
    # every element is written unless one shared vector set collects several views' sums
    gV = (torch.empty if (vb > 1 or bn == 1) else torch.zeros)(v.shape, dtype=dt, device=dev)
    rc = _lib.fn("ude_transform_backward", dt)(1 if affine else 0, bn, vb, N, _lib.ptr(g), _lib.ptr(M), _lib.ptr(v),
                                               _lib.ptr(gM), _lib.ptr(gV), _lib.stream(dev))
    _lib.check(rc, "transform.backward")
    return [gM, gV]


def affineForward(transformation, vector):
    return _forward(True, transformation, vector)


def affineBackward(gradIn, transformation, vector):
    return _backward(True, gradIn, transformation, vector)


def linearForward(transformation, vector):
    return _forward(False, transformation, vector)


def linearBackward(gradIn, transformation, vector):
    return _backward(False, gradIn, transformation, vector)


def linearForwardFacing(transformation, vector):
    """(linearForward(transformation, vector), its camera-facing copy) in one
    pass: the copy is negated where its z component is positive, as the
    reference's renderer does for camera-space normals (render.py:200-203)."""
    check_shapes(transformation, vector)
    dev = _lib.require_gpu(transformation, vector)
    M = _lib.f32(transformation, "transformation")
    v = _lib.f32(vector, "vector")
    bn = M.shape[0]
    vb, N, _ = v.shape
    out = torch.empty((bn, N, 3), dtype=torch.float32, device=dev)
    facing = torch.empty_like(out)
    rc = _lib.lib().ude_transform_forward_facing(bn, vb, N, _lib.ptr(M), _lib.ptr(v), _lib.ptr(out),
                                                 _lib.ptr(facing), _lib.stream(dev))
    _lib.check(rc, "transform.linearForwardFacing")
    return out, facing
