"""interpolation -- weighted attribute interpolation over the sampled splats.

Mirrors FMMKdTreeSampling/interpolation.cpp:31-102.  float32 and float64 are
accepted, as the reference dispatches both (interpolationCuda.cu:68, 153).
"""
import torch

from . import _lib


def check_shapes(indices, weights, attribute):
    """indices / weights (bn, W, H, k), attribute (1 | bn, N, C)."""
    _lib.expect(indices, "indices", None, None, None, None)
    _lib.expect(weights, "weights", *indices.shape)
    _lib.expect(attribute, "attribute", (1, indices.shape[0]), None, None)


def forward(indices, weights, attribute):
    """image (bn, W, H, C) = sum_s weights * attribute[indices]."""
    check_shapes(indices, weights, attribute)
    dev = _lib.require_gpu(indices, weights, attribute)
    indices = _lib.i32(indices, "indices")
    dt, (weights, attribute) = _lib.reals([("weights", weights), ("attribute", attribute)])
    bn, W, H, k = indices.shape
    ab, N, C = attribute.shape
    out = torch.empty((bn, W, H, C), dtype=dt, device=dev)
    rc = _lib.fn("ude_interp_forward", dt)(bn, W, H, k, ab, N, C, _lib.ptr(indices), _lib.ptr(weights),
                                           _lib.ptr(attribute), _lib.ptr(out), _lib.stream(dev))
    _lib.check(rc, "interpolation.forward")
    return out


def backward(gradIn, indices, weights, attribute):
    """[gradWeights, gradAttribute]."""
    check_shapes(indices, weights, attribute)
    _lib.expect(gradIn, "gradIn", *indices.shape[:3], attribute.shape[2])
    dev = _lib.require_gpu(gradIn, indices, weights, attribute)
    indices = _lib.i32(indices, "indices")
    dt, (weights, gradIn, attribute) = _lib.reals([("weights", weights), ("gradIn", gradIn),
                                                   ("attribute", attribute)])
    bn, W, H, k = indices.shape
    ab, N, C = attribute.shape
    gW = torch.empty(weights.shape, dtype=dt, device=dev)
    gA = torch.zeros(attribute.shape, dtype=dt, device=dev)
    rc = _lib.fn("ude_interp_backward", dt)(bn, W, H, k, ab, N, C, _lib.ptr(gradIn), _lib.ptr(indices),
                                            _lib.ptr(weights), _lib.ptr(attribute), _lib.ptr(gW), _lib.ptr(gA),
                                            _lib.stream(dev))
    _lib.check(rc, "interpolation.backward")
    return [gW, gA]
