"""interpolation -- weighted attribute interpolation over the sampled splats.

Mirrors FMMKdTreeSampling/interpolation.cpp:31-102.
"""
import torch

from . import _lib


def forward(indices, weights, attribute):
    """image (bn, W, H, C) = sum_s weights * attribute[indices]."""
    dev = _lib.require_gpu(indices, weights, attribute)
    indices = _lib.i32(indices, "indices")
    weights = _lib.f32(weights, "weights")
    attribute = _lib.f32(attribute, "attribute")
    bn, W, H, k = indices.shape
    ab, N, C = attribute.shape
    out = torch.empty((bn, W, H, C), dtype=torch.float32, device=dev)
    rc = _lib.lib().ude_interp_forward(bn, W, H, k, ab, N, C, _lib.ptr(indices), _lib.ptr(weights),
                                       _lib.ptr(attribute), _lib.ptr(out), _lib.stream(dev))
    _lib.check(rc, "interpolation.forward")
    return out


def backward(gradIn, indices, weights, attribute):
    """[gradWeights, gradAttribute]."""
    dev = _lib.require_gpu(gradIn, indices, weights, attribute)
    gradIn = _lib.f32(gradIn, "gradIn")
    indices = _lib.i32(indices, "indices")
    weights = _lib.f32(weights, "weights")
    attribute = _lib.f32(attribute, "attribute")
    bn, W, H, k = indices.shape
    ab, N, C = attribute.shape
    gW = torch.empty(weights.shape, dtype=torch.float32, device=dev)
    gA = torch.zeros(attribute.shape, dtype=torch.float32, device=dev)
    rc = _lib.lib().ude_interp_backward(bn, W, H, k, ab, N, C, _lib.ptr(gradIn), _lib.ptr(indices),
                                        _lib.ptr(weights), _lib.ptr(attribute), _lib.ptr(gW), _lib.ptr(gA),
                                        _lib.stream(dev))
    _lib.check(rc, "interpolation.backward")
    return [gW, gA]
