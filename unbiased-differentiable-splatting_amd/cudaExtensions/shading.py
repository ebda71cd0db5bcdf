"""shading -- deferred Lambert and second-order spherical-harmonics shading.

Mirrors FMMKdTreeSampling/shading.cpp:51-207.
"""
import torch

from . import _lib


def forward(directions, intensities, normals, diffuse):
    """Lambert: (bn, W, H, 3) = sum_l diffuse/pi * I_l * max(0, d_l . n)."""
    dev = _lib.require_gpu(directions, intensities, normals, diffuse)
    d = _lib.f32(directions, "directions")
    I = _lib.f32(intensities, "intensities")
    n = _lib.f32(normals, "normals")
    a = _lib.f32(diffuse, "diffuse")
    bn, W, H, _ = n.shape
    out = torch.empty((bn, W, H, 3), dtype=torch.float32, device=dev)
    rc = _lib.lib().ude_lambert_forward(bn, W, H, I.shape[1], I.shape[0], _lib.ptr(d), _lib.ptr(I), _lib.ptr(n),
                                        _lib.ptr(a), _lib.ptr(out), _lib.stream(dev))
    _lib.check(rc, "shading.forward")
    return out


def backward(gradIn, directions, intensities, normals, diffuse):
    """[gradDirections, gradIntensities, gradNormals, gradDiffuse]."""
    dev = _lib.require_gpu(gradIn, directions, intensities, normals, diffuse)
    g = _lib.f32(gradIn, "gradIn")
    d = _lib.f32(directions, "directions")
    I = _lib.f32(intensities, "intensities")
    n = _lib.f32(normals, "normals")
    a = _lib.f32(diffuse, "diffuse")
    bn, W, H, _ = n.shape
    gd = torch.zeros(d.shape, dtype=torch.float32, device=dev)
    gI = torch.zeros(I.shape, dtype=torch.float32, device=dev)
    gn = torch.empty(n.shape, dtype=torch.float32, device=dev)
    ga = torch.empty(a.shape, dtype=torch.float32, device=dev)
    rc = _lib.lib().ude_lambert_backward(bn, W, H, I.shape[1], I.shape[0], _lib.ptr(g), _lib.ptr(d), _lib.ptr(I),
                                         _lib.ptr(n), _lib.ptr(a), _lib.ptr(gd), _lib.ptr(gI), _lib.ptr(gn),
                                         _lib.ptr(ga), _lib.stream(dev))
    _lib.check(rc, "shading.backward")
    return [gd, gI, gn, ga]


def forwardSh(shCoefficients, normals, diffuse):
    """(bn, W, H, 3) = diffuse * (n4^T M_c n4) / pi with M_c from 9 SH coefficients per channel."""
    dev = _lib.require_gpu(shCoefficients, normals, diffuse)
    c = _lib.f32(shCoefficients, "shCoefficients")
    n = _lib.f32(normals, "normals")
    a = _lib.f32(diffuse, "diffuse")
    bn, W, H, _ = n.shape
    out = torch.empty((bn, W, H, 3), dtype=torch.float32, device=dev)
    rc = _lib.lib().ude_sh_forward(bn, W, H, c.shape[0], _lib.ptr(c), _lib.ptr(n), _lib.ptr(a), _lib.ptr(out),
                                   _lib.stream(dev))
    _lib.check(rc, "shading.forwardSh")
    return out


def backwardSh(gradIn, shCoefficients, normals, diffuse):
    """[gradShCoefficients, gradNormals, gradDiffuse] (reference conventions, shadingCuda.cu:244-293)."""
    dev = _lib.require_gpu(gradIn, shCoefficients, normals, diffuse)
    g = _lib.f32(gradIn, "gradIn")
    c = _lib.f32(shCoefficients, "shCoefficients")
    n = _lib.f32(normals, "normals")
    a = _lib.f32(diffuse, "diffuse")
    bn, W, H, _ = n.shape
    gc = torch.zeros(c.shape, dtype=torch.float32, device=dev)
    gn = torch.empty(n.shape, dtype=torch.float32, device=dev)
    ga = torch.empty(a.shape, dtype=torch.float32, device=dev)
    rc = _lib.lib().ude_sh_backward(bn, W, H, c.shape[0], _lib.ptr(g), _lib.ptr(c), _lib.ptr(n), _lib.ptr(a),
                                    _lib.ptr(gc), _lib.ptr(gn), _lib.ptr(ga), _lib.stream(dev))
    _lib.check(rc, "shading.backwardSh")
    return [gc, gn, ga]
