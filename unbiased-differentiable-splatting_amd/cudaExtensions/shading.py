"""shading -- deferred Lambert and second-order spherical-harmonics shading.

Mirrors FMMKdTreeSampling/shading.cpp:51-207.  float32 and float64 are
accepted, as the reference dispatches both (shadingCuda.cu:63, 153, 225, 315).
"""
import torch

from . import _lib


def check_lambert(directions, intensities, normals, diffuse):
    """normals / diffuse (bn, W, H, 3), directions (bn, L, 3), intensities (1 | bn, L, 3)."""
    _lib.expect(normals, "normals", None, None, None, 3)
    _lib.expect(diffuse, "diffuse", *normals.shape)
    bn = normals.shape[0]
    _lib.expect(directions, "directions", bn, None, 3)
    _lib.expect(intensities, "intensities", (1, bn), directions.shape[1], 3)


def check_sh(shCoefficients, normals, diffuse):
    """normals / diffuse (bn, W, H, 3), shCoefficients (1 | bn, 3, 9)."""
    _lib.expect(normals, "normals", None, None, None, 3)
    _lib.expect(diffuse, "diffuse", *normals.shape)
    _lib.expect(shCoefficients, "shCoefficients", (1, normals.shape[0]), 3, 9)


def forward(directions, intensities, normals, diffuse):
    """Lambert: (bn, W, H, 3) = sum_l diffuse/pi * I_l * max(0, d_l . n)."""
    check_lambert(directions, intensities, normals, diffuse)
    dev = _lib.require_gpu(directions, intensities, normals, diffuse)
    dt, (n, d, I, a) = _lib.reals([("normals", normals), ("directions", directions), ("intensities", intensities),
                                   ("diffuse", diffuse)])
    bn, W, H, _ = n.shape
    out = torch.empty((bn, W, H, 3), dtype=dt, device=dev)
    rc = _lib.fn("ude_lambert_forward", dt)(bn, W, H, I.shape[1], I.shape[0], _lib.ptr(d), _lib.ptr(I), _lib.ptr(n),
                                            _lib.ptr(a), _lib.ptr(out), _lib.stream(dev))
    _lib.check(rc, "shading.forward")
    return out


def backward(gradIn, directions, intensities, normals, diffuse):
    """[gradDirections, gradIntensities, gradNormals, gradDiffuse]."""
    check_lambert(directions, intensities, normals, diffuse)
    _lib.expect(gradIn, "gradIn", *normals.shape)
    dev = _lib.require_gpu(gradIn, directions, intensities, normals, diffuse)
    dt, (n, g, d, I, a) = _lib.reals([("normals", normals), ("gradIn", gradIn), ("directions", directions),
                                      ("intensities", intensities), ("diffuse", diffuse)])
    bn, W, H, _ = n.shape
    gd = torch.zeros(d.shape, dtype=dt, device=dev)
    gI = torch.zeros(I.shape, dtype=dt, device=dev)
    gn = torch.empty(n.shape, dtype=dt, device=dev)
    ga = torch.empty(a.shape, dtype=dt, device=dev)
    rc = _lib.fn("ude_lambert_backward", dt)(bn, W, H, I.shape[1], I.shape[0], _lib.ptr(g), _lib.ptr(d), _lib.ptr(I),
                                             _lib.ptr(n), _lib.ptr(a), _lib.ptr(gd), _lib.ptr(gI), _lib.ptr(gn),
                                             _lib.ptr(ga), _lib.stream(dev))
    _lib.check(rc, "shading.backward")
    return [gd, gI, gn, ga]


def forwardSh(shCoefficients, normals, diffuse):
    """(bn, W, H, 3) = diffuse * (n4^T M_c n4) / pi with M_c from 9 SH coefficients per channel."""
    check_sh(shCoefficients, normals, diffuse)
    dev = _lib.require_gpu(shCoefficients, normals, diffuse)
    dt, (n, c, a) = _lib.reals([("normals", normals), ("shCoefficients", shCoefficients), ("diffuse", diffuse)])
    bn, W, H, _ = n.shape
    out = torch.empty((bn, W, H, 3), dtype=dt, device=dev)
    rc = _lib.fn("ude_sh_forward", dt)(bn, W, H, c.shape[0], _lib.ptr(c), _lib.ptr(n), _lib.ptr(a), _lib.ptr(out),
                                       _lib.stream(dev))
    _lib.check(rc, "shading.forwardSh")
    return out


def backwardSh(gradIn, shCoefficients, normals, diffuse):
    """[gradShCoefficients, gradNormals, gradDiffuse] (reference conventions, shadingCuda.cu:244-293)."""
    check_sh(shCoefficients, normals, diffuse)
    _lib.expect(gradIn, "gradIn", *normals.shape)
    dev = _lib.require_gpu(gradIn, shCoefficients, normals, diffuse)
    dt, (n, g, c, a) = _lib.reals([("normals", normals), ("gradIn", gradIn), ("shCoefficients", shCoefficients),
                                   ("diffuse", diffuse)])
    bn, W, H, _ = n.shape
    gc = torch.zeros(c.shape, dtype=dt, device=dev)
    gn = torch.empty(n.shape, dtype=dt, device=dev)
    ga = torch.empty(a.shape, dtype=dt, device=dev)
    rc = _lib.fn("ude_sh_backward", dt)(bn, W, H, c.shape[0], _lib.ptr(g), _lib.ptr(c), _lib.ptr(n), _lib.ptr(a),
                                        _lib.ptr(gc), _lib.ptr(gn), _lib.ptr(ga), _lib.stream(dev))
    _lib.check(rc, "shading.backwardSh")
    return [gc, gn, ga]
