"""estimator -- the Horvitz-Thompson image estimate over Poisson-rate samples
and its adjoint, each fused into one per-pixel pass (no reference counterpart:
it composes splatting.forward, the 1 / pi weighting and interpolation.forward
of a three-channel attribute, splatting.cpp:43-82 and interpolation.cpp:31-58,
with their backward ops).

``indices`` / ``invPi`` are sampling.samplePoisson's outputs: per pixel its
samples first, then -1 (0) padding; a pixel's slots after its first -1 are not
read.  The image equals

    interpolation.forward(indices, splatting.forward(...) * invPi, attribute)

bit for bit; the gradients equal the unfused backward ops' up to the order of
their atomic sums.  No per-slot tensor is stored in either pass.
"""
import torch

from . import _lib


def _inputs(indices, invPi, points, normals, stdDevs, attribute):
    indices = _lib.i32(indices, "indices")
    invPi = _lib.f32(invPi, "invPi")
    points = _lib.f32(points, "points")
    normals = _lib.f32(normals, "normals")
    stdDevs = _lib.f32(stdDevs, "stdDevs")
    attribute = _lib.f32(attribute, "attribute")
    if invPi.shape != indices.shape:
        raise ValueError("invPi %s must match indices %s" % (tuple(invPi.shape), tuple(indices.shape)))
    if attribute.dim() != 3 or attribute.shape[2] != 3:
        raise ValueError("attribute must be (1 | bn, N, 3), got %s" % (tuple(attribute.shape),))
    return indices, invPi, points, normals, stdDevs, attribute


def workspace(indices, stdDevs):
    """A workspace for forward and backward of these shapes; passing the same one to
    forward and then to backward(..., prepared=True) reuses the per-splat records."""
    bn = indices.shape[0]
    return _lib.workspace(_lib.lib().ude_ht_workspace_bytes(bn, stdDevs.shape[1]), indices.device)


def _ws(ws, bn, N, dev):
    need = _lib.lib().ude_ht_workspace_bytes(bn, N)
    if ws is None:
        return _lib.workspace(need, dev)
    if ws.dtype != torch.uint8 or ws.numel() < need or ws.device != dev:
        raise ValueError("workspace must be a uint8 tensor of at least %d bytes on %s" % (need, dev))
    return ws


def forward(width, height, focallength, indices, invPi, points, normals, stdDevs, attribute, smoothing=0.0,
            workspace=None):
    """image (bn, W, H, 3) = sum over a pixel's samples of min(w, 1) / pi * attribute."""
    dev = _lib.require_gpu(indices, invPi, points, normals, stdDevs, attribute)
    indices, invPi, points, normals, stdDevs, attribute = _inputs(indices, invPi, points, normals, stdDevs,
                                                                  attribute)
    bn, W, H, k = indices.shape
    N = stdDevs.shape[1]
    out = torch.empty((bn, W, H, 3), dtype=torch.float32, device=dev)
    L = _lib.lib()
    ws = _ws(workspace, bn, N, dev)
    rc = L.ude_ht_forward(bn, N, W, H, k, float(focallength), float(smoothing), _lib.ptr(indices), _lib.ptr(invPi),
                          _lib.ptr(points), _lib.ptr(normals), _lib.ptr(stdDevs), stdDevs.shape[0],
                          attribute.shape[0], _lib.ptr(attribute), _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                          _lib.stream(dev))
    _lib.check(rc, "estimator.forward")
    return out


def backward(width, height, focallength, gradImage, indices, invPi, points, normals, stdDevs, attribute,
             smoothing=0.0, workspace=None, prepared=False):
    """[gradPoints, gradNormals, gradStdDevs, gradAttribute] of sum(gradImage * image).
    prepared=True: `workspace` is the one forward used on the same splats."""
    dev = _lib.require_gpu(gradImage, indices, invPi, points, normals, stdDevs, attribute)
    gradImage = _lib.f32(gradImage, "gradImage")
    indices, invPi, points, normals, stdDevs, attribute = _inputs(indices, invPi, points, normals, stdDevs,
                                                                  attribute)
    bn, W, H, k = indices.shape
    if tuple(gradImage.shape) != (bn, W, H, 3):
        raise ValueError("gradImage must be %s, got %s" % ((bn, W, H, 3), tuple(gradImage.shape)))
    N = stdDevs.shape[1]
    gp = torch.empty(points.shape, dtype=torch.float32, device=dev)
    gn = torch.empty(normals.shape, dtype=torch.float32, device=dev)
    gs = (torch.empty if stdDevs.shape[0] == bn else torch.zeros)(stdDevs.shape, dtype=torch.float32, device=dev)
    ga = (torch.empty if attribute.shape[0] == bn else torch.zeros)(attribute.shape, dtype=torch.float32,
                                                                    device=dev)
    if prepared and workspace is None:
        raise ValueError("prepared=True needs the workspace forward used")
    L = _lib.lib()
    ws = _ws(workspace, bn, N, dev)
    rc = L.ude_ht_backward(bn, N, W, H, k, float(focallength), float(smoothing), _lib.ptr(gradImage),
                           _lib.ptr(indices), _lib.ptr(invPi), _lib.ptr(points), _lib.ptr(normals),
                           _lib.ptr(stdDevs), stdDevs.shape[0], attribute.shape[0], _lib.ptr(attribute),
                           _lib.ptr(gp), _lib.ptr(gn), _lib.ptr(gs), _lib.ptr(ga), int(bool(prepared)), _lib.ptr(ws),
                           ws.numel(), _lib.stream(dev))
    _lib.check(rc, "estimator.backward")
    return [gp, gn, gs, ga]


def l2(image, target):
    """(loss, grad) of the L2 image loss sum (image - target)^2: grad = 2 (image -
    target) in float32, loss a float64 0-d tensor summed in a fixed order."""
    dev = _lib.require_gpu(image, target)
    image = _lib.f32(image, "image")
    target = _lib.f32(target, "target")
    if image.shape != target.shape:
        raise ValueError("target %s must match image %s" % (tuple(target.shape), tuple(image.shape)))
    grad = torch.empty_like(image)
    sums = torch.empty(1 + 1024, dtype=torch.float64, device=dev)
    rc = _lib.lib().ude_l2_grad(image.numel(), _lib.ptr(image), _lib.ptr(target), _lib.ptr(grad), _lib.ptr(sums),
                                _lib.stream(dev))
    _lib.check(rc, "estimator.l2")
    return sums[0], grad
