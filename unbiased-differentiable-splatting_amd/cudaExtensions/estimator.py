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

With ``composite=True`` the samples of a pixel are alpha-composited front to back
in the ray-depth order of filtering.forward (zSorting.cpp:24-48), each with alpha
min(w, 1) / pi: the Poisson-rate counterpart of

    interpolation.forward(indices, filtering.forward(indices, w / pi, ...), attribute)

and, since inclusions are independent, an unbiased estimate of the image
composited over every splat of the sampling support (prod (1 - alpha_t) has
mean prod (1 - w_t)).  Its backward is the exact adjoint (no dropped terms).
"""
import torch

from . import _lib


def check_shapes(width, height, indices, invPi, points, normals, stdDevs, attribute):
    """indices / invPi (bn, width, height, k), points / normals (bn, N, 3), stdDevs
    (1 | bn, N, 2), attribute (1 | bn, N, 3): the shapes the kernels index by."""
    _lib.expect(indices, "indices", None, int(width), int(height), None)
    _lib.expect(invPi, "invPi", *indices.shape)
    bn = indices.shape[0]
    _lib.expect(stdDevs, "stdDevs", (1, bn), None, 2)
    N = stdDevs.shape[1]
    _lib.expect(points, "points", bn, N, 3)
    _lib.expect(normals, "normals", bn, N, 3)
    _lib.expect(attribute, "attribute", (1, bn), N, 3)


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


MAX_COMPOSITE_SAMPLES = 112     # slots per pixel of the composited estimate


def _ws_bytes(shape, N, composite):
    bn, W, H, k = shape
    L = _lib.lib()
    return L.ude_htc_workspace_bytes(bn, N, W, H, k) if composite else L.ude_ht_workspace_bytes(bn, N)


def workspace(indices, stdDevs, composite=False):
    """A workspace for forward and backward of these shapes; passing the same one to
    forward and then to backward(..., prepared=True) reuses the per-splat records."""
    return _lib.workspace(_ws_bytes(tuple(indices.shape), stdDevs.shape[1], composite), indices.device)


def _ws(ws, shape, N, dev, composite):
    need = _ws_bytes(shape, N, composite)
    if ws is None:
        return _lib.workspace(need, dev)
    if ws.dtype != torch.uint8 or ws.numel() < need or ws.device != dev:
        raise ValueError("workspace must be a uint8 tensor of at least %d bytes on %s" % (need, dev))
    return ws


def forward(width, height, focallength, indices, invPi, points, normals, stdDevs, attribute, smoothing=0.0,
            workspace=None, composite=False):
    """image (bn, W, H, 3) = sum over a pixel's samples of min(w, 1) / pi * attribute;
    composite=True: front-to-back compositing of alpha = min(w, 1) / pi in ray-depth order."""
    check_shapes(width, height, indices, invPi, points, normals, stdDevs, attribute)
    dev = _lib.require_gpu(indices, invPi, points, normals, stdDevs, attribute)
    indices, invPi, points, normals, stdDevs, attribute = _inputs(indices, invPi, points, normals, stdDevs,
                                                                  attribute)
    bn, W, H, k = indices.shape
    if composite and k > MAX_COMPOSITE_SAMPLES:
        raise ValueError("the composited estimate takes at most %d slots per pixel" % MAX_COMPOSITE_SAMPLES)
    N = stdDevs.shape[1]
    out = torch.empty((bn, W, H, 3), dtype=torch.float32, device=dev)
    L = _lib.lib()
    ws = _ws(workspace, tuple(indices.shape), N, dev, composite)
    fn = L.ude_htc_forward if composite else L.ude_ht_forward
    rc = fn(bn, N, W, H, k, float(focallength), float(smoothing), _lib.ptr(indices), _lib.ptr(invPi),
            _lib.ptr(points), _lib.ptr(normals), _lib.ptr(stdDevs), stdDevs.shape[0],
            attribute.shape[0], _lib.ptr(attribute), _lib.ptr(out), _lib.ptr(ws), ws.numel(),
            _lib.stream(dev))
    _lib.check(rc, "estimator.forward")
    return out


def backward(width, height, focallength, gradImage, indices, invPi, points, normals, stdDevs, attribute,
             smoothing=0.0, workspace=None, prepared=False, composite=False):
    """[gradPoints, gradNormals, gradStdDevs, gradAttribute] of sum(gradImage * image).
    prepared=True: `workspace` is the one forward used on the same splats.
    composite: as in forward (the depth order held fixed)."""
    check_shapes(width, height, indices, invPi, points, normals, stdDevs, attribute)
    _lib.expect(gradImage, "gradImage", indices.shape[0], int(width), int(height), 3)
    dev = _lib.require_gpu(gradImage, indices, invPi, points, normals, stdDevs, attribute)
    gradImage = _lib.f32(gradImage, "gradImage")
    indices, invPi, points, normals, stdDevs, attribute = _inputs(indices, invPi, points, normals, stdDevs,
                                                                  attribute)
    bn, W, H, k = indices.shape
    if composite and k > MAX_COMPOSITE_SAMPLES:
        raise ValueError("the composited estimate takes at most %d slots per pixel" % MAX_COMPOSITE_SAMPLES)
    N = stdDevs.shape[1]
    gp = torch.empty(points.shape, dtype=torch.float32, device=dev)
    gn = torch.empty(normals.shape, dtype=torch.float32, device=dev)
    gs = (torch.empty if stdDevs.shape[0] == bn else torch.zeros)(stdDevs.shape, dtype=torch.float32, device=dev)
    ga = (torch.empty if attribute.shape[0] == bn else torch.zeros)(attribute.shape, dtype=torch.float32,
                                                                    device=dev)
    if prepared and workspace is None:
        raise ValueError("prepared=True needs the workspace forward used")
    L = _lib.lib()
    ws = _ws(workspace, tuple(indices.shape), N, dev, composite)
    fn = L.ude_htc_backward if composite else L.ude_ht_backward
    rc = fn(bn, N, W, H, k, float(focallength), float(smoothing), _lib.ptr(gradImage),
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
