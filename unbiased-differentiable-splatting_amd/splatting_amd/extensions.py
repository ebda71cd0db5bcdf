"""Autograd wrappers around the ``cudaExtensions`` op surface.

Mirrors splatting/extensions/__init__.py (FmmSampling :104-117, AffineTransform /
LinearTransform :124-164, EWASplatting :174-233, DepthFiltering :372-400,
PixelScatterGather :407-438, LinearInterpolation :445-484, DiffuseShading /
SphericalHarmonicShading :524-581).

``EWASplatting`` has two gradient modes:
  * ``"reference"`` (default) reproduces the reference's post-processing: row-wise
    clipping to norm <= 1 (conditionalNormalization), NaN -> 0, and no gradient
    for the normals (extensions/__init__.py:203-233);
  * ``"raw"`` returns the exact adjoint, including the normals' gradient, with
    no clipping, and composites with the exact compositing adjoint (the
    reference's drops terms where |w| or |1 - w| <= 1e-4) -- needed for
    unbiasedness and finite-difference checks.
"""
import torch as th

from cudaExtensions import sampling, splatting, filtering, transform, interpolation, shading, pixelgather

from .utils import conditionalNormalization

GRAD_MODE = {"ewa": "reference"}


def set_ewa_gradient_mode(mode):
    if mode not in ("reference", "raw"):
        raise ValueError(mode)
    GRAD_MODE["ewa"] = mode


class FmmSampling:
    @th.no_grad()
    def __call__(self, width, height, focallength, points, normals, colors, stdDevs, numSamples, smoothing,
                 precision=0.001, seed=None):
        return sampling.sample(width, height, focallength, points, normals, colors, stdDevs, numSamples, smoothing,
                               precision, seed=seed)


class AffineTransform(th.autograd.Function):
    @staticmethod
    def forward(ctx, transformMatrix, vector):
        ctx.save_for_backward(transformMatrix, vector)
        return transform.affineForward(transformMatrix, vector)

    @staticmethod
    def backward(ctx, gradIn):
        M, v = ctx.saved_tensors
        gM, gV = transform.affineBackward(gradIn.contiguous(), M, v)
        return gM, gV


class LinearTransform(th.autograd.Function):
    @staticmethod
    def forward(ctx, transformMatrix, vector):
        ctx.save_for_backward(transformMatrix, vector)
        return transform.linearForward(transformMatrix, vector)

    @staticmethod
    def backward(ctx, gradIn):
        M, v = ctx.saved_tensors
        gM, gV = transform.linearBackward(gradIn.contiguous(), M, v)
        return gM, gV


class LinearTransformFacing(th.autograd.Function):
    """LinearTransform and the camera-facing copy of its result (negated where
    its z component is positive, no gradient) in one pass: the reference's
    linearForward followed by its backface torch.where (render.py:193-205)."""

    @staticmethod
    def forward(ctx, transformMatrix, vector):
        ctx.save_for_backward(transformMatrix, vector)
        out, facing = transform.linearForwardFacing(transformMatrix, vector)
        ctx.mark_non_differentiable(facing)
        return out, facing

    @staticmethod
    def backward(ctx, gradIn, gradFacing):
        M, v = ctx.saved_tensors
        gM, gV = transform.linearBackward(gradIn.contiguous(), M, v)
        return gM, gV


class EWASplatting(th.autograd.Function):
    @staticmethod
    def forward(ctx, width, height, focallength, indices, points, normals, stdDevs, smoothing):
        weights = splatting.forward(width, height, focallength, indices, points, normals, stdDevs, smoothing)
        ctx.save_for_backward(indices, points, normals, stdDevs)
        ctx.width, ctx.height, ctx.focallength, ctx.smoothing = width, height, focallength, smoothing
        return weights

    @staticmethod
    def backward(ctx, gradWeights):
        indices, points, normals, stdDevs = ctx.saved_tensors
        if GRAD_MODE["ewa"] != "raw" and stdDevs.shape[0] == indices.shape[0]:
            # the reference's post-processing (extensions/__init__.py:219-225), fused
            gp, _, gs = splatting.backward(ctx.width, ctx.height, ctx.focallength, gradWeights.contiguous(), indices,
                                           points, normals, stdDevs, ctx.smoothing, normalized=True)
            return None, None, None, None, gp, None, gs, None
        gp, gn, gs = splatting.backward(ctx.width, ctx.height, ctx.focallength, gradWeights.contiguous(), indices,
                                        points, normals, stdDevs, ctx.smoothing)
        if GRAD_MODE["ewa"] == "raw":
            return None, None, None, None, gp, gn, gs, None
        gp = conditionalNormalization(gp)
        gs = conditionalNormalization(gs)
        gp = th.nan_to_num(gp, nan=0.0, posinf=float("inf"), neginf=float("-inf"))
        gs = th.nan_to_num(gs, nan=0.0, posinf=float("inf"), neginf=float("-inf"))
        return None, None, None, None, gp, None, gs, None


class DepthFiltering(th.autograd.Function):
    @staticmethod
    def forward(ctx, indices, weights, width, height, focallength, points, normals):
        ctx.save_for_backward(indices, weights, points, normals)
        ctx.width, ctx.height, ctx.focallength = width, height, focallength
        return filtering.forward(indices, weights, width, height, focallength, points, normals)

    @staticmethod
    def backward(ctx, gradIn):
        indices, weights, points, normals = ctx.saved_tensors
        gW = filtering.backward(indices, weights, gradIn.contiguous(), ctx.width, ctx.height, ctx.focallength,
                                points, normals, exact=GRAD_MODE["ewa"] == "raw")
        return None, gW, None, None, None, None, None


class PixelScatterGather(th.autograd.Function):
    """values (bn, N, C) -> their rows per sampled slot (bn, W, H, k, C); the
    backward sums the slots' gradients per splat (extensions/__init__.py:407-438)."""

    @staticmethod
    def forward(ctx, indices, values):
        ctx.save_for_backward(indices)
        ctx.numSources = values.shape[1]
        return pixelgather.scatter(indices, values)

    @staticmethod
    def backward(ctx, grad_output):
        indices, = ctx.saved_tensors
        return None, pixelgather.gather(indices, grad_output.contiguous(), ctx.numSources)


class LinearInterpolation(th.autograd.Function):
    @staticmethod
    def forward(ctx, indices, weights, values):
        ctx.save_for_backward(indices, weights, values)
        return interpolation.forward(indices, weights, values)

    @staticmethod
    def backward(ctx, gradIn):
        indices, weights, values = ctx.saved_tensors
        gW, gV = interpolation.backward(gradIn.contiguous(), indices, weights, values)
        return None, gW, gV


class DiffuseShading(th.autograd.Function):
    @staticmethod
    def forward(ctx, directions, intensities, normals, diffuse):
        ctx.save_for_backward(directions, intensities, normals, diffuse)
        return shading.forward(directions, intensities, normals, diffuse)

    @staticmethod
    def backward(ctx, gradIn):
        d, I, n, a = ctx.saved_tensors
        gd, gI, gn, ga = shading.backward(gradIn.contiguous(), d, I, n, a)
        return gd, gI, gn, ga


class SphericalHarmonicShading(th.autograd.Function):
    @staticmethod
    def forward(ctx, coefficients, normals, diffuse):
        ctx.save_for_backward(coefficients, normals, diffuse)
        return shading.forwardSh(coefficients, normals, diffuse)

    @staticmethod
    def backward(ctx, gradIn):
        c, n, a = ctx.saved_tensors
        gc, gn, ga = shading.backwardSh(gradIn.contiguous(), c, n, a)
        return gc, gn, ga
