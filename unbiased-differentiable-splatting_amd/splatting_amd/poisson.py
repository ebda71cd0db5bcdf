"""Autograd layer of the Horvitz-Thompson (Poisson-rate) renderer (no
counterpart in the reference, whose autograd layer, splatting/extensions/
__init__.py:104-233, stores the sampled index tensor between forward and
backward).

``PoissonRender`` maps camera-space splats to the unbiased image estimate
sum_s w_s a_s / pi_s (EWA weights w, albedo a, inclusion probabilities pi,
no occlusion).  The sample set is a pure function of the splats, the Philox key
and the rates, so with ``recompute=True`` the backward draws it again from the
same key instead of keeping the (bn, W, H, kcap) index and 1/pi tensors alive
between the passes: the forward saves only the per-splat inputs.  Both passes
run the fused estimator (EWA weight, 1/pi and interpolation per sample, no
per-slot tensor).  Gradients are the raw adjoints (no clipping), as
unbiasedness requires.
"""
import torch as th

from cudaExtensions import estimator, sampling

from .extensions import AffineTransform, LinearTransform


def _draw(P, Nc, S, args):
    W, H, f, kcap, prec, rate, seed, ts = args
    nb = th.where(Nc[..., 2:3] > 0, -Nc, Nc).contiguous()
    return sampling.samplePoisson(W, H, f, P, nb, None, S, kcap, 0.0, prec, seed=seed, rate=rate, tileScale=ts)


class PoissonRender(th.autograd.Function):
    @staticmethod
    def forward(ctx, points, normals, stdDevs, albedo, width, height, focallength, numSamples, precision, rate, seed,
                tileScale=None, recompute=True):
        P, Nc, S, A = (x.detach().contiguous() for x in (points, normals, stdDevs, albedo))
        args = (int(width), int(height), float(focallength), int(numSamples), float(precision), float(rate),
                int(seed), tileScale)
        ids, ipi = _draw(P, Nc, S, args)
        img = estimator.forward(args[0], args[1], args[2], ids, ipi, P, Nc, S, A)
        ctx.args = args
        ctx.recompute = bool(recompute)
        if ctx.recompute:
            ctx.save_for_backward(P, Nc, S, A)
        else:
            ctx.save_for_backward(P, Nc, S, A, ids, ipi)
        return img

    @staticmethod
    def backward(ctx, gradImage):
        if ctx.recompute:
            P, Nc, S, A = ctx.saved_tensors
            ids, ipi = _draw(P, Nc, S, ctx.args)
        else:
            P, Nc, S, A, ids, ipi = ctx.saved_tensors
        W, H, f = ctx.args[:3]
        gP, gN, gS, gA = estimator.backward(W, H, f, gradImage.contiguous(), ids, ipi, P, Nc, S, A)
        return gP, gN, gS, gA, None, None, None, None, None, None, None, None, None


def render(transformMatrix, positions, normals, stdDevs, albedo, width, height, focallength, numSamples=40,
           precision=0.01, rate=10.0, seed=0, tileScale=None, recompute=True):
    """World-space splats through the view transforms to the image estimate
    (bn, W, H, 3); differentiable in positions, normals, std-devs, albedo and the
    view matrices."""
    P = AffineTransform.apply(transformMatrix, positions)
    Nc = LinearTransform.apply(transformMatrix, normals)
    return PoissonRender.apply(P, Nc, stdDevs, albedo, width, height, focallength, numSamples, precision, rate,
                               seed, tileScale, recompute)
