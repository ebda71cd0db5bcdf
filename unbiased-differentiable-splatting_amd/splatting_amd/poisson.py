"""Autograd layer of the Horvitz-Thompson (Poisson-rate) renderer (no
counterpart in the reference, whose autograd layer, splatting/extensions/
__init__.py:104-233, stores the sampled index tensor between forward and
backward).

``PoissonRender`` maps camera-space splats to the unbiased image estimate
sum_s w_s a_s / pi_s (EWA weights w, albedo a, inclusion probabilities pi,
no occlusion), or with ``composite=True`` to the front-to-back composite of the
samples in ray-depth order with alpha_s = w_s / pi_s, an unbiased estimate of
the image composited over every splat (z-sort and alpha compositing as the
reference's filtering op, zSortingCuda.cu:11-243).  The sample set is a pure function of the splats, the Philox key
and the rates, so with ``recompute=True`` the backward draws it again from the
same key instead of keeping the (bn, W, H, kcap) index and 1/pi tensors alive
between the passes: the forward saves only the per-splat inputs.  Both passes
run the fused estimator (EWA weight, 1/pi and interpolation per sample, no
per-slot tensor).  Gradients are the raw adjoints (no clipping), as
unbiasedness requires.

Without an explicit ``seed`` every forward takes the next key of sampling's
per-process counter (as sampling.sample does; the reference seeds cuRAND with
time(NULL)), so consecutive frames draw independent sample sets; the key is kept
for the backward, which redraws the identical set.  A pixel with more inclusions
than ``numSamples`` loses the extra ones, so the forward raises when the draw
reports it (outside graph capture, where it can synchronise; inside a capture
the flag stays on the device for the caller: ``PoissonRender.last_errors``).
"""
import torch as th

from cudaExtensions import estimator, sampling

from .extensions import AffineTransform, LinearTransform


def _draw(P, Nc, S, args, errors):
    W, H, f, kcap, prec, rate, seed, ts = args
    nb = th.where(Nc[..., 2:3] > 0, -Nc, Nc).contiguous()
    return sampling.samplePoisson(W, H, f, P, nb, None, S, kcap, 0.0, prec, seed=seed, rate=rate, tileScale=ts,
                                  errors=errors)


class PoissonRender(th.autograd.Function):
    last_errors = None      # the device error flag of the latest forward draw

    @staticmethod
    def forward(ctx, points, normals, stdDevs, albedo, width, height, focallength, numSamples, precision, rate, seed,
                tileScale=None, recompute=True, composite=False):
        P, Nc, S, A = (x.detach().contiguous() for x in (points, normals, stdDevs, albedo))
        key = sampling._next_seed() if seed is None else int(seed)
        args = (int(width), int(height), float(focallength), int(numSamples), float(precision), float(rate),
                key, tileScale)
        err = th.zeros(1, dtype=th.int32, device=P.device)
        ids, ipi = _draw(P, Nc, S, args, err)
        PoissonRender.last_errors = err
        if not th.cuda.is_current_stream_capturing():
            sampling.raise_errors(err.item())
        img = estimator.forward(args[0], args[1], args[2], ids, ipi, P, Nc, S, A, composite=composite)
        ctx.args = args
        ctx.composite = bool(composite)
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
            ids, ipi = _draw(P, Nc, S, ctx.args, None)      # the forward's draw, which was checked
        else:
            P, Nc, S, A, ids, ipi = ctx.saved_tensors
        W, H, f = ctx.args[:3]
        gP, gN, gS, gA = estimator.backward(W, H, f, gradImage.contiguous(), ids, ipi, P, Nc, S, A,
                                            composite=ctx.composite)
        return gP, gN, gS, gA, None, None, None, None, None, None, None, None, None, None


def render(transformMatrix, positions, normals, stdDevs, albedo, width, height, focallength, numSamples=40,
           precision=0.01, rate=10.0, seed=None, tileScale=None, recompute=True, composite=False):
    """World-space splats through the view transforms to the image estimate
    (bn, W, H, 3); differentiable in positions, normals, std-devs, albedo and the
    view matrices.  ``seed`` None: the next key of the per-process counter."""
    P = AffineTransform.apply(transformMatrix, positions)
    Nc = LinearTransform.apply(transformMatrix, normals)
    return PoissonRender.apply(P, Nc, stdDevs, albedo, width, height, focallength, numSamples, precision, rate,
                               seed, tileScale, recompute, composite)
