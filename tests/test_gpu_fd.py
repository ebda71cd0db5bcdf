"""The hand-written adjoint of the Horvitz-Thompson frame on the GPU against
FP64 central differences (step 1e-4) on 32 random parameters: 256 splats of
the sphere generator at 64x64, one Poisson-rate sample set (indices and 1/pi
held fixed, as the estimator treats them), loss sum_q G_q . image_q with a
random G.  The differenced forward is the oracle's FP64 EWA weight and
interpolation, so the check is independent of the GPU forward.  A parameter whose
step moves a slot across the weight clamp min(w, 1) is passed over (the loss has
a kink there)."""
import numpy as np
import pytest
import torch

from helpers import scene_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_adjoint_matches_central_differences(dev):
    from cudaExtensions import interpolation, sampling, splatting
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=6.0)
    W, H, f = sc.width, sc.height, sc.focal
    rng = np.random.default_rng(17)
    A = rng.uniform(0.1, 1.0, size=(1, p.shape[0], 3)).astype(np.float32)
    G = rng.normal(size=(1, W, H, 3)).astype(np.float32)
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    P, Nc, NB, S = t(p[None]), t(n[None]), t(nb[None]), t(sd[None])
    ids, ipi = sampling.samplePoisson(W, H, f, P, NB, None, S, 48, 0.0, 0.01, seed=4, rate=6.0)
    w = splatting.forward(W, H, f, ids, P, Nc, S, 0.0)
    wt = (w * ipi).contiguous()
    gW, gA = interpolation.backward(t(G), ids, wt, t(A))
    gP, gN, gS = splatting.backward(W, H, f, (gW * ipi).contiguous(), ids, P, Nc, S, 0.0)
    grads = [g[0].cpu().numpy().astype(np.float64) for g in (gP, gN, gS, gA)]
    idn, ipn = ids.cpu().numpy(), ipi.cpu().numpy().astype(np.float64)
    params = [p[None].astype(np.float64), n[None].astype(np.float64), sd[None].astype(np.float64),
              A.astype(np.float64)]

    def loss(prm):
        wf = O.splat_forward(W, H, f, idn, prm[0], prm[1], prm[2])
        img = O.interp_forward(idn, wf * ipn, prm[3])
        return float((G.astype(np.float64) * img).sum()), wf >= 1.0

    drawn = np.unique(idn[idn >= 0])
    assert drawn.size > 50
    h = 1e-4
    checked = 0
    for q in range(200):
        if checked == 32:
            break
        g = int(rng.integers(0, 4))
        sid = int(rng.choice(drawn))
        c = int(rng.integers(0, params[g].shape[-1]))
        x0 = params[g][0, sid, c]
        params[g][0, sid, c] = x0 + h
        fp, cp = loss(params)
        params[g][0, sid, c] = x0 - h
        fm, cm = loss(params)
        params[g][0, sid, c] = x0
        if not np.array_equal(cp, cm):
            continue            # the step crosses the clamp min(w, 1) of a slot: no derivative there
        fd = (fp - fm) / (2 * h)
        got = grads[g][sid, c]
        scale = max(abs(fd), abs(got), 1e-3 * np.abs(grads[g]).max())
        assert abs(got - fd) <= 2e-3 * scale, (g, sid, c, got, fd)
        checked += 1
    assert checked == 32
