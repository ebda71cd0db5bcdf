"""The reference's hand-written EWA gradient chain (gradEWAWeight,
ewaSplattingGrad.cuh:9-206, contracted in the index order of
linalg.cuh:1174-1340; restated in oracle/ewa_reference_adjoint.py) against the
oracle's exact forward-mode derivatives of the same forward, slot by slot:
weights and the point / normal / std-dev gradients agree to 1e-9 of each
slot's largest component on every unclamped slot of the 256-splat 64x64 sphere
workload (12 slots per pixel mixing near and random splats), with and without
smoothing.  The oracle's derivative is what the GPU adjoint is held to, so the
GPU adjoint is held to the reference's gradient as well."""
import numpy as np
import pytest

from helpers import near_slots, scene_inputs
from oracle import oracle as O
from oracle.ewa_reference_adjoint import ewa_gradient


@pytest.mark.parametrize("smoothing", [0.0, 0.5])
def test_reference_chain_equals_exact_derivative(smoothing):
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    W, H, f = sc.width, sc.height, sc.focal
    idx = near_slots(np.random.default_rng(7), p, W, H, f, 12)
    sl = np.argwhere(idx[0] >= 0)
    ii, jj, ss = sl[:, 0], sl[:, 1], sl[:, 2]
    ids = idx[0][ii, jj, ss]
    M = len(ids)
    P, Nn, S = (a[ids].astype(np.float64) for a in (p, n, sd))
    w, gp, gn, gs = ewa_gradient(ii.astype(np.float64), jj.astype(np.float64), W, H, f, P, Nn, S, smoothing)
    # the oracle per slot: every slot gets its own copy of its splat
    own = np.full((1, W, H, idx.shape[3]), -1, np.int32)
    own[0, ii, jj, ss] = np.arange(M)
    gin = np.zeros(own.shape)
    gin[0, ii, jj, ss] = 1.0
    rp, rn, rs = O.splat_backward(W, H, f, gin, own, P[None], Nn[None], S[None], smoothing)
    wo = O.splat_forward(W, H, f, own, P[None], Nn[None], S[None], smoothing)[0, ii, jj, ss]
    m = w <= 1.0                       # the reference's gradient gate (splattingCuda.cu:173)
    assert m.sum() > 30_000
    assert np.all(np.abs(w[m] - wo[m]) <= 1e-9 * wo[m] + 1e-300)
    for name, a, b in (("points", gp, rp[0]), ("normals", gn, rn[0]), ("std-devs", gs, rs[0])):
        a, b = a[m], b[m]
        scale = np.abs(b).max(1, keepdims=True)
        err = np.abs(a - b)
        assert np.all(err <= 1e-9 * scale + 1e-300), (name, float((err / (scale + 1e-300)).max()))
        # the gradients are far from trivially small where the weight is not
        big = wo[m] > 1e-3
        assert np.median(scale[big]) > 1e-3
