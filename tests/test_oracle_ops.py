"""Pins the oracle's per-slot operators (FP64) by closed forms, known answers,
central differences and torch FP64 autograd:

  EWA weight / covariance   ewaSplatting.cuh:8-91, splattingCuda.cu:18-66, samplingCuda.cu:323-361
  EWA adjoint               ewaSplattingGrad.cuh:163-206, splattingCuda.cu:114-234
  alpha compositing         alphacompositing.cuh:6-68, zSortingCuda.cu:11-62
  compositing adjoint       alphacompositingGrad.cuh:7-26, zSortingCuda.cu:126-178
  transform                 transformCuda.cu:11-117, transformationGrad.cuh:5-40
  interpolation             interpolationCuda.cu:5-129
  SH / Lambert shading      shadingCuda.cu:12-293, diffuseShading.cuh, diffuseShadingGrad.cuh
"""
import math

import numpy as np
import pytest
import torch

from helpers import scene_inputs, near_slots
from oracle import oracle as O


# ------------------------------------------------------------------ EWA known answers
@pytest.mark.parametrize("nz", [1.0, -1.0])
def test_ewa_fronto_parallel_known_answer(nz):
    """A fronto-parallel splat at depth z with isotropic std sigma projects to the
    screen Gaussian N(centre, (f sigma / z)^2 I) times the area factor (z / f)^2:
    w(d) = exp(-|d|^2 z^2 / (2 f^2 sigma^2)) / (2 pi sigma^2)."""
    W = H = 16
    f, z, sig = 10.0, 5.0, 0.5
    pts = np.array([[[0.0, 0.0, -z]]])
    nrm = np.array([[[0.0, 0.0, nz]]])
    sd = np.array([[[sig, sig]]])
    idx = np.zeros((1, W, H, 1), np.int32)
    w = O.splat_forward(W, H, f, idx, pts, nrm, sd)[0, :, :, 0]
    ii, jj = np.meshgrid(np.arange(W), np.arange(H), indexing="ij")
    d2 = (ii - W / 2.0) ** 2 + (jj - H / 2.0) ** 2
    ref = np.exp(-d2 * z * z / (2 * f * f * sig * sig)) / (2 * math.pi * sig * sig)
    assert np.abs(w - ref).max() < 1e-12
    V = O.covariance(W, H, f, pts, nrm, sd)[0, 0]
    assert np.allclose(V, (f * sig / z) ** 2 * np.eye(2), rtol=1e-12, atol=1e-14)
    # smoothing adds s I to the covariance
    V = O.covariance(W, H, f, pts, nrm, sd, smoothing=0.3)[0, 0]
    assert np.allclose(V, ((f * sig / z) ** 2 + 0.3) * np.eye(2), rtol=1e-12)


def test_ewa_anisotropic_fronto_parallel():
    """sd = (s0, s1) on a fronto-parallel plane: V = (f/z)^2 diag(s0^2, s1^2)
    (tangent u0 along screen x, u1 = u0 x n along -y)."""
    W = H = 8
    f, z = 20.0, 4.0
    pts = np.array([[[0.0, 0.0, -z]]])
    nrm = np.array([[[0.0, 0.0, 1.0]]])
    sd = np.array([[[0.3, 0.7]]])
    V = O.covariance(W, H, f, pts, nrm, sd)[0, 0]
    assert np.allclose(V, (f / z) ** 2 * np.diag([0.09, 0.49]), rtol=1e-12, atol=1e-12)


def test_ewa_clamp_and_empty_slots():
    W = H = 8
    pts = np.array([[[0.0, 0.0, -5.0]]])
    nrm = np.array([[[0.0, 0.0, 1.0]]])
    sd = np.array([[[0.05, 0.05]]])            # peak 1 / (2 pi 0.0025) > 1: clamped
    idx = np.zeros((1, W, H, 2), np.int32)
    idx[..., 1] = -1
    w = O.splat_forward(W, H, 10.0, idx, pts, nrm, sd)
    assert w[0, 4, 4, 0] == 1.0 and np.all(w[..., 1] == 0.0)
    g = np.ones_like(w)
    gp, gn, gs = O.splat_backward(W, H, 10.0, g, idx, pts, nrm, sd)
    # the clamped centre slot contributes nothing; the rest is the tail's gradient
    idx2 = idx.copy()
    idx2[0, 4, 4, 0] = -1
    gp2, gn2, gs2 = O.splat_backward(W, H, 10.0, g, idx2, pts, nrm, sd)
    assert np.array_equal(gp, gp2) and np.array_equal(gs, gs2)


# ------------------------------------------------------------------ EWA adjoint vs FD
def _fd_case():
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    rng = np.random.default_rng(11)
    idx = near_slots(rng, p, sc.width, sc.height, sc.focal, 6)
    return sc, p[None].astype(np.float64), nb[None].astype(np.float64), sd[None].astype(np.float64), idx, rng


def test_ewa_backward_matches_central_differences():
    sc, P, Nn, S, idx, rng = _fd_case()
    W, H, f = sc.width, sc.height, sc.focal
    gin = rng.normal(size=idx.shape)
    gp, gn, gs = O.splat_backward(W, H, f, gin, idx, P, Nn, S)
    ids = rng.choice(np.unique(idx[idx >= 0]), size=6, replace=False)
    eps = 1e-6
    for sid in ids:
        for arr, grad, ncomp in ((P, gp, 3), (Nn, gn, 3), (S, gs, 2)):
            for c in range(ncomp):
                a0 = arr[0, sid, c]
                arr[0, sid, c] = a0 + eps
                fp = (gin * O.splat_forward(W, H, f, idx, P, Nn, S)).sum()
                arr[0, sid, c] = a0 - eps
                fm = (gin * O.splat_forward(W, H, f, idx, P, Nn, S)).sum()
                arr[0, sid, c] = a0
                fd = (fp - fm) / (2 * eps)
                scale = max(abs(fd), np.abs(grad[0, sid]).max(), 1e-3)
                assert abs(grad[0, sid, c] - fd) <= 2e-5 * scale, (sid, c, grad[0, sid, c], fd)


def test_ewa_normalisation_sums_to_one():
    """Summed over a fine pixel grid, w |det A| integrates the screen Gaussian:
    sum_pixels w ~ 1 / |det A| ~ (f / z)^2 for a fronto-parallel splat."""
    W = H = 96
    f, z = 30.0, 3.0
    pts = np.array([[[0.0, 0.0, -z]]])
    nrm = np.array([[[0.0, 0.0, 1.0]]])
    sd = np.array([[[0.8, 0.5]]])             # screen std 8 and 5 pixels, peak < 1
    idx = np.zeros((1, W, H, 1), np.int32)
    w = O.splat_forward(W, H, f, idx, pts, nrm, sd)
    assert w.max() < 1.0
    assert abs(w.sum() / (f / z) ** 2 - 1.0) < 1e-7


# ------------------------------------------------------------------ compositing
def test_compositing_closed_form_and_order():
    """Slots blend front to back by ray depth: out_front = w0, out_back = w1 (1 - w0),
    whatever their slot order; empty slots sort last."""
    W = H = 4
    f = 4.0
    pts = np.array([[[0.0, 0.0, -3.0], [0.0, 0.0, -2.0]]], np.float32)
    nrm = np.array([[[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]]], np.float32)
    idx = np.tile(np.array([0, -1, 1], np.int32), (1, W, H, 1))
    w = np.tile(np.array([0.6, 0.9, 0.3]), (1, W, H, 1))
    out = O.filter_forward(idx, w, W, H, f, pts, nrm)
    assert np.allclose(out[..., 2], 0.3) and np.allclose(out[..., 0], 0.6 * 0.7)
    assert np.allclose(out[..., 1], 0.9 * 0.4 * 0.7)


def test_compositing_backward_matches_central_differences():
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    W, H, f = 12, 10, sc.focal * 12 / sc.width
    rng = np.random.default_rng(2)
    k = 7
    idx = rng.integers(0, p.shape[0], size=(1, W, H, k)).astype(np.int32)
    idx[rng.uniform(size=idx.shape) < 0.15] = -1
    w = rng.uniform(0.05, 0.95, size=idx.shape)
    gin = rng.normal(size=idx.shape)
    pts, nrm = p[None], nb[None]
    g = O.filter_backward(idx, w, gin, W, H, f, pts, nrm)
    eps = 1e-6
    fd = np.zeros_like(w)
    for s in range(k):
        wp = w.copy(); wp[..., s] += eps
        wm = w.copy(); wm[..., s] -= eps
        fd[..., s] = ((gin * O.filter_forward(idx, wp, W, H, f, pts, nrm)).sum(-1) -
                      (gin * O.filter_forward(idx, wm, W, H, f, pts, nrm)).sum(-1)) / (2 * eps)
    assert np.abs(g - fd).max() <= 1e-6 * max(np.abs(fd).max(), 1.0)


def test_exact_compositing_backward_at_gated_weights():
    """The exact adjoint (no dropped terms) against central differences where the
    reference's gates act: weights at 1e-5 (|w| <= 1e-4 drops the slot's own term)
    and within 1e-5 of 1 (|1 - w| <= 1e-4 drops its occlusion term), among ordinary
    ones.  Where no gate acts the two adjoints agree; where one does they differ."""
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    W, H, f = 12, 10, sc.focal * 12 / sc.width
    rng = np.random.default_rng(4)
    k = 7
    idx = rng.integers(0, p.shape[0], size=(1, W, H, k)).astype(np.int32)
    w = rng.uniform(0.05, 0.95, size=idx.shape)
    w[..., 2] = 1e-5
    w[..., 4] = 1.0 - 1e-5
    gin = rng.normal(size=idx.shape)
    pts, nrm = p[None], nb[None]
    ge = O.filter_backward(idx, w, gin, W, H, f, pts, nrm, exact=True)
    gr = O.filter_backward(idx, w, gin, W, H, f, pts, nrm)
    eps = 1e-7
    fd = np.zeros_like(w)
    for s in range(k):
        wp = w.copy(); wp[..., s] += eps
        wm = w.copy(); wm[..., s] -= eps
        fd[..., s] = ((gin * O.filter_forward(idx, wp, W, H, f, pts, nrm)).sum(-1) -
                      (gin * O.filter_forward(idx, wm, W, H, f, pts, nrm)).sum(-1)) / (2 * eps)
    assert np.abs(ge - fd).max() <= 1e-6 * max(np.abs(fd).max(), 1.0)
    assert np.abs(ge[..., 2] - gr[..., 2]).max() > 1e-3 and np.abs(ge[..., 4] - gr[..., 4]).max() > 1e-3
    # with every weight away from 0 and 1 the two adjoints are the same derivative
    w2 = rng.uniform(0.05, 0.95, size=idx.shape)
    assert np.allclose(O.filter_backward(idx, w2, gin, W, H, f, pts, nrm, exact=True),
                       O.filter_backward(idx, w2, gin, W, H, f, pts, nrm), rtol=1e-12, atol=1e-14)


# ------------------------------------------------------------------ transform
@pytest.mark.parametrize("affine", [True, False])
def test_transform_vs_autograd(affine):
    rng = np.random.default_rng(3)
    bn, N = 3, 50
    M = rng.normal(size=(bn, 4, 4))
    v = rng.normal(size=(1, N, 3))
    out = O.transform_forward(affine, M, v)
    Mt = torch.tensor(M, requires_grad=True)
    vt = torch.tensor(v, requires_grad=True)
    ref = vt @ Mt[:, :3, :3].transpose(1, 2) + (Mt[:, None, :3, 3] if affine else 0.0)
    assert np.allclose(out, ref.detach().numpy(), rtol=1e-13, atol=1e-13)
    g = rng.normal(size=out.shape)
    (ref * torch.tensor(g)).sum().backward()
    gM, gV = O.transform_backward(affine, g, M, v)
    assert np.allclose(gM, Mt.grad.numpy(), rtol=1e-12, atol=1e-12)
    assert np.allclose(gV, vt.grad.numpy(), rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------------ interpolation
def test_interpolation_vs_autograd():
    rng = np.random.default_rng(4)
    bn, W, H, k, N, C = 2, 6, 5, 4, 30, 3
    idx = rng.integers(-1, N, size=(bn, W, H, k)).astype(np.int32)
    w = rng.uniform(size=idx.shape)
    attr = rng.normal(size=(bn, N, C))
    out = O.interp_forward(idx, w, attr)
    wt = torch.tensor(w, requires_grad=True)
    at = torch.tensor(attr, requires_grad=True)
    valid = torch.tensor(idx >= 0)
    gath = at[torch.arange(bn)[:, None, None, None], torch.tensor(np.maximum(idx, 0)).long()]
    ref = ((wt * valid)[..., None] * gath).sum(3)
    assert np.allclose(out, ref.detach().numpy(), rtol=1e-13, atol=1e-13)
    g = rng.normal(size=out.shape)
    (ref * torch.tensor(g)).sum().backward()
    gW, gA = O.interp_backward(g, idx, w, attr)
    assert np.allclose(gW, (wt.grad * valid).numpy(), atol=1e-12)
    assert np.allclose(gA, at.grad.numpy(), atol=1e-12)


# ------------------------------------------------------------------ shading
C1, C2, C3, C4, C5 = 0.429043, 0.511664, 0.743125, 0.886227, 0.247708


def sh_irradiance(L, n):
    """Ramamoorthi & Hanrahan (2001) eq. 13, coefficients ordered
    L00, L1-1, L10, L11, L2-2, L2-1, L20, L21, L22."""
    x, y, z = n[..., 0], n[..., 1], n[..., 2]
    return (C1 * L[8] * (x * x - y * y) + C3 * L[6] * z * z + C4 * L[0] - C5 * L[6]
            + 2 * C1 * (L[4] * x * y + L[7] * x * z + L[5] * y * z)
            + 2 * C2 * (L[3] * x + L[1] * y + L[2] * z))


def test_sh_forward_closed_form():
    rng = np.random.default_rng(5)
    bn, W, H = 2, 5, 4
    coef = rng.normal(size=(3 * bn, 9))
    nrm = rng.normal(size=(bn, W, H, 3))
    nrm[0, 0, 0] = 0.0                      # degenerate normal -> n4 = (0, 0, 0, 1)
    diff = rng.uniform(size=(bn, W, H, 3))
    out = O.sh_forward(coef, nrm, diff)
    nt = torch.tensor(nrm)
    nn = nt.norm(dim=-1, keepdim=True)
    nh = torch.where(nn > 1e-5, nt / nn, torch.zeros_like(nt))
    for b in range(bn):
        for ch in range(3):
            E = sh_irradiance(torch.tensor(coef[3 * b + ch]), nh[b]).numpy()
            assert np.allclose(out[b, ..., ch], diff[b, ..., ch] * E / math.pi, rtol=1e-12, atol=1e-12)


def test_sh_backward_vs_autograd_with_channel_pairing():
    """Coefficient and albedo gradients are the true ones; the normal gradient
    keeps the reference's pairing of colour channel i with normal component i
    (shadingCuda.cu:283): gN[..., i] = d(g_i out_i) / d n_i."""
    rng = np.random.default_rng(6)
    W, H = 6, 3
    coef = rng.normal(size=(3, 9))
    nrm = rng.normal(size=(1, W, H, 3))
    diff = rng.uniform(size=(1, W, H, 3))
    g = rng.normal(size=(1, W, H, 3))
    gc, gn, gd = O.sh_backward(g, coef, nrm, diff)
    ct = torch.tensor(coef, requires_grad=True)
    dt = torch.tensor(diff, requires_grad=True)
    gN = np.zeros_like(nrm)
    tot = 0.0
    for ch in range(3):
        nt = torch.tensor(nrm, requires_grad=True)
        nh = nt / nt.norm(dim=-1, keepdim=True)
        o = dt[..., ch] * sh_irradiance(ct[ch], nh) / math.pi
        s = (o * torch.tensor(g[..., ch])).sum()
        gn_ch, = torch.autograd.grad(s, nt, retain_graph=True)
        gN[..., ch] = gn_ch[..., ch].numpy()
        tot = tot + s
    tot.backward()
    assert np.allclose(gc, ct.grad.numpy(), rtol=1e-11, atol=1e-11)
    assert np.allclose(gd, dt.grad.numpy(), rtol=1e-11, atol=1e-11)
    assert np.allclose(gn, gN, rtol=1e-11, atol=1e-11)


def test_lambert_vs_autograd():
    rng = np.random.default_rng(7)
    bn, W, H, L = 2, 4, 5, 3
    dirs = rng.normal(size=(bn, L, 3))
    ints = rng.uniform(size=(1, L, 3))
    nrm = rng.normal(size=(bn, W, H, 3))
    diff = rng.uniform(size=(bn, W, H, 3))
    out = O.lambert_forward(dirs, ints, nrm, diff)
    dt = torch.tensor(dirs, requires_grad=True)
    it = torch.tensor(ints, requires_grad=True)
    nt = torch.tensor(nrm, requires_grad=True)
    ft = torch.tensor(diff, requires_grad=True)
    cos = torch.einsum("bld,bwhd->bwhl", dt, nt).clamp(min=0.0)
    ref = torch.einsum("bwhl,lc->bwhc", cos, it[0]) * ft / math.pi
    assert np.allclose(out, ref.detach().numpy(), rtol=1e-12, atol=1e-12)
    g = rng.normal(size=out.shape)
    (ref * torch.tensor(g)).sum().backward()
    gdir, gint, gn, gd = O.lambert_backward(g, dirs, ints, nrm, diff)
    for got, ref_g in ((gdir, dt.grad), (gint, it.grad), (gn, nt.grad), (gd, ft.grad)):
        assert np.allclose(got, ref_g.numpy(), rtol=1e-11, atol=1e-11)
