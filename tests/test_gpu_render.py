"""End-to-end GPU parity of the drop-in renderer (splatting_amd.render.Splatting,
reference: splatting/render.py:268-376 with the autograd layer of
splatting/extensions/__init__.py:104-484) on the 4,096-splat unit-sphere
workload at 128x128 (focal 153.6, camera at distance 2.5, k = 40, precision
0.01, shading "None"), with an L2 sum loss against a target rendered from centres
perturbed by N(0, 0.01^2), in both gradient modes:

  * "reference": the EWA gradients of points and std-devs clipped row-wise to
    norm <= 1 with NaN -> 0, and no gradient of the normals
    (extensions/__init__.py:203-233);
  * "raw": the exact adjoints, the normals' included, and the exact compositing
    adjoint (the reference's drops terms where |w| or |1 - w| <= 1e-4).

The sampled indices must equal the oracle's sampler on the same camera-space
inputs bit for bit; the image and the world-space gradients of positions,
normals, std-devs and albedo must match the oracle's FP64 chain (transform ->
EWA -> z-sort compositing -> interpolation -> L2 and back) elementwise within
1e-4 relative plus a floor of 64 float32 roundoffs of each output's magnitude
sum_t |term_t| carried through the same chain.
"""
import numpy as np
import pytest
import torch

from helpers import assert_close, magnitude
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def _clip(g, eps=1e-4):
    """utils.py:147-150 in FP64: rows divided by max(|g| + eps, 1); returns the scale too."""
    s = np.maximum(np.linalg.norm(g, axis=-1, keepdims=True) + eps, 1.0)
    return g / s, s


@pytest.mark.parametrize("mode", ["reference", "raw"])
def test_config0_frame_matches_oracle_chain(dev, mode):
    from splatting_amd import extensions, render, scenes
    sc = scenes.config(0)
    W, H, f = sc.width, sc.height, sc.focal
    N = sc.n
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
    M = t(sc.view)[None].contiguous()
    opts = render.Options(numSamples=40, precision=0.01, shadingMode="None", device=dev)
    r = render.Splatting(opts, None, W, H, f, 0.1, 10.0, 0.0)
    r.seed = 0
    zero = t(np.zeros((1, 1, 3)))
    rng = np.random.default_rng(0)
    perturbed = (sc.positions + rng.normal(0.0, 0.01, size=sc.positions.shape)).astype(np.float32)
    with torch.no_grad():
        target, _, _ = r.forward(None, None, t(perturbed)[None], t(sc.normals)[None], t(sc.stddevs)[None],
                                 t(sc.albedo)[None], zero, zero, zero, transformMatrix=M)
    leaves = [t(a)[None].requires_grad_(True) for a in (sc.positions, sc.normals, sc.stddevs, sc.albedo)]
    extensions.set_ewa_gradient_mode(mode)
    try:
        image, indices, depthW = r.forward(None, None, *leaves, zero, zero, zero, transformMatrix=M)
        ((image - target) ** 2).sum().backward()
    finally:
        extensions.set_ewa_gradient_mode("reference")
    torch.cuda.synchronize()

    # camera space as the GPU computed it (the sampler's bit-exact check needs its inputs)
    from cudaExtensions import transform
    camP = transform.affineForward(M, leaves[0].detach()).cpu().numpy()[0]
    camN = transform.linearForward(M, leaves[1].detach()).cpu().numpy()[0]
    back = np.where(camN[:, 2:3] > 0, -camN, camN).astype(np.float32)
    idx = indices.cpu().numpy()
    ref_idx = O.sample_view(W, H, f, camP, back, sc.stddevs, 40, 0.0, 0.01, seed=0)
    assert np.array_equal(idx[0], ref_idx), "sampled indices differ from the oracle"
    assert (idx >= 0).sum() > 100_000

    # the oracle's FP64 chain on the same indices, from the GPU's float32 camera-space
    # points and normals (the transform's own rounding, relative 2^-24, would be
    # amplified by the EWA exponent of slots a few std-devs away); the transforms'
    # adjoints are checked against the oracle on their own
    Mv = sc.view.astype(np.float64)[None]
    P64 = camP[None].astype(np.float64)
    N64 = camN[None].astype(np.float64)
    S64 = sc.stddevs[None].astype(np.float64)
    A64 = sc.albedo[None].astype(np.float64)
    # EWA weights: the GPU's (the renderer's kernel on the same inputs), checked against
    # the oracle with the float32 evaluation condition of each weight; the compositing,
    # interpolation and adjoint chain below then starts from them
    from cudaExtensions import splatting
    ewa = splatting.forward(W, H, f, indices, t(camP)[None], t(camN)[None], t(sc.stddevs)[None],
                            0.0).cpu().numpy().astype(np.float64)
    assert_close(ewa, O.splat_forward(W, H, f, idx, P64, N64, S64), RTOL, mode + ": EWA weights",
                 mag=magnitude(O.splat_forward, W, H, f, idx, P64, N64, S64))
    dw = O.filter_forward(idx, ewa, W, H, f, camP[None], camN[None])
    img = O.interp_forward(idx, dw, A64)
    m_ewa = ewa                                                   # non-negative, no accumulation
    m_dw = dw                                                     # products of non-negative factors
    m_img = magnitude(O.interp_forward, idx, m_dw, A64)
    assert_close(image.detach().cpu().numpy(), img, RTOL, mode + ": image", mag=m_img)

    tgt = target.cpu().numpy().astype(np.float64)
    g_img = 2.0 * (img - tgt)
    m_gimg = 2.0 * (m_img + np.abs(tgt))
    g_dw, g_A = O.interp_backward(g_img, idx, dw, A64)
    m_gdw, m_gA = magnitude(O.interp_backward, m_gimg, idx, m_dw, A64)
    exact = mode == "raw"
    g_ewa = O.filter_backward(idx, ewa, g_dw, W, H, f, camP[None], camN[None], exact=exact)
    m_gewa = magnitude(O.filter_backward, idx, m_ewa, m_gdw, W, H, f, camP[None], camN[None], exact=exact)
    g_cP, g_cN, g_S = O.splat_backward(W, H, f, g_ewa, idx, P64, N64, S64)
    m_cP, m_cN, m_S = magnitude(O.splat_backward, W, H, f, m_gewa, idx, P64, N64, S64)
    if mode == "reference":
        g_cP, sP = _clip(g_cP)
        g_S, sS = _clip(g_S)
        m_cP, m_S = m_cP / sP, m_S / sS
    _, g_P = O.transform_backward(True, g_cP, Mv, sc.positions[None])
    _, m_P = magnitude(O.transform_backward, True, m_cP, Mv, sc.positions[None])
    # elements twelve orders of magnitude below a tensor's largest carry only terms
    # of weights below e^-87, which float32 evaluation flushes to zero
    floor = lambda ref: 1e-12 * float(np.abs(ref).max())
    assert_close(leaves[0].grad.cpu().numpy(), g_P, RTOL, mode + ": grad positions", mag=m_P, atol=floor(g_P))
    assert_close(leaves[2].grad.cpu().numpy(), g_S, RTOL, mode + ": grad std-devs", mag=m_S, atol=floor(g_S))
    assert_close(leaves[3].grad.cpu().numpy(), g_A, RTOL, mode + ": grad albedo", mag=m_gA, atol=floor(g_A))
    if mode == "reference":
        # EWASplatting returns no normal gradient (a zero tensor where another op's
        # backward still reaches the normals)
        assert leaves[1].grad is None or not bool(leaves[1].grad.any())
    else:
        _, g_N = O.transform_backward(False, g_cN, Mv, sc.normals[None])
        _, m_N = magnitude(O.transform_backward, False, m_cN, Mv, sc.normals[None])
        assert_close(leaves[1].grad.cpu().numpy(), g_N, RTOL, mode + ": grad normals", mag=m_N, atol=floor(g_N))
    assert np.abs(g_P).max() > 0 and np.abs(g_S).max() > 0


def test_fused_backface_equals_torch_where(dev):
    """The renderer's camera-facing normals from the transform pass
    (fuseBackface) give the same indices and image as the reference's
    linearForward + torch.where path, and the same gradients up to the order
    of their atomic sums."""
    from splatting_amd import render, scenes
    sc = scenes.config(0)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
    M = t(sc.view)[None].contiguous()
    zero = t(np.zeros((1, 1, 3)))
    outs = []
    for fused in (True, False):
        opts = render.Options(numSamples=40, precision=0.01, shadingMode="None", device=dev)
        r = render.Splatting(opts, None, sc.width, sc.height, sc.focal, 0.1, 10.0, 0.0)
        r.seed, r.fuseBackface = 3, fused
        leaves = [t(a)[None].requires_grad_(True) for a in (sc.positions, sc.normals, sc.stddevs, sc.albedo)]
        image, indices, _ = r.forward(None, None, *leaves, zero, zero, zero, transformMatrix=M)
        (image ** 2).sum().backward()
        outs.append((image.detach(), indices, [a.grad for a in leaves]))
    (i1, x1, g1), (i2, x2, g2) = outs
    assert torch.equal(x1, x2)
    assert torch.equal(i1, i2)
    for a, b in zip(g1, g2):
        if a is None or b is None:
            assert a is None and b is None
            continue
        assert torch.allclose(a, b, rtol=1e-5, atol=1e-5 * float(b.abs().max()))
