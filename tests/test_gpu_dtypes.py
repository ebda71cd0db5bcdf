"""The float64 and half op surface (the reference dispatches splatting,
transform, interpolation and shading on float and double and filtering also on
half: splattingCuda.cu:84,214; transformCuda.cu:53,135; interpolationCuda.cu:68,153;
shadingCuda.cu:63,153,225,315; zSortingCuda.cu:100,220):

  * float64: every op and adjoint against the oracle's FP64 restatement,
    elementwise within 1e-10 relative plus 1024 float64 roundoffs of each
    output's magnitude (its terms' sum of |term| times their evaluation
    condition; the floor matters only far in the Gaussian tails, e.g. a weight
    of 1e-259 on an edge-on surfel, where it is 2e-9 relative);
  * float16 filtering: the float32 op on the same (half-rounded) inputs, within
    half a half-precision ulp of its float32 result plus one float32 rounding
    (the half kernel rounds its last product straight to half, once, where the
    float32 op rounds to float32 first; either is a correct rounding);
  * FP64 central differences on the GPU: the hand-written adjoint of the
    reference's op chain splatting -> filtering -> interpolation (fixed sampled
    indices, 256 splats at 64x64) against central differences of the chain's
    float64 forward at steps 1e-4 and 2.5e-5, 32 random parameters of each kind;
    a parameter whose two differences disagree (a depth-order change of two
    splats within the step makes the chain discontinuous there: about one point
    coordinate in five on this overlapping 256-splat sphere) is replaced by
    another draw, at most 48 times per kind.

"""
import numpy as np
import pytest
import torch

from helpers import EPS64, assert_close, magnitude, near_slots, random_slots, scene_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def t(x, dev, dt=torch.float64):
    x = np.ascontiguousarray(x)
    return torch.from_numpy(x).to(dev) if x.dtype.kind in "iu" else torch.from_numpy(x).to(dev, dt)


def close64(got, ref, what, mag=None):
    return assert_close(got.cpu().numpy(), ref, RTOL, what, mag=mag, ops=1024, eps=EPS64, atol=1e-300)


def within_half_rounding(h16, f32):
    """h16 (float16) is a rounding to half of a value within one float32 rounding of f32."""
    a = h16.double().cpu().numpy()
    b = f32.double().cpu().numpy()
    r16 = f32.half().cpu().numpy()
    ulp = np.maximum(np.spacing(np.abs(r16)), np.spacing(np.abs(h16.cpu().numpy()))).astype(np.float64)
    bound = 0.5 * ulp + np.abs(b) * 2.0 ** -24
    return bool(np.all(np.abs(a - b) <= bound)), float(np.mean(a != r16.astype(np.float64)))


@pytest.fixture(scope="module")
def case():
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    rng = np.random.default_rng(7)
    idx = near_slots(rng, p, sc.width, sc.height, sc.focal, 12)
    return sc, p.astype(np.float64), n.astype(np.float64), sd.astype(np.float64), idx


def test_splatting_f64(dev, case):
    from cudaExtensions import splatting
    sc, p, n, sd, idx = case
    W, H, f = sc.width, sc.height, sc.focal
    w = splatting.forward(W, H, f, t(idx, dev), t(p[None], dev), t(n[None], dev), t(sd[None], dev), 0.0)
    assert w.dtype == torch.float64
    args = (W, H, f, idx, p[None], n[None], sd[None])
    close64(w, O.splat_forward(*args), "f64 splat forward", magnitude(O.splat_forward, *args))
    g = np.random.default_rng(11).normal(size=idx.shape)
    grads = splatting.backward(W, H, f, t(g, dev), t(idx, dev), t(p[None], dev), t(n[None], dev), t(sd[None], dev),
                               0.0)
    ref = O.splat_backward(W, H, f, g, idx, p[None], n[None], sd[None])
    mag = magnitude(O.splat_backward, W, H, f, g, idx, p[None], n[None], sd[None])
    for a, r, m, nm in zip(grads, ref, mag, ("points", "normals", "std-devs")):
        assert a.dtype == torch.float64
        close64(a, r, "f64 splat grad " + nm, m)


@pytest.mark.parametrize("k", [12, 100])
def test_filtering_f64_and_f16(dev, k):
    from cudaExtensions import filtering
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    W, H, f = sc.width, sc.height, sc.focal
    rng = np.random.default_rng(5 + k)
    idx = random_slots(rng, 1, W, H, k, p.shape[0])
    w = rng.uniform(0.0, 1.0, size=idx.shape)
    w[idx < 0] = 0.0
    g = rng.normal(size=idx.shape)
    # float64: depths in FP64 too; the oracle orders by FP32 depths, so compare on a
    # scene whose depths are FP32-representable orderings (distinct splats differ by far more)
    out = filtering.forward(t(idx, dev), t(w, dev), W, H, f, t(p[None], dev), t(n[None], dev))
    ref = O.filter_forward(idx, w, W, H, f, p[None], n[None])
    close64(out, ref, "f64 filter forward")
    gw = filtering.backward(t(idx, dev), t(w, dev), t(g, dev), W, H, f, t(p[None], dev), t(n[None], dev))
    rg = O.filter_backward(idx, w, g, W, H, f, p[None], n[None])
    close64(gw, rg, "f64 filter backward", magnitude(O.filter_backward, idx, w, g, W, H, f, p[None], n[None]))
    # float16: the float32 op on the same rounded inputs, output rounded to half
    h = lambda x: t(x, dev, torch.float16)
    wh, gh, ph, nh = h(w), h(g), h(p[None]), h(n[None])
    f = float(f)
    o16 = filtering.forward(t(idx, dev), wh, W, H, f, ph, nh)
    o32 = filtering.forward(t(idx, dev), wh.float(), W, H, f, ph.float(), nh.float())
    assert o16.dtype == torch.float16
    ok, frac = within_half_rounding(o16, o32)
    assert ok and frac < 0.01, frac
    b16 = filtering.backward(t(idx, dev), wh, gh, W, H, f, ph, nh)
    b32 = filtering.backward(t(idx, dev), wh.float(), gh.float(), W, H, f, ph.float(), nh.float())
    ok, frac = within_half_rounding(b16, b32)
    assert ok and frac < 0.01, frac


@pytest.mark.parametrize("affine", [True, False])
def test_transform_f64(dev, affine):
    from cudaExtensions import transform
    rng = np.random.default_rng(1)
    M, v, g = rng.normal(size=(3, 4, 4)), rng.normal(size=(1, 5000, 3)), rng.normal(size=(3, 5000, 3))
    fw = transform.affineForward if affine else transform.linearForward
    bw = transform.affineBackward if affine else transform.linearBackward
    close64(fw(t(M, dev), t(v, dev)), O.transform_forward(affine, M, v), "f64 transform forward",
            magnitude(O.transform_forward, affine, M, v))
    gM, gV = bw(t(g, dev), t(M, dev), t(v, dev))
    rM, rV = O.transform_backward(affine, g, M, v)
    mM, mV = magnitude(O.transform_backward, affine, g, M, v)
    close64(gM, rM, "f64 transform grad M", mM)
    close64(gV, rV, "f64 transform grad v", mV)


def test_interpolation_f64(dev):
    from cudaExtensions import interpolation
    rng = np.random.default_rng(3)
    bn, W, H, k, N = 2, 24, 20, 9, 300
    idx = random_slots(rng, bn, W, H, k, N)
    w, a, g = rng.uniform(size=idx.shape), rng.normal(size=(1, N, 3)), rng.normal(size=(bn, W, H, 3))
    close64(interpolation.forward(t(idx, dev), t(w, dev), t(a, dev)), O.interp_forward(idx, w, a),
            "f64 interp forward", magnitude(O.interp_forward, idx, w, a))
    gW, gA = interpolation.backward(t(g, dev), t(idx, dev), t(w, dev), t(a, dev))
    rW, rA = O.interp_backward(g, idx, w, a)
    mW, mA = magnitude(O.interp_backward, g, idx, w, a)
    close64(gW, rW, "f64 interp grad w", mW)
    close64(gA, rA, "f64 interp grad attr", mA)


def test_shading_f64(dev):
    from cudaExtensions import shading
    rng = np.random.default_rng(21)
    bn, W, H, L = 2, 16, 12, 3
    coef, nrm = rng.normal(size=(1, 3, 9)), rng.normal(size=(bn, W, H, 3))
    diff, g = rng.uniform(size=(bn, W, H, 3)), rng.normal(size=(bn, W, H, 3))
    close64(shading.forwardSh(t(coef, dev), t(nrm, dev), t(diff, dev)), O.sh_forward(coef, nrm, diff),
            "f64 sh forward", magnitude(O.sh_forward, coef, nrm, diff))
    got = shading.backwardSh(t(g, dev), t(coef, dev), t(nrm, dev), t(diff, dev))
    for a, r, m, nm in zip(got, O.sh_backward(g, coef, nrm, diff), magnitude(O.sh_backward, g, coef, nrm, diff),
                           ("coef", "normals", "diffuse")):
        close64(a, r, "f64 sh grad " + nm, m)
    dirs, ints = rng.normal(size=(bn, L, 3)), rng.uniform(size=(1, L, 3))
    close64(shading.forward(t(dirs, dev), t(ints, dev), t(nrm, dev), t(diff, dev)),
            O.lambert_forward(dirs, ints, nrm, diff), "f64 lambert forward",
            magnitude(O.lambert_forward, dirs, ints, nrm, diff))
    got = shading.backward(t(g, dev), t(dirs, dev), t(ints, dev), t(nrm, dev), t(diff, dev))
    ref = O.lambert_backward(g, dirs, ints, nrm, diff)
    mag = magnitude(O.lambert_backward, g, dirs, ints, nrm, diff)
    for a, r, m, nm in zip(got, ref, mag, ("dir", "int", "normal", "diffuse")):
        close64(a, r, "f64 lambert grad " + nm, m)


def test_mixed_dtypes_rejected(dev, case):
    from cudaExtensions import splatting
    sc, p, n, sd, idx = case
    with pytest.raises(TypeError):
        splatting.forward(sc.width, sc.height, sc.focal, t(idx, dev), t(p[None], dev), t(n[None], dev, torch.float32),
                          t(sd[None], dev), 0.0)


def test_gpu_fp64_central_differences_of_the_op_chain(dev):
    """L = sum g . interpolation(filtering(splatting(p, n, sd)), albedo) on fixed
    indices (the reference's sampler draws them), all in float64 on the GPU;
    the raw adjoints (splatting.backward of the exact filtering.backward -- the
    reference's drops terms at |w| or |1 - w| <= 1e-4 -- of
    interpolation.backward) against (L(x + h) - L(x - h)) / 2h at h = 2.5e-5, where
    the difference at h = 1e-4 agrees with it (the chain is smooth over the step)."""
    from cudaExtensions import filtering, interpolation, sampling, splatting
    sc, p, n, nb, sd = scene_inputs(2, sd_scale=8.0)
    W, H, f = sc.width, sc.height, sc.focal
    idx = sampling.sample(W, H, f, t(p[None], dev, torch.float32), t(nb[None], dev, torch.float32), None,
                          t(sd[None], dev, torch.float32), 16, 0.0, 0.01, seed=3)
    rng = np.random.default_rng(9)
    P, Nn, S = t(p[None], dev), t(n[None], dev), t(sd[None], dev)
    A = t(sc.albedo[None], dev)
    g = t(rng.normal(size=(1, W, H, 3)), dev)

    def loss(P_, N_, S_):
        w = splatting.forward(W, H, f, idx, P_, N_, S_, 0.0)
        dw = filtering.forward(idx, w, W, H, f, P_, N_)
        return float((interpolation.forward(idx, dw, A) * g).sum())

    w = splatting.forward(W, H, f, idx, P, Nn, S, 0.0)
    dw = filtering.forward(idx, w, W, H, f, P, Nn)
    gDW, _ = interpolation.backward(g, idx, dw, A)
    gW = filtering.backward(idx, w, gDW, W, H, f, P, Nn, exact=True)
    gP, gN, gS = splatting.backward(W, H, f, gW, idx, P, Nn, S, 0.0)
    h = 1e-4
    checked = 0
    for X, G, name in ((P, gP, "points"), (Nn, gN, "normals"), (S, gS, "std-devs")):
        flat = G.reshape(-1)
        cand = torch.nonzero(flat.abs() > 1e-3 * flat.abs().max()).reshape(-1).cpu().numpy()
        scale = float(flat.abs().max())

        def central(q, step):
            Xp, Xm = X.clone(), X.clone()
            Xp.view(-1)[q] += step
            Xm.view(-1)[q] -= step
            args_p = [Xp if Y is X else Y for Y in (P, Nn, S)]
            args_m = [Xm if Y is X else Y for Y in (P, Nn, S)]
            return (loss(*args_p) - loss(*args_m)) / (2 * step)

        draws = rng.permutation(cand)
        redrawn = 0
        done = 0
        for q in draws:
            fd, fd4 = central(q, h), central(q, h / 4)
            if abs(fd - fd4) > 1e-4 * abs(fd4) + 1e-6 * scale:
                redrawn += 1                  # not smooth over [x - h, x + h]
                assert redrawn <= 48, (name, "too many depth-order changes", int(q), fd, fd4)
                continue
            an = float(flat[q])
            assert abs(fd4 - an) <= 1e-4 * abs(an) + 1e-6 * scale, (name, int(q), fd4, an)
            done += 1
            if done == 32:
                break
        print("central differences of %s: 32 checked, %d redrawn (a depth-order change within the step)"
              % (name, redrawn))
        checked += done
    assert checked == 96
