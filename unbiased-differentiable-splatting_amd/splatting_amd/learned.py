"""Per-tile rate network for Poisson-rate sampling (no counterpart in the reference).

``RateNet`` maps 32 features of a 16x16 pixel tile to the logarithm of the
tile's sampling-rate multiplier.  Inference, the loss gradient and the Adam
update run in the HIP library (bfloat16 matrix-core products, FP32
accumulation); this module only holds the FP32 parameter block and the feature
standardisation, and drives training.  The multipliers only change how many
splats are drawn where: the inclusion weights stay 1/pi, so estimates stay
unbiased whatever the network predicts.
"""
import math

import numpy as np
import torch

from cudaExtensions import _lib

N_IN, N_HID = 32, 64


def _call(name, *args):
    _lib.check(getattr(_lib.lib(), name)(*args), name)


class RateNet:
    def __init__(self, params, feat_mean, feat_std):
        self.params = params                      # FP32 (6337,) on the GPU
        self.mean = feat_mean                     # (32,) feature standardisation
        self.std = feat_std

    @staticmethod
    def param_count():
        return int(_lib.lib().ude_mlp_param_count())

    @classmethod
    def init(cls, device, seed=0):
        """He-initialised weights, zero biases, identity standardisation."""
        g = np.random.default_rng(seed)
        p = np.zeros(cls.param_count(), np.float32)
        o = 0
        for fan_in, fan_out in ((N_IN, N_HID), (N_HID, N_HID), (N_HID, 1)):
            n = fan_in * fan_out
            p[o:o + n] = g.normal(0.0, math.sqrt(2.0 / fan_in), size=n)
            o += n + fan_out
        p[-1] = 0.0
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device)
        return cls(t(p), t(np.zeros(N_IN)), t(np.ones(N_IN)))

    def standardize(self, feats):
        return ((feats - self.mean) / self.std).contiguous()

    def forward(self, feats):
        """Predicted log multipliers for (T, 32) raw features."""
        X = self.standardize(_lib.f32(feats, "features"))
        T = X.shape[0]
        y = torch.empty(T, dtype=torch.float32, device=X.device)
        ws = _lib.workspace(_lib.lib().ude_mlp_workspace_bytes(T), X.device)
        _call("ude_mlp_forward", T, _lib.ptr(X), _lib.ptr(self.params), _lib.ptr(y), _lib.ptr(ws), ws.numel(),
              _lib.stream(X.device))
        return y

    def loss_grad(self, X, target, weight=None):
        """Weighted MSE of standardised features X (T, 32) against target (T,) and its gradient."""
        T = X.shape[0]
        grad = torch.empty_like(self.params)
        loss = torch.empty(1, dtype=torch.float32, device=X.device)
        ws = _lib.workspace(_lib.lib().ude_mlp_workspace_bytes(T), X.device)
        _call("ude_mlp_loss_grad", T, _lib.ptr(X), _lib.ptr(self.params), _lib.ptr(target), _lib.ptr(weight),
              _lib.ptr(grad), _lib.ptr(loss), _lib.ptr(ws), ws.numel(), _lib.stream(X.device))
        return loss, grad

    def fit(self, feats, target, weight=None, steps=2000, lr=3e-3, log_every=0):
        """Adam on the GPU; sets the feature standardisation from feats first."""
        feats = _lib.f32(feats, "features")
        self.mean = feats.mean(0)
        self.std = feats.std(0).clamp(min=1e-6)
        X = self.standardize(feats)
        target = _lib.f32(target, "target")
        weight = None if weight is None else _lib.f32(weight, "weight")
        m = torch.zeros_like(self.params)
        v = torch.zeros_like(self.params)
        hist = []
        for s in range(1, steps + 1):
            loss, grad = self.loss_grad(X, target, weight)
            _call("ude_adam_step", self.params.numel(), _lib.ptr(self.params), _lib.ptr(grad), _lib.ptr(m),
                  _lib.ptr(v), lr, 0.9, 0.999, 1e-8, s, _lib.stream(X.device))
            if log_every and (s % log_every == 0 or s == 1):
                hist.append((s, float(loss.item())))
        return hist

    def tile_scale(self, feats, counts, lo=0.75, hi=4.0):
        """Rate multipliers for tiles with features (..., 32) and plain-rate sample
        counts (...): exp(prediction), rescaled so that the expected sample count
        sum(scale * counts) is unchanged, and kept within [lo, hi]."""
        shape = counts.shape
        y = self.forward(feats.reshape(-1, N_IN)).reshape(shape).double()
        y = torch.nan_to_num(y, nan=0.0, posinf=20.0, neginf=-20.0)     # a non-finite feature: plain rate
        c = counts.double()
        tot = c.sum().clamp(min=1e-30)
        y = y - (y * c).sum() / tot
        s = torch.exp(y.clamp(-20.0, 20.0))
        for _ in range(3):
            s = (s * (tot / (s * c).sum().clamp(min=1e-30))).clamp(lo, hi)
        return s.float().contiguous()

    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy(), mean=self.mean.cpu().numpy(), std=self.std.cpu().numpy())

    @classmethod
    def load(cls, path, device):
        d = np.load(path)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device)
        return cls(t(d["params"]), t(d["mean"]), t(d["std"]))


# --------------------------------------------------------------------------- Poisson-rate frames
class PoissonFrame:
    """One fwd+bwd frame of the Horvitz-Thompson renderer over one or more views
    (``scene.view`` (4, 4) or (V, 4, 4)): world positions / normals through the
    affine / linear transforms, Poisson-rate sampling of all views (optionally
    with per-tile rate multipliers), EWA weights times 1/pi, interpolated albedo
    (shading "None"), L2 loss against ``target``, and the raw adjoints of every
    op.  ``composite``: the samples of a pixel are z-sorted and alpha-composited
    front to back (alpha = min(w, 1) / pi), an unbiased estimate of the image
    composited over all splats; otherwise they are summed without occlusion.
    Gradients are with respect to world positions / normals, std-devs and albedo,
    summed over the views."""

    def __init__(self, scene, device, rate=10.0, kcap=40, precision=0.01, target=None, composite=False,
                 proposal="bands", footprint_scale=None):
        from cudaExtensions import transform
        self.sc = scene
        self.dev = device
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device)
        view = np.asarray(scene.view, np.float32)
        self.M = (t(view) if view.ndim == 3 else t(view)[None]).contiguous()
        self.V = self.M.shape[0]
        self.Pw = t(scene.positions)[None].contiguous()
        self.Nw = t(scene.normals)[None].contiguous()
        self.S = t(scene.stddevs)[None].contiguous()
        self.A = t(scene.albedo)[None].contiguous()
        self.rate, self.kcap, self.prec = rate, kcap, precision
        self.composite = bool(composite)
        self.proposal = proposal                 # samplePoisson's proposal: "bands" or "footprint"
        self.footprint_scale = footprint_scale
        self.target = target
        self.W, self.H, self.f = scene.width, scene.height, scene.focal
        self.dens = torch.zeros((self.V, 2, self.W, self.H), dtype=torch.float32, device=device)
        self._transform = transform
        self._ws = None              # fused-estimator workspace shared by forward and backward
        self._ws_shape = None
        # error bits of every draw of this frame object (ORed on the device, never
        # cleared): a capacity overflow drops samples, so callers check it
        self.errors = torch.zeros(1, dtype=torch.int32, device=device)

    def check_errors(self):
        """Raise if any draw so far overflowed its capacity or its walk bound (synchronises)."""
        from cudaExtensions import sampling
        sampling.raise_errors(self.errors.item())

    def camera(self):
        tr = self._transform
        P = tr.affineForward(self.M, self.Pw)
        Nc, nb = tr.linearForwardFacing(self.M, self.Nw)
        return P, Nc, nb

    def image(self, seed, tile_scale=None, rate=None, kcap=None, seed_tensor=None):
        """Sampled image estimate (V, W, H, 3) (EWA weights, 1/pi and albedo
        interpolation fused in one pass) and the sample tensors."""
        from cudaExtensions import sampling, estimator
        P, Nc, nb = self.camera()
        ids, ipi = sampling.samplePoisson(self.W, self.H, self.f, P, nb, None, self.S, kcap or self.kcap, 0.0,
                                          self.prec, seed=seed, rate=rate or self.rate, tileScale=tile_scale,
                                          densities=self.dens, seedTensor=seed_tensor, errors=self.errors,
                                          proposal=self.proposal, footprintScale=self.footprint_scale)
        if self._ws is None or self._ws_shape != tuple(ids.shape):
            self._ws = estimator.workspace(ids, self.S, composite=self.composite)
            self._ws_shape = tuple(ids.shape)
        img = estimator.forward(self.W, self.H, self.f, ids, ipi, P, Nc, self.S, self.A, workspace=self._ws,
                                composite=self.composite)
        self.last_ids = ids
        return img, (P, Nc, ids, ipi)

    def step(self, seed, tile_scale=None, rate=None, kcap=None, seed_tensor=None, with_loss=True):
        """Returns (loss, image, grads (positions, normals, std-devs, albedo), dL/dimage);
        loss (a float64 0-d tensor) is None without with_loss."""
        from cudaExtensions import estimator
        tr = self._transform
        img, (P, Nc, ids, ipi) = self.image(seed, tile_scale, rate, kcap, seed_tensor)
        loss, g = estimator.l2(img, self.target)          # one pass, no memset (graph-safe)
        if not with_loss:
            loss = None
        gP, gN, gS, gA = estimator.backward(self.W, self.H, self.f, g, ids, ipi, P, Nc, self.S, self.A,
                                            workspace=self._ws, prepared=True, composite=self.composite)
        _, gPw = tr.affineBackward(gP, self.M, self.Pw)
        _, gNw = tr.linearBackward(gN, self.M, self.Nw)
        return loss, img, (gPw, gNw, gS, gA), g


def mean_image(frame, seeds, rate=None):
    """Average of the sampled image estimates over the given seeds."""
    acc = None
    for s in seeds:
        img, _ = frame.image(s, rate=rate)
        acc = img.double() if acc is None else acc + img.double()
    return (acc / len(seeds)).float().contiguous()


def tile_counts(ids):
    """Number of samples per 16x16 tile, (bn, ceil(W/16), ceil(H/16))."""
    bn, W, H, _ = ids.shape
    c = (ids >= 0).sum(-1).float()
    pw, ph = (-W) % 16, (-H) % 16
    c = torch.nn.functional.pad(c, (0, ph, 0, pw))
    return c.reshape(bn, c.shape[1] // 16, 16, c.shape[2] // 16, 16).sum((2, 4))


def tile_sum(img):
    bn, W, H = img.shape[:3]
    x = img.reshape(bn, W, H, -1).sum(-1)
    x = torch.nn.functional.pad(x, (0, (-H) % 16, 0, (-W) % 16))
    return x.reshape(bn, x.shape[1] // 16, 16, x.shape[2] // 16, 16).sum((2, 4))


def splat_tiles(frame):
    """Tile index (ti * tH + tj) of every splat's screen centre, -1 outside the image."""
    P, Nc, nb = frame.camera()
    p = P[0]
    sx = -(p[:, 0] * frame.f) / p[:, 2] + 0.5 * frame.W
    sy = (p[:, 1] * frame.f) / p[:, 2] + 0.5 * frame.H
    ok = (p[:, 2] < 0) & (sx >= 0) & (sx < frame.W) & (sy >= 0) & (sy < frame.H)
    tH = (frame.H + 15) // 16
    idx = (sx.clamp(0, frame.W - 1).long() >> 4) * tH + (sy.clamp(0, frame.H - 1).long() >> 4)
    return torch.where(ok, idx, torch.full_like(idx, -1))


def tile_variance(frame, seeds, rate=None):
    """Per 16x16 tile: the summed per-parameter variance of the sampled gradient
    over the splats centred in the tile (sampling is independent per pixel, so a
    splat's gradient variance comes from the pixels around it), the mean sample
    count, and the mean image gradient."""
    s1 = s2 = cnt = g_img = None
    for s in seeds:
        _, _, grads, g = frame.step(s, rate=rate)
        v = torch.cat([grads[0][0], grads[1][0], grads[2][0], grads[3][0]], dim=-1).double()   # (N, 11)
        s1 = v if s1 is None else s1 + v
        s2 = v * v if s2 is None else s2 + v * v
        g_img = g if g_img is None else g_img + g
        _, (P, Nc, ids, ipi) = frame.image(s, rate=rate)
        c = tile_counts(ids).double()
        cnt = c if cnt is None else cnt + c
    S = len(seeds)
    mean = s1 / S
    var = ((s2 - S * mean * mean) / (S - 1)).clamp(min=0).sum(-1)            # per splat
    tiles = splat_tiles(frame)
    V = torch.zeros(cnt.numel(), dtype=torch.float64, device=frame.dev)
    ok = tiles >= 0
    V.index_add_(0, tiles[ok], var[ok])
    return V, (cnt / S).reshape(-1), g_img / S


def variance_targets(frame, seeds, low=0.25):
    """Plain-rate variance measurements of one scene at rate lambda and lambda *
    low.  Under Poisson sampling a tile's gradient variance behaves like
    V_t(s) = a_t / s + b_t in its rate multiplier s (saturated inclusions, pi = 1,
    contribute only once s drops), so two rates give the marginal cost
    a_t = (V_t(low) - V_t(1)) / (1 / low - 1).  The Neyman allocation of a fixed
    budget sum_t s_t n_t (n_t = plain samples in the tile) minimises sum_t a_t / s_t
    with s_t proportional to sqrt(a_t / n_t): the training target is
    y_t = log sqrt(a_t / n_t) (centred per view), weighted by n_t.
    Returns (features (T, 32), target (T,), weight (T,))."""
    from cudaExtensions import sampling
    V1, n, g_img = tile_variance(frame, seeds)
    Vl, _, _ = tile_variance(frame, [s + 7919 for s in seeds], rate=frame.rate * low)
    a = ((Vl - V1) / (1.0 / low - 1.0)).clamp(min=0)
    P, Nc, nb = frame.camera()
    feats = sampling.tileFeatures(frame.W, frame.H, frame.f, P, nb, frame.S, 0.0, frame.prec, frame.dens,
                                  g_img.contiguous())
    amax = float(a.max().item())
    y = 0.5 * torch.log((a + 1e-9 * amax + 1e-30) / (n + 1e-3))
    # only the allocation within a view matters: centre the targets (count-weighted)
    y = y - (y * n).sum() / n.sum().clamp(min=1e-30)
    return feats.reshape(-1, N_IN), y.float(), n.float()


def gradient_variance(frame, seeds, tile_scale=None, rate=None, kcap=None, groups=1, out_groups=None):
    """Mean over all parameters of the per-parameter variance of the sampled
    gradient across seeds (raw adjoints), and the mean gradient.  With groups > 1
    the seeds are split into that many consecutive groups and the median of the
    groups' values is returned: the estimator's heavy tails (a rare pair with a
    small inclusion probability and a large weight) can put one seed's gradient
    far out, and the median of groups keeps such a seed from deciding the value.
    ``out_groups``: a list that receives the groups' values."""
    S = len(seeds)
    G = max(1, min(int(groups), S // 2))
    bounds = [S * q // G for q in range(G + 1)]
    vals, total = [], None
    for q in range(G):
        s1 = s2 = None
        for s in seeds[bounds[q]:bounds[q + 1]]:
            _, _, grads, _ = frame.step(s, tile_scale, rate, kcap)
            x = torch.cat([gq.reshape(-1) for gq in grads]).double()
            s1 = x if s1 is None else s1 + x
            s2 = x * x if s2 is None else s2 + x * x
        n = bounds[q + 1] - bounds[q]
        mean = s1 / n
        vals.append(float(((s2 - n * mean * mean) / (n - 1)).clamp(min=0).mean().item()))
        total = s1 if total is None else total + s1
    if out_groups is not None:
        out_groups.extend(vals)
    return float(np.median(vals)), total / S


def kappa_interval(plain_groups, learned_groups, resamples=4000, seed=0):
    """95 % bootstrap interval of kappa = median(learned) / median(plain) over the
    seed groups (groups resampled with replacement within each set)."""
    rng = np.random.default_rng(seed)
    p = np.asarray(plain_groups, np.float64)
    q = np.asarray(learned_groups, np.float64)
    r = [np.median(q[rng.integers(0, q.size, q.size)]) / np.median(p[rng.integers(0, p.size, p.size)])
         for _ in range(resamples)]
    return [float(np.percentile(r, 2.5)), float(np.percentile(r, 97.5))]


def samples_per_pixel(frame, seed, rate=None, kcap=None):
    frame.image(seed, rate=rate, kcap=kcap)
    return float((frame.last_ids >= 0).float().sum(-1).mean().item())


def compare_proposals(frame, seeds, groups=8, kcap=None):
    """The band proposal against the footprint proposal on one frame at the same
    expected sample count: the footprint rate is set so that its mean samples per
    pixel equal the bands' at frame.rate (two secant steps on the count, which
    grows with the rate), then gradient variances are measured over disjoint seed
    sets as in compare_rates (median of seed groups); a second band measurement
    gives the noise.  kappa = V_footprint / V_bands at equal samples, so the
    footprint proposal reaches the bands' variance at about rate_equal * kappa.
    The footprint draws use capacity ``kcap`` (default twice the frame's: its
    samples concentrate on the pixels a splat covers).  The frame's proposal is
    left as it was."""
    S = len(seeds)
    kf = kcap or 2 * frame.kcap
    base = seeds[0]
    keep = frame.proposal
    try:
        frame.proposal = "bands"
        n_b = samples_per_pixel(frame, base - 1)
        gb, gr, gf = [], [], []
        v_b, _ = gradient_variance(frame, [base + s for s in range(S)], groups=groups, out_groups=gb)
        v_rep, _ = gradient_variance(frame, [base + S + s for s in range(S)], groups=groups, out_groups=gr)
        frame.proposal = "footprint"
        r = frame.rate
        n_f = samples_per_pixel(frame, base - 1, rate=r, kcap=kf)
        for _ in range(2):
            r = r * n_b / max(n_f, 1e-9)
            n_f = samples_per_pixel(frame, base - 1, rate=r, kcap=kf)
        v_f, _ = gradient_variance(frame, [base + 2 * S + s for s in range(S)], rate=r, kcap=kf, groups=groups,
                                   out_groups=gf)
    finally:
        frame.proposal = keep
    return {"band_grad_variance": v_b, "band_grad_variance_repeat": v_rep,
            "measurement_noise": abs(v_rep / v_b - 1.0),
            "footprint_rate_equal_samples": r, "footprint_grad_variance_equal_samples": v_f,
            "kappa": v_f / v_b, "kappa_ci95": kappa_interval(gb + gr, gf),
            "samples_per_pixel_bands": n_b, "samples_per_pixel_footprint": n_f, "kcap_footprint": kf,
            "variance_seeds": S, "variance_seed_groups": groups}


def rate_features(frame, net, seed):
    """One plain frame: tile features of its densities and image gradient, and the
    learned tile multipliers for the plain per-tile sample counts."""
    from cudaExtensions import sampling
    _, _, _, g = frame.step(seed)
    P, Nc, nb = frame.camera()
    feats = sampling.tileFeatures(frame.W, frame.H, frame.f, P, nb, frame.S, 0.0, frame.prec, frame.dens, g)
    counts = tile_counts(frame.last_ids)[0]
    return feats, counts, net.tile_scale(feats[0], counts)[None].contiguous()


def max_count(ids):
    return int((ids >= 0).sum(-1).max().item())


def compare_rates(frame, net, seeds, kcap_learned=None, groups=8):
    """Plain against learned per-tile rates on one frame at the same expected
    sample count.  Gradient variances (mean per-parameter variance of the raw
    adjoints) are measured over disjoint seed sets; a second plain measurement on
    other seeds gives the noise of the measurement itself.  Variance behaves like
    1 / rate, so the learned rates reach the plain variance at rate * kappa with
    kappa = V_learned / V_plain.  Returns a dict of the measurements and the tile
    multipliers.  Each variance is the median over `groups` seed groups."""
    S = len(seeds)
    base = seeds[0]
    feats, counts, ts = rate_features(frame, net, base - 1)
    kcl = kcap_learned or frame.kcap * 2
    G = groups                             # median of seed groups (heavy tails)
    gp, gr, gl = [], [], []
    v_plain, _ = gradient_variance(frame, [base + s for s in range(S)], groups=G, out_groups=gp)
    v_rep, _ = gradient_variance(frame, [base + S + s for s in range(S)], groups=G, out_groups=gr)
    v_learned, _ = gradient_variance(frame, [base + 2 * S + s for s in range(S)], tile_scale=ts, kcap=kcl, groups=G,
                                     out_groups=gl)
    frame.image(base - 2)
    n_plain = float((frame.last_ids >= 0).float().sum(-1).mean().item())
    over_plain = max_count(frame.last_ids) >= frame.kcap
    frame.image(base - 2, tile_scale=ts, kcap=kcl)
    n_learned = float((frame.last_ids >= 0).float().sum(-1).mean().item())
    over_learned = max_count(frame.last_ids) >= kcl
    return {"plain_grad_variance": v_plain, "plain_grad_variance_repeat": v_rep,
            "measurement_noise": abs(v_rep / v_plain - 1.0),
            "learned_grad_variance_equal_budget": v_learned, "kappa": v_learned / v_plain,
            "kappa_ci95": kappa_interval(gp + gr, gl),
            "samples_per_pixel_plain": n_plain, "samples_per_pixel_learned": n_learned,
            "kcap_plain": frame.kcap, "kcap_learned": kcl, "capacity_reached": over_plain or over_learned,
            "tile_scale_range": [float(ts.min().item()), float(ts.max().item())], "variance_seeds": S,
            "variance_seed_groups": G,
            "tile_scale": ts}


def make_target(scene, device, rate, kcap, precision=0.01, seeds=range(10_000, 10_016), composite=False):
    """Target image: the mean Poisson-rate estimate of the scene with perturbed centres."""
    from . import scenes
    pert = scenes.Scene(scenes.perturbed_target_positions(scene), scene.normals, scene.stddevs, scene.albedo,
                        scene.width, scene.height, scene.focal, scene.view)
    fr = PoissonFrame(pert, device, rate=rate, kcap=kcap, precision=precision, composite=composite)
    return mean_image(fr, list(seeds))


def train_rate_net(device, scenes_k=range(20), seeds=8, rate=10.0, kcap=40, steps=3000, n_override=None,
                   log=print):
    """Trains a RateNet on plain-rate variance measurements of random ellipsoid
    scenes (all on the GPU: sampling, rendering and the network's own kernels)."""
    from . import scenes
    F, Y, Wt = [], [], []
    for k in scenes_k:
        sc = scenes.training_scene(k, n_override)
        target = make_target(sc, device, rate, kcap)
        fr = PoissonFrame(sc, device, rate=rate, kcap=kcap, target=target)
        f, y, w = variance_targets(fr, list(range(100 * k, 100 * k + seeds)))
        F.append(f); Y.append(y); Wt.append(w)
        log("rate net data: scene %d, %d tiles" % (k, f.shape[0]))
    F, Y, Wt = torch.cat(F), torch.cat(Y), torch.cat(Wt)
    net = RateNet.init(device, seed=0)
    hist = net.fit(F, Y, Wt, steps=steps, lr=3e-3, log_every=max(steps // 6, 1))
    for s, l in hist:
        log("rate net training: step %d loss %.4f" % (s, l))
    return net, hist


DEFAULT_WEIGHTS = __file__.rsplit("/", 1)[0] + "/rate_net.npz"


def main():
    import argparse
    ap = argparse.ArgumentParser(description="train the per-tile rate network on the GPU")
    ap.add_argument("--scenes", type=int, default=20)
    ap.add_argument("--seeds", type=int, default=8)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--splats", type=int, default=None)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    args = ap.parse_args()
    dev = torch.device("cuda:0")
    net, hist = train_rate_net(dev, range(args.scenes), args.seeds, steps=args.steps, n_override=args.splats,
                               log=lambda m: print(m, flush=True))
    net.save(args.out)
    print("saved", args.out, flush=True)


if __name__ == "__main__":
    main()
