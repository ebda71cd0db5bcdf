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

    def tile_scale(self, feats, counts):
        """Rate multipliers exp(prediction), rescaled so that the expected sample
        count over the tiles, sum(scale * counts), is unchanged."""
        s = torch.exp(self.forward(feats).clamp(-8.0, 8.0))
        budget = counts.sum()
        got = (s * counts).sum()
        return torch.where(got > 0, s * (budget / got.clamp(min=1e-30)), s)

    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy(), mean=self.mean.cpu().numpy(), std=self.std.cpu().numpy())

    @classmethod
    def load(cls, path, device):
        d = np.load(path)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device)
        return cls(t(d["params"]), t(d["mean"]), t(d["std"]))
