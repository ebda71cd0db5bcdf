"""Shading modules (reference: splatting/shading.py)."""
import math

import torch as th

from .extensions import SphericalHarmonicShading


class Lambert(th.nn.Module):
    """Per-splat Lambert shading (shading.py:12-29): every splat's diffuse colour
    times the ambient light plus, for each lamp l, its intensity I_l scaled by
    clamp(d_l . n, 0, 1) / pi.  lampDirections (bn, L, 3), lampIntensities
    (1 | bn, L, 3), normals / diffuse (bn, N, 3) -> (bn, N, 3)."""

    def forward(self, ambientLight, lampDirections, lampIntensities, normals, diffuse):
        # cosine of every (lamp, splat) pair: (bn, L, N)
        cosines = th.matmul(lampDirections, normals.transpose(1, 2)).clamp(0.0, 1.0)
        # sum_l I_l cos_l: (bn, N, 3)
        direct = th.matmul(cosines.transpose(1, 2), lampIntensities.expand(cosines.shape[0], -1, -1))
        return diffuse * (ambientLight + direct / math.pi)


def _init_sh(batched):
    c = th.empty(batched, 3, 9)
    c.uniform_(-0.1, 0.1)
    c[..., 0] = 0.5
    return c


class DeferredSphericalHaromic(th.nn.Module):
    """Deferred SH shading on interpolated normals, shading.py:64-81."""

    def __init__(self, batched=1):
        super().__init__()
        self.coefficients = th.nn.Parameter(_init_sh(batched), requires_grad=True)
        self.shading = SphericalHarmonicShading.apply

    def forward(self, normals, diffuse):
        return self.shading(self.coefficients, normals.contiguous(), diffuse.contiguous())


class ForwardSphericalHaromic(th.nn.Module):
    """Per-splat SH shading in torch, shading.py:83-129."""

    def __init__(self):
        super().__init__()
        self.coefficients = th.nn.Parameter(_init_sh(1), requires_grad=True)

    def buildShMatrix(self):
        c1, c2, c3, c4, c5 = 0.429043, 0.511664, 0.743125, 0.886227, 0.247708
        k = self.coefficients
        M = th.empty([1, 3, 4, 4], device=k.device, dtype=k.dtype)
        M[..., 0, 0] = c1 * k[..., 8]; M[..., 1, 0] = c1 * k[..., 4]; M[..., 2, 0] = c1 * k[..., 7]; M[..., 3, 0] = c2 * k[..., 3]
        M[..., 0, 1] = c1 * k[..., 4]; M[..., 1, 1] = -c1 * k[..., 8]; M[..., 2, 1] = c1 * k[..., 5]; M[..., 3, 1] = c2 * k[..., 1]
        M[..., 0, 2] = c1 * k[..., 7]; M[..., 1, 2] = c1 * k[..., 5]; M[..., 2, 2] = c3 * k[..., 6]; M[..., 3, 2] = c2 * k[..., 2]
        M[..., 0, 3] = c2 * k[..., 3]; M[..., 1, 3] = c2 * k[..., 1]; M[..., 2, 3] = c2 * k[..., 2]
        M[..., 3, 3] = c4 * k[..., 0] - c5 * k[..., 6]
        return M

    def forward(self, normals, diffuse):
        shMatrix = self.buildShMatrix().view(1, 1, 3, 4, 4)
        n = th.cat([normals, th.ones([1, normals.shape[1], 1], device=diffuse.device, dtype=diffuse.dtype)], dim=-1)
        n = n.view(1, normals.shape[1], 1, 1, 4)
        image = th.sum((n @ shMatrix).mul(n), dim=-1)
        return diffuse * image.squeeze(-1) / math.pi
