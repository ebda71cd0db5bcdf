"""Differentiable surface-splatting renderer (reference: splatting/render.py:193-376)."""
import torch as th

from .extensions import (AffineTransform, LinearTransform, LinearTransformFacing, FmmSampling, EWASplatting,
                         DepthFiltering, LinearInterpolation, DiffuseShading)
from .shading import Lambert, DeferredSphericalHaromic, ForwardSphericalHaromic
from .utils import world2Camera, transformForward, normalize


def _normalizeNormals(normals):
    norms = normals.norm(dim=-1, keepdim=True)
    normals = th.where(norms < 0.01, th.ones_like(normals), normals)
    return normalize(normals)


class BackfaceCulling(th.nn.Module):
    """Orient normals so that camera-space z <= 0 (render.py:193-205)."""

    def forward(self, normals, camNormals):
        normalAngle = camNormals[..., 2].unsqueeze(-1)
        normals = th.where(normalAngle > 0, -normals, normals)
        camNormals = th.where(normalAngle > 0, -camNormals, camNormals)
        return normals, camNormals


class Options:
    """The subset of the reference's argparse options the renderer reads."""

    def __init__(self, numSamples=40, precision=0.01, shadingMode="None", forward="-z",
                 dtype=th.float32, device="cuda"):
        self.numSamples = numSamples
        self.precision = precision
        self.shadingMode = shadingMode
        self.forward = forward
        self.dtype = dtype
        self.device = device


class Splatting(th.nn.Module):
    """Splatting(opts, logger, width, height, focallength, zmin, zmax, smoothing) as in render.py:230-266."""

    def __init__(self, opts, logger, width, height, focallength, zmin, zmax, smoothing):
        super().__init__()
        self.logger = logger
        self.width, self.height, self.focallength = width, height, focallength
        self.coordForward = opts.forward
        self.smoothing = smoothing
        self.numSamples = opts.numSamples
        self.precision = opts.precision
        self.shadingMode = opts.shadingMode
        self.dtype, self.device = opts.dtype, opts.device
        self.affineTransform = AffineTransform.apply
        self.linearTransform = LinearTransform.apply
        self.backface = BackfaceCulling()
        self.sampling = FmmSampling()
        self.ewa = EWASplatting.apply
        self.depthBlending = DepthFiltering.apply
        self.linearInterpolation = LinearInterpolation.apply
        self.deferredShading = DiffuseShading.apply
        self.forwardShading = Lambert()
        self.sphericalHarmonics = DeferredSphericalHaromic()
        self.forwardHarmonics = ForwardSphericalHaromic()
        self.seed = None
        self.keepIndicesOnHost = True
        self.fuseBackface = True      # camera-facing normals from the transform pass where possible

    def forward(self, camPositions, camOrientations, pointPositions, pointNormals, pointStdDevs, pointDiffuse,
                ambientLight, lampDirections, lampIntensities, alpha=None, transformMatrix=None):
        """Returns (image, indices, depthWeights) (render.py:268-376).  ``transformMatrix``
        (bn, 4, 4) may be given directly instead of positions / orientations."""
        if transformMatrix is None:
            transformMatrix = world2Camera(camPositions, camOrientations)
            transformMatrix = transformForward(transformMatrix, self.coordForward).squeeze(1)
        transformMatrix = transformMatrix.contiguous()
        camPoints = self.affineTransform(transformMatrix, pointPositions)
        if (self.fuseBackface and pointNormals.dtype == th.float32
                and self.shadingMode not in ("sphericalHarmonics", "forwardHarmonics")):
            # only the spherical-harmonics modes read the flipped world normals:
            # the camera-facing copy comes out of the transform pass itself
            camNormals, backfaceNormals = LinearTransformFacing.apply(transformMatrix, pointNormals)
        else:
            camNormals = self.linearTransform(transformMatrix, pointNormals)
            pointNormals, backfaceNormals = self.backface(pointNormals, camNormals)
        with th.no_grad():
            indices = self.sampling(self.width, self.height, self.focallength, camPoints.detach().contiguous(),
                                    backfaceNormals.detach().contiguous(), pointDiffuse, pointStdDevs.detach(),
                                    self.numSamples, self.smoothing, self.precision, seed=self.seed)
            if self.keepIndicesOnHost:
                self.indices = indices.cpu().clone()
        ewaWeights = self.ewa(self.width, self.height, self.focallength, indices, camPoints, camNormals,
                              pointStdDevs, self.smoothing)
        depthWeights = self.depthBlending(indices, ewaWeights, self.width, self.height, self.focallength,
                                          camPoints, camNormals)
        camLampDirections = self.linearTransform(transformMatrix, lampDirections)
        mode = self.shadingMode
        if mode == "deferred":
            ewaNormals = _normalizeNormals(self.linearInterpolation(indices, depthWeights, camNormals))
            ewaDiffuse = self.linearInterpolation(indices, depthWeights, pointDiffuse)
            image = self.deferredShading(camLampDirections, lampIntensities, ewaNormals.contiguous(),
                                         ewaDiffuse.contiguous())
        elif mode == "forward":
            shadedPoints = self.forwardShading(ambientLight, camLampDirections, lampIntensities, camNormals,
                                               pointDiffuse)
            image = self.linearInterpolation(indices, depthWeights, shadedPoints.contiguous())
        elif mode == "sphericalHarmonics":
            ewaNormals = self.linearInterpolation(indices, depthWeights, pointNormals.contiguous())
            ewaDiffuse = self.linearInterpolation(indices, depthWeights, pointDiffuse)
            image = self.sphericalHarmonics(ewaNormals, ewaDiffuse)
        elif mode == "forwardHarmonics":
            shadedPoints = self.forwardHarmonics(pointNormals, pointDiffuse)
            image = self.linearInterpolation(indices, depthWeights, shadedPoints.contiguous())
        elif mode == "None":
            image = self.linearInterpolation(indices, depthWeights, pointDiffuse)
        elif mode == "silhouette":
            image = th.clamp(th.sum(ewaWeights, dim=-1, keepdim=True), 0, 1)
        else:
            raise ValueError("Unkown shading mode")
        return image, indices, depthWeights
