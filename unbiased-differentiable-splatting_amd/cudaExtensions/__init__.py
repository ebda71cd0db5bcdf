"""Drop-in op surface of the reference's torch extensions (``cudaExtensions``),
implemented on hand-written gfx950 HIP kernels behind a C ABI.

Modules and functions mirror the reference bindings (setup.py:11-70):
    sampling.sample / covariance / approximateFMMWeights
    splatting.forward / backward
    filtering.forward / backward
    transform.affineForward / affineBackward / linearForward / linearBackward
    interpolation.forward / backward
    shading.forward / backward / forwardSh / backwardSh
    pixelgather.scatter / gather
    nearestsamples.nearest
and, with no reference counterpart, the fused Horvitz-Thompson image estimate
over sampling.samplePoisson's output (estimator.forward / backward).
"""
from . import sampling, splatting, filtering, transform, interpolation, shading, estimator, pixelgather, \
    nearestsamples  # noqa: F401
