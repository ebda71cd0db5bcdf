"""pixelgather -- per-slot copies of the sampled splats' rows and their adjoint.

Mirrors splatting/extensions/pixelGather.cpp:10-23 (``scatter``, ``gather``,
bound as cudaExtensions.pixelgather, setup.py:27-31): scatter dispatches float,
double and half, gather float and double (pixelGatherCuda.cu:60, 139).
"""
import torch

from . import _lib


def scatter(indices, values):
    """out (bn, W, H, k, C): values[b, indices[b, i, j, s], :] per slot, 0 where the
    index is -1.  indices (bn, W, H, k) int32, values (bn, N, C)."""
    _lib.expect(indices, "indices", None, None, None, None)
    _lib.expect(values, "values", indices.shape[0], None, None)
    dev = _lib.require_gpu(indices, values)
    indices = _lib.i32(indices, "indices")
    dt, (values,) = _lib.reals([("values", values)], (torch.float32, torch.float64, torch.float16))
    bn, W, H, k = indices.shape
    N, C = values.shape[1], values.shape[2]
    out = torch.empty((bn, W, H, k, C), dtype=dt, device=dev)
    if out.numel() == 0:
        return out
    rc = _lib.fn("ude_pixel_scatter", dt)(bn, W, H, k, N, C, _lib.ptr(indices), _lib.ptr(values), _lib.ptr(out),
                                          _lib.stream(dev))
    _lib.check(rc, "pixelgather.scatter")
    return out


def gather(indices, values, numSources):
    """out (bn, numSources, C): the sum over every slot holding splat n of that
    slot's row of values (bn, W, H, k, C) -- the adjoint of scatter."""
    _lib.expect(indices, "indices", None, None, None, None)
    _lib.expect(values, "values", *indices.shape, None)
    dev = _lib.require_gpu(indices, values)
    indices = _lib.i32(indices, "indices")
    dt, (values,) = _lib.reals([("values", values)])
    bn, W, H, k = indices.shape
    N, C = int(numSources), values.shape[4]
    out = torch.zeros((bn, N, C), dtype=dt, device=dev)
    if values.numel() == 0 or N == 0:
        return out
    rc = _lib.fn("ude_pixel_gather", dt)(bn, W, H, k, N, C, _lib.ptr(indices), _lib.ptr(values), _lib.ptr(out),
                                         _lib.stream(dev))
    _lib.check(rc, "pixelgather.gather")
    return out
