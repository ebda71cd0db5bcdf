"""Loader of the gfx950 splatting library (libudsplat.so, C ABI in include/udsplat.h).

Every op of this package runs on the HIP library; there is no CPU or PyTorch
fallback.  A missing library, a non-GPU tensor or a failing kernel launch
raises.
"""
import ctypes
import os

import torch

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libudsplat.so")

_lib = None

UDE_OK = 0
_ERRORS = {-1: "invalid argument", -2: "kernel launch failed", -3: "workspace too small",
           -4: "sizes exceed the static scratch of the allocation-free entry point"}

c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_ptr = ctypes.c_void_p
c_size = ctypes.c_size_t

_SIGS = {
    "ude_sample_workspace_bytes": (c_size, [c_int] * 4),
    "ude_sample": (c_int, [c_int] * 5 + [c_dbl] * 3 + [c_ptr] * 3 + [c_int, ctypes.c_uint64, c_ptr, c_ptr, c_ptr, c_size, c_ptr]),
    "ude_fmm_weights_workspace_bytes": (c_size, [c_int] * 3),
    "ude_fmm_weights": (c_int, [c_int] * 4 + [c_dbl] * 3 + [c_ptr] * 3 + [c_int, c_ptr, c_ptr, c_size, c_ptr]),
    "ude_poisson_workspace_bytes": (c_size, [c_int] * 4),
    "ude_sample_poisson": (c_int, [c_int] * 5 + [c_dbl] * 4 + [c_ptr] * 4 + [c_int, ctypes.c_uint64] + [c_ptr] * 6 +
                           [c_size, c_ptr]),
    "ude_tile_features_workspace_bytes": (c_size, [c_int] * 2),
    "ude_tile_features": (c_int, [c_int] * 4 + [c_dbl] * 3 + [c_ptr] * 3 + [c_int] + [c_ptr] * 4 + [c_size, c_ptr]),
    "ude_mlp_param_count": (c_int, []),
    "ude_mlp_workspace_bytes": (c_size, [c_int]),
    "ude_mlp_forward": (c_int, [c_int] + [c_ptr] * 4 + [c_size, c_ptr]),
    "ude_mlp_loss_grad": (c_int, [c_int] + [c_ptr] * 7 + [c_size, c_ptr]),
    "ude_adam_step": (c_int, [c_int] + [c_ptr] * 4 + [ctypes.c_float] * 4 + [c_int, c_ptr]),
    "ude_set_poisson_proposal": (c_int, [c_int, c_dbl]),
    "ude_get_poisson_proposal": (c_int, [c_ptr]),
    "ude_set_fixed_degree": (c_int, [c_int]),
    "ude_get_fixed_degree": (c_int, []),
    "ude_set_tile_config": (c_int, [c_int] * 5),
    "ude_get_tile_config": (c_int, [c_int, c_ptr]),
    "ude_profile_sampler_begin": (c_int, [c_int]),
    "ude_profile_sampler_end": (c_int, [c_ptr, c_ptr]),
    "ude_covariance": (c_int, [c_int] * 4 + [c_dbl] * 2 + [c_ptr] * 3 + [c_int, c_ptr, c_ptr]),
    "ude_splat_workspace_bytes_f64": (c_size, [c_int] * 2),
    "ude_splat_forward_f64": (c_int, [c_int] * 5 + [c_dbl] * 2 + [c_ptr] * 4 + [c_int, c_ptr, c_ptr, c_size, c_ptr]),
    "ude_splat_backward_f64": (c_int, [c_int] * 5 + [c_dbl] * 2 + [c_ptr] * 5 + [c_int] + [c_ptr] * 4 + [c_size,
                                                                                                          c_ptr]),
    "ude_splat_backward_normalized_f64": (c_int, [c_int] * 5 + [c_dbl] * 2 + [c_ptr] * 9 + [c_size, c_ptr]),
    "ude_filter_forward_f64": (c_int, [c_int] * 5 + [c_dbl] + [c_ptr] * 4 + [c_int, c_ptr, c_ptr]),
    "ude_filter_backward_f64": (c_int, [c_int] * 5 + [c_dbl] + [c_ptr] * 5 + [c_int, c_ptr, c_ptr]),
    "ude_filter_forward_f16": (c_int, [c_int] * 5 + [c_dbl] + [c_ptr] * 4 + [c_int, c_ptr, c_ptr]),
    "ude_filter_backward_f16": (c_int, [c_int] * 5 + [c_dbl] + [c_ptr] * 5 + [c_int, c_ptr, c_ptr]),
    "ude_transform_forward_f64": (c_int, [c_int] * 4 + [c_ptr] * 4),
    "ude_transform_backward_f64": (c_int, [c_int] * 4 + [c_ptr] * 6),
    "ude_interp_forward_f64": (c_int, [c_int] * 7 + [c_ptr] * 5),
    "ude_interp_backward_f64": (c_int, [c_int] * 7 + [c_ptr] * 7),
    "ude_sh_forward_f64": (c_int, [c_int] * 4 + [c_ptr] * 5),
    "ude_sh_backward_f64": (c_int, [c_int] * 4 + [c_ptr] * 8),
    "ude_lambert_forward_f64": (c_int, [c_int] * 5 + [c_ptr] * 6),
    "ude_lambert_backward_f64": (c_int, [c_int] * 5 + [c_ptr] * 10),
    "ude_splat_workspace_bytes": (c_size, [c_int] * 2),
    "ude_splat_forward": (c_int, [c_int] * 5 + [c_dbl] * 2 + [c_ptr] * 4 + [c_int, c_ptr, c_ptr, c_size, c_ptr]),
    "ude_splat_backward": (c_int, [c_int] * 5 + [c_dbl] * 2 + [c_ptr] * 5 + [c_int] + [c_ptr] * 4 + [c_size, c_ptr]),
    "ude_splat_backward_static": (c_int, [c_int] * 5 + [c_dbl] * 2 + [c_ptr] * 5 + [c_int] + [c_ptr] * 4),
    "ude_splat_backward_normalized": (c_int, [c_int] * 5 + [c_dbl] * 2 + [c_ptr] * 9 + [c_size, c_ptr]),
    "ude_filter_forward": (c_int, [c_int] * 5 + [c_dbl] + [c_ptr] * 4 + [c_int, c_ptr, c_ptr]),
    "ude_filter_backward": (c_int, [c_int] * 5 + [c_dbl] + [c_ptr] * 5 + [c_int, c_ptr, c_ptr]),
    "ude_filter_backward_exact": (c_int, [c_int] * 5 + [c_dbl] + [c_ptr] * 5 + [c_int, c_ptr, c_ptr]),
    "ude_filter_backward_exact_f64": (c_int, [c_int] * 5 + [c_dbl] + [c_ptr] * 5 + [c_int, c_ptr, c_ptr]),
    "ude_transform_forward": (c_int, [c_int] * 4 + [c_ptr] * 4),
    "ude_transform_forward_facing": (c_int, [c_int] * 3 + [c_ptr] * 5),
    "ude_transform_backward": (c_int, [c_int] * 4 + [c_ptr] * 6),
    "ude_interp_forward": (c_int, [c_int] * 7 + [c_ptr] * 5),
    "ude_interp_backward": (c_int, [c_int] * 7 + [c_ptr] * 7),
    "ude_ht_workspace_bytes": (c_size, [c_int] * 2),
    "ude_ht_forward": (c_int, [c_int] * 5 + [c_dbl] * 2 + [c_ptr] * 5 + [c_int] * 2 + [c_ptr] * 3 + [c_size, c_ptr]),
    "ude_ht_backward": (c_int, [c_int] * 5 + [c_dbl] * 2 + [c_ptr] * 6 + [c_int] * 2 + [c_ptr] * 5 + [c_int, c_ptr,
                                                                                                    c_size, c_ptr]),
    "ude_htc_workspace_bytes": (c_size, [c_int] * 5),
    "ude_htc_forward": (c_int, [c_int] * 5 + [c_dbl] * 2 + [c_ptr] * 5 + [c_int] * 2 + [c_ptr] * 3 + [c_size, c_ptr]),
    "ude_htc_backward": (c_int, [c_int] * 5 + [c_dbl] * 2 + [c_ptr] * 6 + [c_int] * 2 + [c_ptr] * 5 + [c_int, c_ptr,
                                                                                                     c_size, c_ptr]),
    "ude_l2_grad": (c_int, [ctypes.c_long] + [c_ptr] * 5),
    "ude_pixel_scatter": (c_int, [c_int] * 6 + [c_ptr] * 4),
    "ude_pixel_scatter_f64": (c_int, [c_int] * 6 + [c_ptr] * 4),
    "ude_pixel_scatter_f16": (c_int, [c_int] * 6 + [c_ptr] * 4),
    "ude_pixel_gather": (c_int, [c_int] * 6 + [c_ptr] * 4),
    "ude_pixel_gather_f64": (c_int, [c_int] * 6 + [c_ptr] * 4),
    "ude_nearest_workspace_bytes": (c_size, [c_int] * 2),
    "ude_nearest_samples": (c_int, [c_int] * 6 + [c_ptr] * 5 + [c_size, c_ptr]),
    "ude_nearest_samples_f64": (c_int, [c_int] * 6 + [c_ptr] * 5 + [c_size, c_ptr]),
    "ude_sh_forward": (c_int, [c_int] * 4 + [c_ptr] * 5),
    "ude_sh_backward": (c_int, [c_int] * 4 + [c_ptr] * 8),
    "ude_lambert_forward": (c_int, [c_int] * 5 + [c_ptr] * 6),
    "ude_lambert_backward": (c_int, [c_int] * 5 + [c_ptr] * 10),
}


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libudsplat.so not found at %s; build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C unbiased-differentiable-splatting_amd/csrc`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what):
    if rc != UDE_OK:
        raise RuntimeError("%s failed: %s (%d)" % (what, _ERRORS.get(rc, "error"), rc))


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


def stream(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def expect(t, name, *dims):
    """ValueError unless ``t`` has len(dims) axes of the given sizes (None: any
    size, a tuple: one of its sizes).  Every op checks the shapes its kernels
    index by before it launches them, so a wrong shape raises here instead of
    reading out of bounds on the device."""
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s: expected a torch.Tensor" % name)
    ok = t.dim() == len(dims)
    if ok:
        for got, want in zip(t.shape, dims):
            if want is None:
                continue
            if isinstance(want, tuple):
                if int(got) not in want:
                    ok = False
            elif int(got) != int(want):
                ok = False
    if not ok:
        fmt = ", ".join("*" if d is None else ("|".join(str(x) for x in sorted(set(d))) if isinstance(d, tuple)
                                                 else str(d)) for d in dims)
        raise ValueError("%s must have shape (%s), got %s" % (name, fmt, tuple(t.shape)))
    return t


def require_gpu(*tensors, dtype=torch.float32):
    dev = None
    for t in tensors:
        if not isinstance(t, torch.Tensor):
            raise TypeError("expected a torch.Tensor")
        if not t.is_cuda:
            raise RuntimeError("splatting ops run on the GPU only: got a tensor on %s" % t.device)
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise RuntimeError("all tensors must be on the same device")
    return dev


SUFFIX = {torch.float32: "", torch.float64: "_f64", torch.float16: "_f16"}


def reals(named, allowed=(torch.float32, torch.float64)):
    """The common dtype of the (name, tensor) pairs, which must be one of
    ``allowed`` (the reference dispatches float and double, filtering also
    half), and the tensors made contiguous.  Returns (dtype, [tensors])."""
    dt = named[0][1].dtype
    if dt not in allowed:
        raise TypeError("%s must be one of %s (got %s)" % (named[0][0], ", ".join(str(a) for a in allowed), dt))
    for name, t in named[1:]:
        if t.dtype != dt:
            raise TypeError("%s must have the dtype of %s (%s), got %s" % (name, named[0][0], dt, t.dtype))
    return dt, [t.contiguous() for _, t in named]


def fn(name, dtype):
    """The entry point of ``name`` for ``dtype`` (ude_x, ude_x_f64 or ude_x_f16)."""
    return getattr(lib(), name + SUFFIX[dtype])


def f32(t, name):
    if t.dtype != torch.float32:
        raise TypeError("%s must be float32 (got %s)" % (name, t.dtype))
    return t.contiguous()


def i32(t, name):
    if t.dtype != torch.int32:
        raise TypeError("%s must be int32 (got %s)" % (name, t.dtype))
    return t.contiguous()


def workspace(nbytes, device):
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
