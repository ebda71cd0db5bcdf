"""The allocation-free splatting.backward (ude_splat_backward_static), called
through the C ABI: its gradients equal ude_splat_backward's (through
splatting.backward) to the order of their atomic sums -- elementwise within
1e-5 relative plus 64 float32 roundoffs of each output's magnitude -- on two
calls in a row, the second with fewer views, splats and pixels (the static
scratch keeps the first call's contents, which must not leak), and sizes beyond
its static scratch are refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import assert_close, magnitude, near_slots, scene_inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def static_backward(W, H, f, g, idx, P, Nn, S):
    """ude_splat_backward_static through ctypes; outputs allocated by the caller."""
    from cudaExtensions import _lib
    bn, N = P.shape[0], P.shape[1]
    gp = torch.empty_like(P)
    gn = torch.empty_like(Nn)
    gs = (torch.empty_like if S.shape[0] == bn else torch.zeros_like)(S)
    dev = P.device
    rc = _lib.lib().ude_splat_backward_static(bn, N, W, H, idx.shape[3], float(f), 0.0, _lib.ptr(g), _lib.ptr(idx),
                                              _lib.ptr(P), _lib.ptr(Nn), _lib.ptr(S), S.shape[0], _lib.ptr(gp),
                                              _lib.ptr(gn), _lib.ptr(gs), _lib.stream(dev))
    _lib.check(rc, "ude_splat_backward_static")
    return gp, gn, gs


def case(cfg, views, sd_scale, k, seed):
    sc, p, n, nb, sd = scene_inputs(cfg, sd_scale=sd_scale)
    rng = np.random.default_rng(seed)
    P = np.stack([p * np.float32([1.0, (-1.0) ** v, 1.0]) for v in range(views)])
    Nn = np.stack([n * np.float32([1.0, (-1.0) ** v, 1.0]) for v in range(views)])
    idx = np.concatenate([near_slots(rng, P[v], sc.width, sc.height, sc.focal, k) for v in range(views)])
    g = rng.normal(size=idx.shape).astype(np.float32)
    return sc, P, Nn, sd[None], idx, g


def test_static_scratch_backward_equals_workspace_backward(dev):
    from cudaExtensions import splatting
    # first the larger call (4096 splats, 2 views, 128x128), then a smaller one (256 splats, 1 view, 64x64)
    for cfg, views, sds, k, seed in ((0, 2, 4.0, 16, 1), (2, 1, 8.0, 12, 2)):
        sc, P, Nn, S, idx, g = case(cfg, views, sds, k, seed)
        W, H, f = sc.width, sc.height, sc.focal
        args = [t(x, dev) for x in (g, idx, P, Nn, S)]
        got = static_backward(W, H, f, *args)
        ref = splatting.backward(W, H, f, *args, 0.0)
        mags = magnitude(O.splat_backward, W, H, f, g, idx, P, Nn, S)
        for a, b, m, nm in zip(got, ref, mags, ("points", "normals", "std-devs")):
            assert_close(a.cpu().numpy(), b.cpu().numpy(), 1e-5, "static scratch %s (%d splats)" % (nm, P.shape[1]),
                         mag=m)
        assert float(ref[0].abs().max()) > 0


def test_static_scratch_refuses_oversized_calls(dev):
    from cudaExtensions import _lib
    L = _lib.lib()
    N = 1 << 20                       # 1M splats x 2 views x 64 B > 64 MiB: refused before any launch
    rc = L.ude_splat_backward_static(2, N, 8, 8, 4, 10.0, 0.0, *(ctypes.c_void_p(0),) * 5, 1,
                                     *(ctypes.c_void_p(0),) * 3, _lib.stream(dev))
    assert rc == -4
    with pytest.raises(RuntimeError, match="static scratch"):
        _lib.check(rc, "ude_splat_backward_static")
