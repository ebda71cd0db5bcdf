"""ctypes front end of the CPU oracle (oracle/ud_oracle.c).

Test infrastructure only: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product.  All functions take and
return numpy arrays; FP64 restatements take float64 inputs, the FP32 sampler
mirror takes float32 inputs.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.o_expf.restype = ctypes.c_float
        _lib.o_expf.argtypes = [ctypes.c_float]
        _lib.o_logf.restype = ctypes.c_float
        _lib.o_logf.argtypes = [ctypes.c_float]
        _lib.o_hermite_error.restype = ctypes.c_float
        _lib.o_hermite_error.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_int]
        _lib.o_morton2.restype = ctypes.c_uint32
        _lib.o_morton2.argtypes = [ctypes.c_float, ctypes.c_float]
        _lib.o_uniform_export.restype = ctypes.c_float
        _lib.o_uniform_export.argtypes = [ctypes.c_uint32] * 6
        _lib.o_num_launches_export.restype = ctypes.c_int
        _lib.o_num_launches_export.argtypes = [ctypes.c_int]
        _lib.o_tree_export.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_void_p] * 10 + [ctypes.c_long, ctypes.c_void_p]
        _lib.o_tree_eval.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float] + [ctypes.c_void_p] * 5 + [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5
        _lib.o_sample_view.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 3 + [ctypes.c_void_p] * 3 + [ctypes.c_uint32] * 3 + [ctypes.c_void_p] * 3
        _lib.o_fmm_weights_view.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_void_p] * 4
        _lib.o_poisson_view.argtypes = ([ctypes.c_int] * 4 + [ctypes.c_double] * 4 + [ctypes.c_void_p] * 4 +
                                       [ctypes.c_uint32] * 3 + [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_double])
        _lib.o_foot_view.argtypes = ([ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_void_p] * 3 +
                                    [ctypes.c_void_p] * 4)
        _lib.o_foot_view.restype = ctypes.c_int
        _lib.o_foot_terms.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_float] + [ctypes.c_void_p] * 2
        _lib.o_foot_bounds.argtypes = [ctypes.c_int, ctypes.c_float] + [ctypes.c_void_p] * 2
        _lib.o_set_fixed_degree.restype = ctypes.c_int
        _lib.o_set_fixed_degree.argtypes = [ctypes.c_int]
        _lib.o_poisson_wide_h.restype = ctypes.c_float
        _lib.o_poisson_wide_h.argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_void_p] * 3
        _lib.o_set_abs_terms.argtypes = [ctypes.c_int]
        _lib.o_sample_view_rows.argtypes = ([ctypes.c_int] * 4 + [ctypes.c_double] * 3 + [ctypes.c_void_p] * 3 +
                                           [ctypes.c_uint32] * 3 + [ctypes.c_void_p] * 3 + [ctypes.c_int] * 2 +
                                           [ctypes.c_void_p])
    return _lib


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops (the runtime reads
    OMP_NUM_THREADS only once, so later changes go through this call)."""
    lib().omp_set_num_threads(int(n))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------- scalars
def expf(x):
    return lib().o_expf(float(x))


def logf(x):
    return lib().o_logf(float(x))


def hermite_error(r, n, p):
    return lib().o_hermite_error(float(r), float(n), int(p))


def morton2(ux, uy):
    return lib().o_morton2(float(ux), float(uy))


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().o_philox4x32_10(c, k, o)
    return list(o)


def uniform(seed, view, pixel, launch, ordinal):
    return lib().o_uniform_export(seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF, view, pixel, launch, ordinal)


def num_launches(k):
    return lib().o_num_launches_export(int(k))


class abs_terms:
    """Inside this block the per-slot operators return, per output element, the
    sum of the absolute values of the terms they accumulate instead of the sum
    (the magnitude an FP32 accumulation of the same terms is measured against)."""

    def __enter__(self):
        lib().o_set_abs_terms(1)
        return self

    def __exit__(self, *exc):
        lib().o_set_abs_terms(0)
        return False


# ---------------------------------------------------------------- per-slot ops (FP64)
def splat_forward(W, H, f, idx, pts, nrm, sd, smoothing=0.0):
    idx = _c(idx, np.int32); pts = _c(pts, np.float64); nrm = _c(nrm, np.float64); sd = _c(sd, np.float64)
    bn, _, _, k = idx.shape
    N = pts.shape[1]
    out = np.zeros(idx.shape, np.float64)
    L = lib()
    L.o_splat_forward(ctypes.c_int(bn), ctypes.c_int(N), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(k),
                      ctypes.c_double(f), ctypes.c_double(smoothing), _p(idx), _p(pts), _p(nrm), _p(sd),
                      ctypes.c_int(sd.shape[0]), _p(out))
    return out


def splat_backward(W, H, f, gin, idx, pts, nrm, sd, smoothing=0.0):
    idx = _c(idx, np.int32); pts = _c(pts, np.float64); nrm = _c(nrm, np.float64); sd = _c(sd, np.float64)
    gin = _c(gin, np.float64)
    bn, _, _, k = idx.shape
    N = pts.shape[1]
    gp = np.zeros(pts.shape, np.float64); gn = np.zeros(nrm.shape, np.float64); gs = np.zeros(sd.shape, np.float64)
    lib().o_splat_backward(ctypes.c_int(bn), ctypes.c_int(N), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(k),
                           ctypes.c_double(f), ctypes.c_double(smoothing), _p(gin), _p(idx), _p(pts), _p(nrm),
                           _p(sd), ctypes.c_int(sd.shape[0]), _p(gp), _p(gn), _p(gs))
    return gp, gn, gs


def covariance(W, H, f, pts, nrm, sd, smoothing=0.0):
    pts = _c(pts, np.float64); nrm = _c(nrm, np.float64); sd = _c(sd, np.float64)
    bn, N, _ = pts.shape
    out = np.zeros((bn, N, 2, 2), np.float64)
    lib().o_covariance(ctypes.c_int(bn), ctypes.c_int(N), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_double(f),
                       ctypes.c_double(smoothing), _p(pts), _p(nrm), _p(sd), ctypes.c_int(sd.shape[0]), _p(out))
    return out


def filter_forward(idx, w, W, H, f, pts, nrm):
    """pts/nrm float32 (depth ordering is computed in FP32 like the GPU)."""
    idx = _c(idx, np.int32); w = _c(w, np.float64); pts = _c(pts, np.float32); nrm = _c(nrm, np.float32)
    bn, _, _, k = idx.shape
    out = np.zeros(idx.shape, np.float64)
    rc = lib().o_filter_forward(ctypes.c_int(bn), ctypes.c_int(pts.shape[1]), ctypes.c_int(W), ctypes.c_int(H),
                                ctypes.c_int(k), ctypes.c_double(f), _p(idx), _p(w), _p(pts), _p(nrm),
                                ctypes.c_int(pts.shape[0]), _p(out))
    assert rc == 0
    return out


def filter_backward(idx, w, gin, W, H, f, pts, nrm, exact=False):
    """filtering.backward; exact=True: the exact derivative (no dropped terms)."""
    idx = _c(idx, np.int32); w = _c(w, np.float64); gin = _c(gin, np.float64)
    pts = _c(pts, np.float32); nrm = _c(nrm, np.float32)
    bn, _, _, k = idx.shape
    out = np.zeros(idx.shape, np.float64)
    rc = (lib().o_filter_backward_exact if exact else lib().o_filter_backward)(ctypes.c_int(bn), ctypes.c_int(pts.shape[1]), ctypes.c_int(W), ctypes.c_int(H),
                                 ctypes.c_int(k), ctypes.c_double(f), _p(idx), _p(w), _p(gin), _p(pts), _p(nrm),
                                 ctypes.c_int(pts.shape[0]), _p(out))
    assert rc == 0
    return out


def htc(W, H, f, idx, ipi, pts, nrm, sd, attr, gimg=None, smoothing=0.0):
    """Composited Horvitz-Thompson estimate (FP64) and, with gimg (bn, W, H, 3), its
    exact adjoint.  pts / nrm float32 (bn, N, 3) camera space, sd (1|bn, N, 2),
    attr (1|bn, N, 3).  Returns image, or (image, (gP, gN, gS, gA))."""
    idx = _c(idx, np.int32); ipi = _c(ipi, np.float64)
    pts = _c(pts, np.float32); nrm = _c(nrm, np.float32); sd = _c(sd, np.float64); attr = _c(attr, np.float64)
    bn, _, _, k = idx.shape
    N = pts.shape[1]
    out = np.zeros((bn, W, H, 3), np.float64)
    gP = np.zeros(pts.shape, np.float64); gN = np.zeros(nrm.shape, np.float64)
    gS = np.zeros(sd.shape, np.float64); gA = np.zeros(attr.shape, np.float64)
    g = None if gimg is None else _c(gimg, np.float64)
    rc = lib().o_htc(ctypes.c_int(bn), ctypes.c_int(N), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(k),
                     ctypes.c_double(f), ctypes.c_double(smoothing), _p(idx), _p(ipi), _p(pts), _p(nrm), _p(sd),
                     ctypes.c_int(sd.shape[0]), ctypes.c_int(attr.shape[0]), _p(attr),
                     None if g is None else _p(g), _p(out), _p(gP), _p(gN), _p(gS), _p(gA))
    assert rc == 0
    return out if gimg is None else (out, (gP, gN, gS, gA))


def pixel_scatter(idx, values):
    """pixelgather.scatter (pixelGatherCuda.cu:7-35): out[b, i, j, s, :] =
    values[b, idx[b, i, j, s], :] where idx >= 0, else 0."""
    idx = np.asarray(idx)
    values = np.asarray(values)
    bn = idx.shape[0]
    out = np.zeros(idx.shape + (values.shape[2],), values.dtype)
    for b in range(bn):
        m = idx[b] >= 0
        out[b][m] = values[b][idx[b][m]]
    return out


def pixel_gather(idx, values, N):
    """pixelgather.gather (pixelGatherCuda.cu:76-109): out[b, n, :] = sum of
    values[b, i, j, s, :] over the slots with idx[b, i, j, s] == n, FP64."""
    idx = np.asarray(idx)
    values = np.asarray(values, np.float64)
    bn = idx.shape[0]
    out = np.zeros((bn, N, values.shape[4]), np.float64)
    for b in range(bn):
        m = idx[b] >= 0
        np.add.at(out[b], idx[b][m], values[b][m])
    return out


def nearest_samples(idx, w, pts, K):
    """nearestsamples.nearest (nearestSamplesCuda.cu:290-404): per pixel the surface
    splat is the first slot of largest weight; every other splat sampled there is a
    neighbour candidate of it at float32 squared distance ((dx dx + dy dy) + dz dz);
    per splat the K nearest distinct candidates by (distance, index), -1 padded.
    Returns (1, N, K) int32."""
    idx = np.asarray(idx)
    w = np.asarray(w)
    P = np.asarray(pts)[0]
    N = P.shape[0]
    flat_i = idx.reshape(-1, idx.shape[-1])
    flat_w = w.reshape(-1, w.shape[-1])
    surf = flat_i[np.arange(flat_i.shape[0]), np.argmax(flat_w, axis=1)]
    cand = {}
    for s, row in zip(surf, flat_i):
        if s < 0:
            continue
        for c in row:
            if c >= 0 and c != s:
                cand.setdefault(int(s), set()).add(int(c))
    out = np.full((1, N, K), -1, np.int32)
    for s, cs in cand.items():
        cs = np.array(sorted(cs))
        d = (P[s][None, :] - P[cs]).astype(np.float32)
        d2 = d[:, 0] * d[:, 0]
        d2 = (d2 + d[:, 1] * d[:, 1]).astype(np.float32)
        d2 = (d2 + d[:, 2] * d[:, 2]).astype(np.float32)
        order = np.lexsort((cs, d2))[:K]
        out[0, s, :len(order)] = cs[order]
    return out


def direct_sums(screen, h, c, queries):
    """FP64 direct Gaussian sums c sum_i exp(-|x - s_i|^2 / 2h^2) at each query x
    over the screen points s (n, 2) float32 (OpenMP)."""
    s = _c(screen, np.float32)
    x = _c(queries, np.float64)
    out = np.zeros(x.shape[0], np.float64)
    lib().o_direct_sums(ctypes.c_long(s.shape[0]), _p(s), ctypes.c_double(h), ctypes.c_double(c),
                        ctypes.c_long(x.shape[0]), _p(x), _p(out))
    return out


def transform_forward(affine, M, v):
    M = _c(M, np.float64); v = _c(v, np.float64)
    bn = M.shape[0]; vb, N, _ = v.shape
    out = np.zeros((bn, N, 3), np.float64)
    lib().o_transform_forward(ctypes.c_int(0 if affine else 1), ctypes.c_int(bn), ctypes.c_int(vb), ctypes.c_int(N),
                              _p(M), _p(v), _p(out))
    return out


def transform_backward(affine, g, M, v):
    g = _c(g, np.float64); M = _c(M, np.float64); v = _c(v, np.float64)
    bn = M.shape[0]; vb, N, _ = v.shape
    gM = np.zeros(M.shape, np.float64); gV = np.zeros(v.shape, np.float64)
    lib().o_transform_backward(ctypes.c_int(0 if affine else 1), ctypes.c_int(bn), ctypes.c_int(vb), ctypes.c_int(N),
                               _p(g), _p(M), _p(v), _p(gM), _p(gV))
    return gM, gV


def interp_forward(idx, w, attr):
    idx = _c(idx, np.int32); w = _c(w, np.float64); attr = _c(attr, np.float64)
    bn, W, H, k = idx.shape
    ab, N, C = attr.shape
    out = np.zeros((bn, W, H, C), np.float64)
    lib().o_interp_forward(ctypes.c_int(bn), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(k), ctypes.c_int(ab),
                           ctypes.c_int(N), ctypes.c_int(C), _p(idx), _p(w), _p(attr), _p(out))
    return out


def interp_backward(g, idx, w, attr):
    g = _c(g, np.float64); idx = _c(idx, np.int32); w = _c(w, np.float64); attr = _c(attr, np.float64)
    bn, W, H, k = idx.shape
    ab, N, C = attr.shape
    gW = np.zeros(idx.shape, np.float64); gA = np.zeros(attr.shape, np.float64)
    lib().o_interp_backward(ctypes.c_int(bn), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(k), ctypes.c_int(ab),
                            ctypes.c_int(N), ctypes.c_int(C), _p(g), _p(idx), _p(w), _p(attr), _p(gW), _p(gA))
    return gW, gA


def sh_forward(coef, nrm, diff):
    coef = _c(coef, np.float64); nrm = _c(nrm, np.float64); diff = _c(diff, np.float64)
    bn, W, H, _ = nrm.shape
    out = np.zeros(nrm.shape, np.float64)
    lib().o_sh_forward(ctypes.c_int(bn), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(coef.shape[0]),
                       _p(coef), _p(nrm), _p(diff), _p(out))
    return out


def sh_backward(g, coef, nrm, diff):
    g = _c(g, np.float64); coef = _c(coef, np.float64); nrm = _c(nrm, np.float64); diff = _c(diff, np.float64)
    bn, W, H, _ = nrm.shape
    gc = np.zeros(coef.shape, np.float64); gn = np.zeros(nrm.shape, np.float64); gd = np.zeros(diff.shape, np.float64)
    lib().o_sh_backward(ctypes.c_int(bn), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(coef.shape[0]),
                        _p(g), _p(coef), _p(nrm), _p(diff), _p(gc), _p(gn), _p(gd))
    return gc, gn, gd


def lambert_forward(dirs, ints, nrm, diff):
    dirs = _c(dirs, np.float64); ints = _c(ints, np.float64); nrm = _c(nrm, np.float64); diff = _c(diff, np.float64)
    bn, W, H, _ = nrm.shape
    out = np.zeros(nrm.shape, np.float64)
    lib().o_lambert_forward(ctypes.c_int(bn), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(dirs.shape[1]),
                            ctypes.c_int(ints.shape[0]), _p(dirs), _p(ints), _p(nrm), _p(diff), _p(out))
    return out


def lambert_backward(g, dirs, ints, nrm, diff):
    g = _c(g, np.float64); dirs = _c(dirs, np.float64); ints = _c(ints, np.float64)
    nrm = _c(nrm, np.float64); diff = _c(diff, np.float64)
    bn, W, H, _ = nrm.shape
    gdir = np.zeros(dirs.shape, np.float64); gint = np.zeros(ints.shape, np.float64)
    gn = np.zeros(nrm.shape, np.float64); gd = np.zeros(diff.shape, np.float64)
    lib().o_lambert_backward(ctypes.c_int(bn), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(dirs.shape[1]),
                             ctypes.c_int(ints.shape[0]), _p(g), _p(dirs), _p(ints), _p(nrm), _p(diff),
                             _p(gdir), _p(gint), _p(gn), _p(gd))
    return gdir, gint, gn, gd


# ---------------------------------------------------------------- sampler mirror (FP32)
class Tree:
    """Exported oracle tree of one view (for stage-by-stage comparisons)."""

    def __init__(self, W, H, f, pts, nrm, sd, smoothing=0.0, precision=0.01):
        pts = _c(pts, np.float32); nrm = _c(nrm, np.float32); sd = _c(sd, np.float32)
        N = pts.shape[0]
        self.N = N
        nn = 2 * max(N, 1)
        self.hk = np.zeros(max(N, 1), np.float32)
        self.params = np.zeros(8, np.float32)
        self.morton = np.zeros(max(N, 1), np.uint32)
        self.ids = np.zeros(max(N, 1), np.int32)
        self.ibuf = np.zeros(8 * nn, np.int32)
        self.fbuf = np.zeros(7 * nn, np.float32)
        cap = 290 * max(N, 1)
        self.coef_all = np.zeros(cap, np.float32)
        ncoef = ctypes.c_long(0)
        rc = lib().o_tree_export(ctypes.c_int(N), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_double(f),
                                 ctypes.c_double(smoothing), ctypes.c_double(precision), _p(pts), _p(nrm), _p(sd),
                                 _p(self.hk), _p(self.params), _p(self.morton), _p(self.ids), _p(self.ibuf),
                                 _p(self.fbuf), _p(self.coef_all), ctypes.c_long(cap), ctypes.byref(ncoef))
        assert rc >= 0, rc
        self.nc = rc
        self.ncoef = ncoef.value
        self.h, self.radius, self.lo, self.hix, self.hiy, self.s2h = [float(x) for x in self.params[:6]]
        self.precision = precision
        nc = self.nc
        nn_used = max(2 * nc - 1, 0)
        ib = self.ibuf.reshape(8, nn)
        self.left, self.right, self.parent, self.count = (ib[q, :nn_used] for q in range(4))
        self.deg, self.off, self.first, self.last = (ib[q, :nn_used] for q in range(4, 8))
        fb = self.fbuf
        self.box = fb[:4 * nn].reshape(nn, 4)[:nn_used]
        self.cen = fb[4 * nn:6 * nn].reshape(nn, 2)[:nn_used]
        self.r2 = fb[6 * nn:7 * nn][:nn_used]
        self.coef = self.coef_all[:self.ncoef]
        self.morton_sorted = self.morton[:nc]
        self.ids_sorted = self.ids[:nc]

    def weight(self, nodes, qx, qy):
        """FGT weight W(node, q) (FP32 mirror) and the FP64 direct Gaussian sum."""
        nodes = _c(nodes, np.int32); qx = _c(qx, np.float32); qy = _c(qy, np.float32)
        nq = nodes.shape[0]
        w = np.zeros(nq, np.float32); d = np.zeros(nq, np.float64)
        lib().o_tree_eval(ctypes.c_int(self.nc), ctypes.c_float(self.h), ctypes.c_float(self.precision),
                          ctypes.c_float(self.s2h), _p(self.morton), _p(self.ids), _p(self.ibuf), _p(self.fbuf),
                          _p(self.coef_all), ctypes.c_int(self.N), ctypes.c_int(nq), _p(nodes), _p(qx), _p(qy),
                          _p(w), _p(d))
        return w, d


def sample_view(W, H, f, pts, nrm, sd, k, smoothing=0.0, precision=0.01, seed=0, view=0):
    """sampling.sample for one view: int32 (W, H, k)."""
    pts = _c(pts, np.float32); nrm = _c(nrm, np.float32); sd = _c(sd, np.float32)
    N = pts.shape[0]
    out = np.zeros((W, H, k), np.int32)
    h = ctypes.c_float(0.0)
    nc = ctypes.c_int(0)
    rc = lib().o_sample_view(ctypes.c_int(N), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(k), ctypes.c_double(f),
                             ctypes.c_double(smoothing), ctypes.c_double(precision), _p(pts), _p(nrm), _p(sd),
                             ctypes.c_uint32(seed & 0xFFFFFFFF), ctypes.c_uint32((seed >> 32) & 0xFFFFFFFF),
                             ctypes.c_uint32(view), _p(out), ctypes.byref(h), ctypes.byref(nc))
    if rc != 0:
        raise RuntimeError("oracle sample_view failed: %d" % rc)
    return out


def sample_view_rows(W, H, f, pts, nrm, sd, k, i0, i1, smoothing=0.0, precision=0.01, seed=0, view=0):
    """sampling.sample restricted to pixel rows i0 <= x < i1; returns (indices, tree build seconds)."""
    pts = _c(pts, np.float32); nrm = _c(nrm, np.float32); sd = _c(sd, np.float32)
    N = pts.shape[0]
    out = np.zeros((W, H, k), np.int32)
    h = ctypes.c_float(0.0)
    nc = ctypes.c_int(0)
    tt = ctypes.c_double(0.0)
    rc = lib().o_sample_view_rows(ctypes.c_int(N), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(k),
                                  ctypes.c_double(f), ctypes.c_double(smoothing), ctypes.c_double(precision),
                                  _p(pts), _p(nrm), _p(sd), ctypes.c_uint32(seed & 0xFFFFFFFF),
                                  ctypes.c_uint32((seed >> 32) & 0xFFFFFFFF), ctypes.c_uint32(view), _p(out),
                                  ctypes.byref(h), ctypes.byref(nc), ctypes.c_int(i0), ctypes.c_int(i1),
                                  ctypes.byref(tt))
    if rc != 0:
        raise RuntimeError("oracle sample_view_rows failed: %d" % rc)
    return out, tt.value


def tie_pixels():
    """Pixels of the last sample / sample_view(_rows) call whose sorted key list
    met two equal keys (where the reference's insertion order differs from
    sorting by (key, draw number))."""
    lib().o_tie_pixels.restype = ctypes.c_long
    return int(lib().o_tie_pixels())


def sample(W, H, f, pts, nrm, sd, k, smoothing=0.0, precision=0.01, seed=0):
    """sampling.sample for a batch: pts/nrm (bn,N,3), sd (1|bn,N,2) -> (bn,W,H,k)."""
    bn = pts.shape[0]
    outs = []
    for b in range(bn):
        sdb = sd[b] if sd.shape[0] > 1 else sd[0]
        outs.append(sample_view(W, H, f, pts[b], nrm[b], sdb, k, smoothing, precision, seed, b))
    return np.stack(outs)


def set_fixed_degree(p):
    """Degree of the fixed-order expansions (densities of approximateFMMWeights and
    of Poisson-rate sampling): 4 .. 8, default 6."""
    if lib().o_set_fixed_degree(ctypes.c_int(int(p))) != 0:
        raise ValueError("fixed degree must be in 4 .. 8")


def fmm_weights_view(W, H, f, pts, nrm, sd, smoothing=0.0, precision=0.01):
    """sampling.approximateFMMWeights for one view: float32 (W, H)."""
    pts = _c(pts, np.float32); nrm = _c(nrm, np.float32); sd = _c(sd, np.float32)
    out = np.zeros((W, H), np.float32)
    rc = lib().o_fmm_weights_view(ctypes.c_int(pts.shape[0]), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_double(f),
                                  ctypes.c_double(smoothing), ctypes.c_double(precision), _p(pts), _p(nrm), _p(sd),
                                  _p(out))
    if rc != 0:
        raise RuntimeError("oracle fmm_weights_view failed: %d" % rc)
    return out


def fmm_weights(W, H, f, pts, nrm, sd, smoothing=0.0, precision=0.01):
    """sampling.approximateFMMWeights for a batch: (bn, W, H)."""
    outs = []
    for b in range(pts.shape[0]):
        sdb = sd[b] if sd.shape[0] > 1 else sd[0]
        outs.append(fmm_weights_view(W, H, f, pts[b], nrm[b], sdb, smoothing, precision))
    return np.stack(outs)


PROPOSALS = {"bands": 0, "footprint": 1}


def poisson_view(W, H, f, pts, nrm, sd, kcap, rate, smoothing=0.0, precision=0.01, seed=0, view=0, tile_scale=None,
                 proposal="bands", footprint_scale=0.0):
    """Poisson-rate sampling of one view: (ids (W,H,kcap) int32, 1/pi (W,H,kcap) float32,
    densities (D0, D1) (2,W,H) float32, overflow flag).  tile_scale: optional
    (ceil(W/16), ceil(H/16)) rate multipliers.  proposal "footprint": the wide
    band plus each splat's own footprint, normalised by footprint_scale (> 0) or
    by the view's mean footprint mass over h_1^1.5."""
    pts = _c(pts, np.float32); nrm = _c(nrm, np.float32); sd = _c(sd, np.float32)
    ts = None
    if tile_scale is not None:
        ts = _c(tile_scale, np.float32)
        assert ts.shape == ((W + 15) // 16, (H + 15) // 16)
    ids = np.zeros((W, H, kcap), np.int32)
    ipi = np.zeros((W, H, kcap), np.float32)
    dens = np.zeros((2, W, H), np.float32)
    rc = lib().o_poisson_view(ctypes.c_int(pts.shape[0]), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_int(kcap),
                              ctypes.c_double(f), ctypes.c_double(smoothing), ctypes.c_double(precision),
                              ctypes.c_double(rate), _p(ts) if ts is not None else None, _p(pts), _p(nrm), _p(sd),
                              ctypes.c_uint32(seed & 0xFFFFFFFF), ctypes.c_uint32((seed >> 32) & 0xFFFFFFFF),
                              ctypes.c_uint32(view), _p(ids), _p(ipi), _p(dens),
                              ctypes.c_int(PROPOSALS[proposal]), ctypes.c_double(footprint_scale))
    if rc < 0:
        raise RuntimeError("oracle poisson_view failed: %d" % rc)
    return ids, ipi, dens, rc


def foot_view(W, H, f, pts, nrm, sd, smoothing=0.0, precision=0.01):
    """The footprint proposal's records of one view's sorted leaves: (leaf (nc, 8)
    float32: x, y, t00, t01, t11, amp, 0, 0; sorted splat ids (nc,) int32; parent
    links (2 nc - 1,) int32; total footprint mass; the cap's decay a1)."""
    pts = _c(pts, np.float32); nrm = _c(nrm, np.float32); sd = _c(sd, np.float32)
    N = pts.shape[0]
    leaf = np.zeros((max(N, 1), 8), np.float32)
    ids = np.zeros(max(N, 1), np.int32)
    par = np.zeros(max(2 * N, 1), np.int32)
    mass = np.zeros(2, np.float64)
    nc = lib().o_foot_view(ctypes.c_int(N), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_double(f),
                           ctypes.c_double(smoothing), ctypes.c_double(precision), _p(pts), _p(nrm), _p(sd),
                           _p(leaf), _p(ids), _p(par), _p(mass))
    if nc < 0:
        raise RuntimeError("oracle foot_view failed: %d" % nc)
    return leaf[:nc], ids[:nc], par[:max(2 * nc - 1, 0)], float(mass[0]), np.float32(mass[1])


def foot_terms(leaf, a1, q):
    """Footprint terms min(1, w(q), E(|q - s|)) of leaf records (n, 8) at points q (n, 2)."""
    leaf = _c(leaf, np.float32); q = _c(q, np.float32)
    out = np.zeros(leaf.shape[0], np.float32)
    lib().o_foot_terms(ctypes.c_int(leaf.shape[0]), _p(leaf), ctypes.c_float(a1), _p(q), _p(out))
    return out


def foot_bounds(a1, d2):
    """Node bounds of the footprint term at squared box distances d2 (n,)."""
    d2 = _c(d2, np.float32)
    out = np.zeros(d2.shape[0], np.float32)
    lib().o_foot_bounds(ctypes.c_int(d2.shape[0]), ctypes.c_float(a1), _p(d2), _p(out))
    return out


def poisson(W, H, f, pts, nrm, sd, kcap, rate, smoothing=0.0, precision=0.01, seed=0, tile_scale=None,
            proposal="bands", footprint_scale=0.0):
    """Poisson-rate sampling for a batch: (ids (bn,W,H,kcap), 1/pi (bn,W,H,kcap))."""
    ids, ipi = [], []
    for b in range(pts.shape[0]):
        sdb = sd[b] if sd.shape[0] > 1 else sd[0]
        a, p_, _, _ = poisson_view(W, H, f, pts[b], nrm[b], sdb, kcap, rate, smoothing, precision, seed, b,
                                   None if tile_scale is None else tile_scale[b], proposal, footprint_scale)
        ids.append(a); ipi.append(p_)
    return np.stack(ids), np.stack(ipi)


def poisson_wide_h(W, H, f, pts, nrm, sd, smoothing=0.0, precision=0.01):
    """The wide proposal bandwidth of Poisson-rate sampling for one view."""
    pts = _c(pts, np.float32); nrm = _c(nrm, np.float32); sd = _c(sd, np.float32)
    return lib().o_poisson_wide_h(ctypes.c_int(pts.shape[0]), ctypes.c_int(W), ctypes.c_int(H), ctypes.c_double(f),
                                  ctypes.c_double(smoothing), ctypes.c_double(precision), _p(pts), _p(nrm), _p(sd))

