"""Host-side logic that runs without a GPU: argument checking of the op
surface (GPU-only, no fallback), camera / gradient helpers
(splatting/extensions/utils.py), scene generators, the bench contract."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ------------------------------------------------------------------ op surface refuses CPU tensors
def test_ops_refuse_cpu_tensors():
    from cudaExtensions import sampling, splatting, filtering, interpolation, transform, shading
    p = torch.zeros(1, 4, 3)
    s = torch.ones(1, 4, 2)
    idx = torch.zeros(1, 2, 2, 3, dtype=torch.int32)
    w = torch.zeros(1, 2, 2, 3)
    M = torch.eye(4)[None]
    calls = [
        lambda: sampling.sample(2, 2, 1.0, p, p, None, s, 3, 0.0),
        lambda: sampling.covariance(2, 2, 1.0, p, p, s, 0.0),
        lambda: splatting.forward(2, 2, 1.0, idx, p, p, s, 0.0),
        lambda: splatting.backward(2, 2, 1.0, w, idx, p, p, s, 0.0),
        lambda: filtering.forward(idx, w, 2, 2, 1.0, p, p),
        lambda: filtering.backward(idx, w, w, 2, 2, 1.0, p, p),
        lambda: interpolation.forward(idx, w, p),
        lambda: interpolation.backward(torch.zeros(1, 2, 2, 3), idx, w, p),
        lambda: transform.affineForward(M, p),
        lambda: transform.linearForward(M, p),
        lambda: shading.forwardSh(torch.zeros(1, 3, 9), torch.zeros(1, 2, 2, 3), torch.zeros(1, 2, 2, 3)),
    ]
    for c in calls:
        with pytest.raises(RuntimeError, match="GPU only"):
            c()


def test_ops_reject_shapes_before_launch():
    """Shapes the kernels index by are checked on the host, before any device
    work: e.g. (1, N, 3) points with two views would be read past their end."""
    from cudaExtensions import sampling, splatting, filtering, interpolation, transform, shading, estimator
    p = torch.zeros(2, 4, 3)
    p1 = torch.zeros(1, 4, 3)
    s = torch.ones(1, 4, 2)
    idx = torch.zeros(2, 2, 2, 3, dtype=torch.int32)
    w = torch.zeros(2, 2, 2, 3)
    M = torch.eye(4)[None].repeat(2, 1, 1)
    bad = [
        lambda: splatting.forward(2, 2, 1.0, idx, p1, p1, s, 0.0),                 # broadcast points
        lambda: splatting.forward(3, 2, 1.0, idx, p, p, s, 0.0),                   # width != indices
        lambda: splatting.forward(2, 2, 1.0, idx, p, p, torch.ones(1, 5, 2), 0.0),   # N mismatch
        lambda: splatting.backward(2, 2, 1.0, w[..., :2], idx, p, p, s, 0.0),       # gradIn shape
        lambda: estimator.forward(2, 2, 1.0, idx, w, p1, p1, s, torch.zeros(1, 4, 3)),
        lambda: estimator.forward(2, 2, 1.0, idx, w, p, p, s, torch.zeros(3, 4, 3)),
        lambda: estimator.backward(2, 2, 1.0, torch.zeros(2, 2, 2, 4), idx, w, p, p, s, torch.zeros(1, 4, 3)),
        lambda: filtering.forward(idx, w[..., :2], 2, 2, 1.0, p, p),
        lambda: filtering.forward(idx, w, 2, 2, 1.0, torch.zeros(3, 4, 3), torch.zeros(3, 4, 3)),
        lambda: interpolation.forward(idx, w, torch.zeros(3, 4, 3)),
        lambda: interpolation.backward(torch.zeros(2, 2, 2, 5), idx, w, torch.zeros(1, 4, 3)),
        lambda: transform.affineForward(M, torch.zeros(3, 4, 3)),
        lambda: transform.affineForward(torch.eye(3)[None], p1),
        lambda: transform.linearBackward(torch.zeros(2, 5, 3), M, p),
        lambda: shading.forward(torch.zeros(1, 2, 3), torch.zeros(1, 2, 3), torch.zeros(2, 2, 2, 3),
                                torch.zeros(2, 2, 2, 3)),
        lambda: shading.forwardSh(torch.zeros(3, 9), torch.zeros(1, 2, 2, 3), torch.zeros(1, 2, 2, 3)),
        lambda: sampling.sample(2, 2, 1.0, p, p1, None, s, 3, 0.0),
        lambda: sampling.samplePoisson(2, 2, 1.0, p, p, None, torch.ones(2, 5, 2), 3, 0.0),
    ]
    for i, c in enumerate(bad):
        with pytest.raises(ValueError):
            c()


def test_dtype_checks():
    from cudaExtensions import _lib
    with pytest.raises(TypeError, match="float32"):
        _lib.f32(torch.zeros(2, dtype=torch.float64), "points")
    with pytest.raises(TypeError, match="int32"):
        _lib.i32(torch.zeros(2, dtype=torch.int64), "indices")
    with pytest.raises(TypeError):
        _lib.require_gpu(np.zeros(3))


def test_seed_counter():
    from cudaExtensions import sampling
    sampling.manual_seed(41)
    assert sampling._next_seed() == 41 and sampling._next_seed() == 42
    sampling.manual_seed(0)


def test_no_oracle_in_product():
    """The product package never imports or loads the CPU oracle."""
    pkg = os.path.join(ROOT, "unbiased-differentiable-splatting_amd")
    for dp, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".cpp", "Makefile")):
                src = open(os.path.join(dp, fn), errors="ignore").read()
                assert "oracle" not in src.replace("the oracle", ""), os.path.join(dp, fn)


# ------------------------------------------------------------------ utils
def test_world2camera_is_rotation_of_quaternion():
    from splatting_amd.utils import world2Camera
    rng = np.random.default_rng(0)
    q = rng.normal(size=(3, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    pos = rng.normal(size=(3, 3))
    M = world2Camera(torch.tensor(pos), torch.tensor(q)).numpy()[:, 0]
    for b in range(3):
        x, y, z, w = q[b]
        # rotation matrix of the unit quaternion w + xi + yj + zk
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        assert np.allclose(M[b, :3, :3], R, atol=1e-12)
        assert np.allclose(M[b, :3, 3], pos[b]) and np.allclose(M[b, 3], [0, 0, 0, 1])


def test_world2camera_scales_with_quaternion_norm():
    """A non-unit quaternion gives |q|^2 times the rotation of q / |q| (the entries
    are homogeneous quadratics, as the reference's x^2 - y^2 - z^2 + w^2 ...)."""
    from splatting_amd.utils import world2Camera
    rng = np.random.default_rng(1)
    q = rng.normal(size=(4, 4)) * 1.7
    pos = rng.normal(size=(4, 3))
    M = world2Camera(torch.tensor(pos), torch.tensor(q)).numpy()[:, 0]
    Mu = world2Camera(torch.tensor(pos), torch.tensor(q / np.linalg.norm(q, axis=1, keepdims=True))).numpy()[:, 0]
    s2 = (q * q).sum(1)
    for b in range(4):
        assert np.allclose(M[b, :3, :3], s2[b] * Mu[b, :3, :3], atol=1e-12)
        R = Mu[b, :3, :3]
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12) and abs(np.linalg.det(R) - 1.0) < 1e-12


def test_lambert_module_is_ambient_plus_clamped_cosines():
    """Per splat: diffuse * (ambient + sum_l I_l clamp(d_l . n, 0, 1) / pi), with the
    lamp intensities shared over the batch when their batch is 1 (shading.py:12-29)."""
    import math
    from splatting_amd.shading import Lambert
    rng = np.random.default_rng(2)
    bn, L, N = 2, 3, 50
    d = rng.normal(size=(bn, L, 3))
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    I = rng.uniform(size=(1, L, 3))
    n = rng.normal(size=(bn, N, 3))
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    a = rng.uniform(size=(bn, N, 3))
    amb = rng.uniform(size=(bn, 1, 3))
    out = Lambert()(*(torch.tensor(x) for x in (amb, d, I, n, a))).numpy()
    ref = np.zeros((bn, N, 3))
    for b in range(bn):
        for i in range(N):
            acc = amb[b, 0] * a[b, i]
            for l in range(L):
                c = min(max(float(d[b, l] @ n[b, i]), 0.0), 1.0)
                acc = acc + a[b, i] / math.pi * I[0, l] * c
            ref[b, i] = acc
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-14)


def test_transform_forward_flip():
    from splatting_amd.utils import transformForward
    M = torch.eye(4)[None, None]
    out = transformForward(M, "z")[0, 0]
    assert torch.equal(out.diag(), torch.tensor([1.0, -1.0, -1.0, 1.0]))
    assert torch.equal(transformForward(M, "-z"), M)


def test_conditional_normalization_and_normalize():
    from splatting_amd.utils import conditionalNormalization, normalize
    x = torch.tensor([[0.3, 0.4, 0.0], [3.0, 4.0, 0.0]])
    y = conditionalNormalization(x)
    assert torch.allclose(y[0], x[0]) and abs(float(y[1].norm()) - 1.0) < 1e-4
    n = normalize(torch.tensor([[0.0, 3.0, 4.0]]))
    assert abs(float(n.norm()) - 5.0 / 5.0001) < 1e-6


def test_ewa_gradient_mode_switch():
    from splatting_amd import extensions
    with pytest.raises(ValueError):
        extensions.set_ewa_gradient_mode("exact")
    extensions.set_ewa_gradient_mode("raw")
    assert extensions.GRAD_MODE["ewa"] == "raw"
    extensions.set_ewa_gradient_mode("reference")


def test_backface_culling():
    from splatting_amd.render import BackfaceCulling
    n = torch.tensor([[[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]]])
    a, b = BackfaceCulling()(n, n)
    assert torch.equal(b[..., 2], torch.tensor([[-1.0, -1.0]]))


# ------------------------------------------------------------------ scenes
def test_scene_configs_deterministic_and_well_formed():
    from splatting_amd import scenes
    for cfg, n, W in [(0, 4096, 128), (2, 256, 64)]:
        a = scenes.config(cfg)
        b = scenes.config(cfg)
        assert a.n == n and a.width == a.height == W
        assert np.array_equal(a.positions, b.positions) and np.array_equal(a.stddevs, b.stddevs)
        assert a.positions.dtype == np.float32 and a.stddevs.shape == (n, 2)
        assert np.allclose(np.linalg.norm(a.normals, axis=1), 1.0, atol=1e-5)
        p, nrm = a.camera_space()
        assert np.all(p[:, 2] < 0)              # in front of the camera (looking down -z)
    big = scenes.config(1, n_override=64 * 100)
    assert big.n == 6400 and big.width == 512 and abs(big.focal - 614.4) < 1e-9


def test_ellipsoid_normals_are_surface_normals():
    from splatting_amd import scenes
    sc = scenes.config(1, n_override=64 * 50)
    assert np.allclose(np.linalg.norm(sc.normals, axis=1), 1.0, atol=1e-5)


# ------------------------------------------------------------------ bench contract
def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_bench_algorithmic_bytes():
    sys.path.insert(0, ROOT)
    import bench
    b = bench.algorithmic_bytes(1_000_000, 512 * 512, 40)
    assert b["sample"] == 32 * 1_000_000 + 4 * 512 * 512 * 40
    assert all(v > 0 for v in b.values())


# ------------------------------------------------------------------ box room and cameras
def test_box_room_surfels_lie_on_faces():
    from splatting_amd import scenes
    p, n, s, a = scenes.box_room(20_000, n_boxes=12, seed=5)
    assert p.shape == (20_000, 3) and n.shape == (20_000, 3) and s.shape == (20_000, 2) and a.shape == (20_000, 3)
    assert np.all(np.abs(np.linalg.norm(n, axis=1) - 1) < 1e-6)
    assert np.all((np.abs(n) == 1).sum(1) == 1)                   # axis-aligned normals
    assert np.all((s >= 0.01) & (s <= 0.03))
    lo, hi = scenes.room_boxes(12, seed=5)
    ax = np.argmax(np.abs(n), axis=1)
    on_face = np.zeros(len(p), bool)
    room = np.array([8.0, 6.0, 3.0])
    for q in range(3):
        sel = ax == q
        # room walls: normal points inward
        on_face |= sel & (np.isclose(p[:, q], 0.0, atol=1e-5) & (n[:, q] > 0))
        on_face |= sel & (np.isclose(p[:, q], room[q], atol=1e-4) & (n[:, q] < 0))
        for b in range(len(lo)):
            inside = np.all((p >= lo[b] - 1e-4) & (p <= hi[b] + 1e-4), axis=1)
            on_face |= sel & inside & np.isclose(p[:, q], lo[b, q], atol=1e-4) & (n[:, q] < 0)
            on_face |= sel & inside & np.isclose(p[:, q], hi[b, q], atol=1e-4) & (n[:, q] > 0)
    assert on_face.all()


def test_room_cameras_look_along_yaw_pitch_and_avoid_boxes():
    from splatting_amd import scenes
    M = scenes.look_matrix([1.0, 2.0, 1.5], 0.3, -0.2)
    R = M[:3, :3].astype(np.float64)
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)
    d = np.array([np.cos(-0.2) * np.cos(0.3), np.cos(-0.2) * np.sin(0.3), np.sin(-0.2)])
    x = np.array([1.0, 2.0, 1.5]) + 2.0 * d                     # 2 m ahead of the camera
    c = R @ x + M[:3, 3]
    assert np.allclose(c, [0.0, 0.0, -2.0], atol=1e-5)
    boxes = scenes.room_boxes(200)
    V = scenes.room_cameras(16, boxes=boxes)
    assert V.shape == (16, 4, 4)
    centres = np.stack([-(m[:3, :3].T @ m[:3, 3]) for m in V.astype(np.float64)])
    assert np.all((centres > 0.3 - 1e-4) & (centres < np.array([8.0, 6.0, 3.0]) - 0.3 + 1e-4))
    for c in centres:
        assert not np.any(np.all((c > boxes[0] - 0.3) & (c < boxes[1] + 0.3), axis=1))


# ------------------------------------------------------------------ launch shapes and the tuner's cache
def test_tile_config_roundtrip_and_autotune_cache(tmp_path):
    from cudaExtensions import sampling
    from splatting_amd import autotune
    saved = {k: sampling.get_tile_config(k) for k in autotune.KERNELS}
    try:
        sampling.set_tile_config("poisson", 16, 4, 2, 1)
        assert sampling.get_tile_config("poisson") == (16, 4, 2, 1)
        with pytest.raises(RuntimeError):
            sampling.set_tile_config("fmm_eval", 8, 4, 4)
        assert autotune.workload_key(640, 480, 16, 2_000_000, 40) == "640x480_v16_n2000000_k40"
        assert autotune.workload_key(512, 512, 1, 767_129, 40) == "512x512_v1_n770000_k40"
        path = str(tmp_path / "tune.json")
        res = {"fmm_eval": {"config": [4, 16, 8, 1], "ms": 1.0, "default_ms": 1.1, "candidates": 32},
               "poisson": {"config": [32, 2, 1, 0], "ms": 2.0, "default_ms": 2.0, "candidates": 32}}
        autotune.save(path, "k1", res)
        assert autotune.load(path, "k1") == res and autotune.load(path, "k2") is None
        calls = []
        got = autotune.tune_cached(lambda: calls.append(1), path, "k1")
        assert got == res and not calls                      # cached: nothing is run
        assert sampling.get_tile_config("fmm_eval") == (4, 16, 8, 1)
        assert sampling.get_tile_config("poisson") == (32, 2, 1, 0)
    finally:
        for k, c in saved.items():
            sampling.set_tile_config(k, *c)


def test_autotune_degree_picks_the_fastest_order():
    import time as _t
    from cudaExtensions import sampling
    from splatting_amd import autotune
    cost = {4: 0.004, 5: 0.001, 6: 0.003, 7: 0.006, 8: 0.008}
    try:
        r = autotune.tune_degree(lambda: _t.sleep(cost[sampling.get_fixed_degree()]), repeats=2, sync=lambda: None)
        assert r["degree"] == 5 and sampling.get_fixed_degree() == 5
        assert set(r["table"]) == set(cost)
        autotune.apply({"degree": {"degree": 7}})
        assert sampling.get_fixed_degree() == 7
        with pytest.raises(RuntimeError):
            sampling.set_fixed_degree(9)
    finally:
        sampling.set_fixed_degree(6)
