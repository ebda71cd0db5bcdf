"""Seeded synthetic scenes (surfel sets) for the benchmark and test workloads.

All generators return float32 numpy arrays: positions (N,3), normals (N,3),
std-devs (N,2), albedo (N,3), plus the camera (world->camera 4x4, width,
height, focal length).  The camera sits at distance 2.5 on +z looking down
-z, so visible camera-space points have z <= 0 (the reference's cull,
samplingCuda.cu:47).
"""
import numpy as np


def camera_matrix(distance=2.5):
    M = np.eye(4, dtype=np.float32)
    M[2, 3] = -distance
    return M


def unit_sphere(n, seed, sd_lo=0.005, sd_hi=0.01):
    """Centres uniform on the unit sphere, radial normals, log-uniform std-devs."""
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    nrm = x.copy()
    sd = np.exp(rng.uniform(np.log(sd_lo), np.log(sd_hi), size=(n, 2)))
    albedo = rng.uniform(0.0, 1.0, size=(n, 3))
    return x.astype(np.float32), nrm.astype(np.float32), sd.astype(np.float32), albedo.astype(np.float32)


def _ellipsoid_surface(rng, n, axes):
    """Area-uniform points on an ellipsoid by rejection on the area element."""
    a, b, c = axes
    gmax = max(a * b, a * c, b * c)
    pts = []
    got = 0
    while got < n:
        m = int((n - got) * 2.2) + 64
        u = rng.normal(size=(m, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        g = np.sqrt((b * c * u[:, 0]) ** 2 + (a * c * u[:, 1]) ** 2 + (a * b * u[:, 2]) ** 2)
        keep = rng.uniform(0.0, gmax, size=m) < g
        u = u[keep]
        pts.append(u * np.array([a, b, c]))
        got += u.shape[0]
    p = np.concatenate(pts)[:n]
    nrm = p / np.array([a * a, b * b, c * c])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return p, nrm


def ellipsoids(n_ellipsoids=64, per=15625, seed=1, sd_lo=0.003, sd_hi=0.008):
    """Random ellipsoids (semi-axes U[0.05,0.3], centres U[-1,1]^3), area-uniform surfels."""
    rng = np.random.default_rng(seed)
    P, Nn = [], []
    for _ in range(n_ellipsoids):
        axes = rng.uniform(0.05, 0.3, size=3)
        centre = rng.uniform(-1.0, 1.0, size=3)
        p, nrm = _ellipsoid_surface(rng, per, axes)
        P.append(p + centre)
        Nn.append(nrm)
    pos = np.concatenate(P)
    nrm = np.concatenate(Nn)
    n = pos.shape[0]
    sd = rng.uniform(sd_lo, sd_hi, size=(n, 2))
    albedo = rng.uniform(0.0, 1.0, size=(n, 3))
    return pos.astype(np.float32), nrm.astype(np.float32), sd.astype(np.float32), albedo.astype(np.float32)


class Scene:
    def __init__(self, positions, normals, stddevs, albedo, width, height, focal, view=None):
        self.positions, self.normals, self.stddevs, self.albedo = positions, normals, stddevs, albedo
        self.width, self.height, self.focal = width, height, focal
        self.view = camera_matrix() if view is None else view

    @property
    def n(self):
        return self.positions.shape[0]

    def camera_space(self):
        """(camPoints, camNormals) float32 under the single view matrix."""
        M = self.view.astype(np.float64)
        p = self.positions.astype(np.float64) @ M[:3, :3].T + M[:3, 3]
        nrm = self.normals.astype(np.float64) @ M[:3, :3].T
        return p.astype(np.float32), nrm.astype(np.float32)


def config(idx, n_override=None):
    """Workloads: 0 = 4,096 splats on the unit sphere at 128x128 (focal 153.6);
    1 = 1M surfels on 64 ellipsoids at 512x512 (focal 614.4);
    2 = 256 splats of the sphere generator at 64x64 (focal 76.8)."""
    if idx == 0:
        p, n, s, a = unit_sphere(n_override or 4096, seed=0)
        return Scene(p, n, s, a, 128, 128, 1.2 * 128)
    if idx == 1:
        per = 15625 if n_override is None else max(1, n_override // 64)
        p, n, s, a = ellipsoids(64, per, seed=1)
        return Scene(p, n, s, a, 512, 512, 1.2 * 512)
    if idx == 2:
        p, n, s, a = unit_sphere(n_override or 256, seed=2)
        return Scene(p, n, s, a, 64, 64, 1.2 * 64)
    raise ValueError("unknown workload %r" % idx)


def training_scene(k, n_override=None):
    """The k-th random ellipsoid scene used to train the rate network: 64 ellipsoids with
    random axes and centres, 1M surfels, 512x512 (focal 614.4), ellipsoid seed 1000 + k."""
    per = 15625 if n_override is None else max(1, n_override // 64)
    p, n, s, a = ellipsoids(64, per, seed=1000 + k)
    return Scene(p, n, s, a, 512, 512, 1.2 * 512)


def perturbed_target_positions(scene, sigma=0.01, seed=100):
    rng = np.random.default_rng(seed)
    return (scene.positions + rng.normal(0.0, sigma, size=scene.positions.shape)).astype(np.float32)
