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


def _box_faces(lo, hi, inward):
    """The six faces of an axis-aligned box as (axis, value, lo, hi, normal sign)."""
    faces = []
    for ax in range(3):
        for side, v in ((0, lo[ax]), (1, hi[ax])):
            out = 1.0 if side == 1 else -1.0
            faces.append((ax, v, lo, hi, -out if inward else out))
    return faces


def room_boxes(n_boxes=200, seed=3, size=(8.0, 6.0, 3.0)):
    """n_boxes random axis-aligned boxes with sides U[0.2, 1.0] m inside the room: (lo, hi) arrays."""
    rng = np.random.default_rng(seed)
    side = rng.uniform(0.2, 1.0, size=(n_boxes, 3))
    lo = rng.uniform(0.0, 1.0, size=(n_boxes, 3)) * (np.array(size) - side)
    return lo, lo + side


def box_room(n=2_000_000, n_boxes=200, seed=3, size=(8.0, 6.0, 3.0), sd_lo=0.01, sd_hi=0.03):
    """Surfels area-uniform on the inside of a size[0] x size[1] x size[2] m room
    (normals inward) and on the outside of the room_boxes (normals outward);
    std-devs U[sd_lo, sd_hi], random albedo."""
    rng = np.random.default_rng(seed + 1)
    faces = _box_faces(np.zeros(3), np.array(size, np.float64), True)
    for lo, hi in zip(*room_boxes(n_boxes, seed, size)):
        faces += _box_faces(lo, hi, False)
    area = np.array([np.prod([(f[3][a] - f[2][a]) for a in range(3) if a != f[0]]) for f in faces])
    counts = rng.multinomial(n, area / area.sum())
    P, Nn = [], []
    for (ax, v, lo, hi, sgn), c in zip(faces, counts):
        if c == 0:
            continue
        p = rng.uniform(lo, hi, size=(c, 3))
        p[:, ax] = v
        nrm = np.zeros((c, 3))
        nrm[:, ax] = sgn
        P.append(p)
        Nn.append(nrm)
    pos = np.concatenate(P)
    nrm = np.concatenate(Nn)
    perm = rng.permutation(n)
    pos, nrm = pos[perm], nrm[perm]
    sd = rng.uniform(sd_lo, sd_hi, size=(n, 2))
    albedo = rng.uniform(0.0, 1.0, size=(n, 3))
    return pos.astype(np.float32), nrm.astype(np.float32), sd.astype(np.float32), albedo.astype(np.float32)


def look_matrix(centre, yaw, pitch):
    """World -> camera matrix of a camera at `centre` looking along (yaw, pitch)
    (z up); the camera looks down its -z axis."""
    d = np.array([np.cos(pitch) * np.cos(yaw), np.cos(pitch) * np.sin(yaw), np.sin(pitch)])
    r = np.cross(d, [0.0, 0.0, 1.0])
    r /= np.linalg.norm(r)
    u = np.cross(r, d)
    R = np.stack([r, u, -d])
    M = np.eye(4)
    M[:3, :3] = R
    M[:3, 3] = -R @ np.asarray(centre, np.float64)
    return M.astype(np.float32)


def room_cameras(n_views=16, seed=4, size=(8.0, 6.0, 3.0), margin=0.3, pitch=0.4, boxes=None):
    """Cameras uniform inside the room, at least `margin` m from the walls and from
    every box in `boxes` ((lo, hi) arrays), with yaw U[0, 2 pi) and pitch
    U[-pitch, pitch]: (n_views, 4, 4) world -> camera matrices."""
    rng = np.random.default_rng(seed)
    lo, hi = np.full(3, margin), np.array(size) - margin
    mats = []
    while len(mats) < n_views:
        c = rng.uniform(lo, hi)
        yaw, pt = rng.uniform(0.0, 2.0 * np.pi), rng.uniform(-pitch, pitch)
        if boxes is not None and np.any(np.all((c > boxes[0] - margin) & (c < boxes[1] + margin), axis=1)):
            continue
        mats.append(look_matrix(c, yaw, pt))
    return np.stack(mats)


def room_scene(n=2_000_000, n_views=16, width=640, height=480, focal=400.0):
    """Batched-view workload: the 8 x 6 x 3 m box room with n surfels and 200 boxes,
    seen by n_views cameras at width x height (Scene.view holds (n_views, 4, 4))."""
    p, nrm, s, a = box_room(n)
    return Scene(p, nrm, s, a, width, height, focal, room_cameras(n_views, boxes=room_boxes()))


def perturbed_target_positions(scene, sigma=0.01, seed=100):
    rng = np.random.default_rng(seed)
    return (scene.positions + rng.normal(0.0, sigma, size=scene.positions.shape)).astype(np.float32)
