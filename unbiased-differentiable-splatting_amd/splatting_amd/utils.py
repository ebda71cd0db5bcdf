"""Camera and gradient helpers with the conventions of the reference's
splatting/extensions/utils.py (view matrix from a position and an (x, y, z, w)
quaternion, the 'z' / '-z' forward axis, the row-wise gradient clip)."""
import torch as th


def normalize(x, eps=1e-4):
    """x / (|x| + eps) along the last axis (utils.py:7-9)."""
    return x / x.norm(dim=-1, keepdim=True).add(eps)


def _skew(v):
    """(..., 3) -> (..., 3, 3) cross-product matrices [v]_x, [v]_x u = v x u."""
    x, y, z = v.unbind(-1)
    o = th.zeros_like(x)
    return th.stack([th.stack([o, -z, y], -1), th.stack([z, o, -x], -1), th.stack([-y, x, o], -1)], -2)


def world2Camera(camPositions, camQuaternion):
    """(bn, 1, 4, 4) view matrices [R | t; 0 0 0 1] (utils.py:11-49).

    R is the homogeneous quadratic form of the quaternion q = (v, w), v = (x, y, z):
    R = (w^2 - |v|^2) I + 2 v v^T + 2 w [v]_x, which is the rotation of q / |q|
    scaled by |q|^2 (the reference's Euler-Rodrigues entries, e.g.
    R_00 = x^2 - y^2 - z^2 + w^2); t is the camera position."""
    v = camQuaternion[..., :3]
    w = camQuaternion[..., 3:4]
    eye = th.eye(3, dtype=camQuaternion.dtype, device=camQuaternion.device)
    R = ((w * w - (v * v).sum(-1, keepdim=True))[..., None] * eye + 2.0 * v[..., :, None] * v[..., None, :]
         + 2.0 * w[..., None] * _skew(v))
    top = th.cat([R, camPositions[..., :, None]], dim=-1)                       # (bn, 3, 4)
    last = th.zeros_like(top[..., :1, :])
    last[..., 0, 3] = 1.0
    return th.cat([top, last], dim=-2).unsqueeze(1)


def transformForward(viewMatrix, forward):
    """forward == 'z': the camera looks along +z, so the view's y and z rows change
    sign (the reference's flip @ mirror, utils.py:51-60); '-z' keeps the matrix."""
    if forward != "z":
        return viewMatrix
    sign = th.tensor([1.0, -1.0, -1.0, 1.0], dtype=viewMatrix.dtype, device=viewMatrix.device)
    return viewMatrix * sign[:, None]


def conditionalNormalization(dx, radius=1.0, eps=1e-4):
    """Rows of norm above 1 - eps scaled to norm |dx| / (|dx| + eps), others kept
    (utils.py:147-150)."""
    return dx / dx.norm(dim=-1, keepdim=True).add(eps).clamp(min=1.0)
