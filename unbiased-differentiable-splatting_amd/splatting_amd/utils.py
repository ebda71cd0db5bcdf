"""Camera and gradient helpers (reference: splatting/extensions/utils.py)."""
import torch as th


def normalize(x, eps=1e-4):
    norm = x.norm(dim=-1, keepdim=True)
    return x / (norm + eps)


def world2Camera(camPositions, camQuaternion):
    """(bn, 1, 4, 4) view matrices from positions and (x, y, z, w) quaternions
    (utils.py:11-49, Euler-Rodrigues rotation)."""
    bn = camPositions.shape[0]
    M = th.zeros((bn, 1, 4, 4), dtype=camPositions.dtype, device=camPositions.device)
    x, y, z, w = camQuaternion.split(1, dim=-1)
    x2, y2, z2, w2 = x * x, y * y, z * z, w * w
    xy, zw, xz, yw, yz, xw = x * y, z * w, x * z, y * w, y * z, x * w
    M[..., 0, 0] = x2 - y2 - z2 + w2
    M[..., 1, 0] = 2 * (xy + zw)
    M[..., 2, 0] = 2 * (xz - yw)
    M[..., 0, 1] = 2 * (xy - zw)
    M[..., 1, 1] = -x2 + y2 - z2 + w2
    M[..., 2, 1] = 2 * (yz + xw)
    M[..., 0, 2] = 2 * (xz + yw)
    M[..., 1, 2] = 2 * (yz - xw)
    M[..., 2, 2] = -x2 - y2 + z2 + w2
    M[..., :3, 3] = camPositions.unsqueeze(1)
    M[..., 3, 3] = 1.0
    return M


def transformForward(viewMatrix, forward):
    """'z': flip y and mirror z (utils.py:51-60); '-z': identity."""
    if forward == "z":
        mirror = th.eye(4, dtype=viewMatrix.dtype, device=viewMatrix.device).view(1, 1, 4, 4).clone()
        mirror[..., 2, 2] = -1.0
        flip = th.eye(4, dtype=viewMatrix.dtype, device=viewMatrix.device).view(1, 1, 4, 4).clone()
        flip[..., 1, 1] = -1.0
        viewMatrix = flip @ mirror @ viewMatrix
    return viewMatrix


def conditionalNormalization(dx, radius=1.0, eps=1e-4):
    """Row-wise clip to norm <= 1 (utils.py:147-150)."""
    norm = dx.norm(dim=-1, keepdim=True)
    scale = th.clamp(norm + eps, 1.0)
    return dx.div(scale)
