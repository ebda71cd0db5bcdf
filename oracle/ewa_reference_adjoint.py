"""Restatement of the reference's hand-written EWA weight gradient, FP64 numpy.

TEST INFRASTRUCTURE ONLY (the checker): only tests/ import this module.  The
product's adjoint (csrc/ud_ops.hip: k_splat_moments + k_splat_finalize) is
derived independently (moments of the whitened offset chained through forward
mode duals); this module restates the reference's own chain rule step by step
so the two can be compared:

  gradEWAWeight             ewaSplattingGrad.cuh:163-206
    gradScreenspace         transformationGrad.cuh:42-54
    gradJacobian            ewaSplattingGrad.cuh:58-99
      gradConstructInvJacobian  ewaSplattingGrad.cuh:9-56
      gradProjectOntoPlane  transformationGrad.cuh:70-88
      gradBackproject       transformationGrad.cuh:56-68
    gradCovariance          ewaSplattingGrad.cuh:101-132
    gradGaussian            ewaSplattingGrad.cuh:134-161
  tensor contractions in the index order of linalg.cuh:1174-1340
    (T322 * M33 and T322 * M23: A[i,k,l] = sum_j T[j,k,l] M[i,j];
     M22 * T2222: A[k,l] = sum_ij M[i,j] T[k,l,i,j];
     M22 * T322: v[i] = sum_kl M[k,l] T[i,k,l];
     T222 * M32: A[k,i,j] = sum_l M[k,l] T[l,i,j];
     V2 * M32: v[i] = sum_j u[j] M[i,j];  M22 * T222: v[k] = sum_ij M[i,j] T[k,i,j])
and the forward quantities of ewaSplatting.cuh:8-91.  Every function works on M
slots at once (leading axis).
"""
import numpy as np


def _forward(px, py, W, H, f, p, n, sd, smoothing):
    """ewaSplatting.cuh:8-91: screen point, x~0, x~1, invJ, V and the Gaussian parts."""
    s = np.stack([-p[:, 0] * f / p[:, 2] + W / 2.0, p[:, 1] * f / p[:, 2] + H / 2.0], 1)

    def backproject(y):
        return np.stack([(W - 2.0 * y[:, 0]) / (2.0 * f), -(H - 2.0 * y[:, 1]) / (2.0 * f), np.ones(len(y))], 1)

    x0 = backproject(s + np.array([1.0, 0.0]))
    x1 = backproject(s + np.array([0.0, 1.0]))
    na = (p * n).sum(1)
    xt0 = x0 * (na / (x0 * n).sum(1))[:, None] - p
    xt1 = x1 * (na / (x1 * n).sum(1))[:, None] - p
    nrm0 = np.linalg.norm(xt0, axis=1)
    u0 = xt0 / nrm0[:, None]
    u1 = np.cross(u0, n)
    invJ = np.zeros((len(p), 2, 2))
    invJ[:, 0, 0] = nrm0
    invJ[:, 0, 1] = (xt1 * u0).sum(1)
    invJ[:, 1, 1] = (xt1 * u1).sum(1)
    J = np.linalg.inv(invJ)
    D2 = np.zeros((len(p), 2, 2))
    D2[:, 0, 0], D2[:, 1, 1] = sd[:, 0] ** 2, sd[:, 1] ** 2
    V = J @ D2 @ np.transpose(J, (0, 2, 1)) + smoothing * np.eye(2)
    return s, x0, x1, xt0, xt1, invJ, V


def _grad_construct_inv_jacobian(a, b, c):
    """ewaSplattingGrad.cuh:9-56 with a = x~0, b = x~1, c = the camera normal:
    T[i, k, l] = d invJ(k, l) / d a_i (and of b_i, c_i)."""
    M = len(a)
    q = (a * a).sum(1)
    inv = 1.0 / np.sqrt(q)
    ipn = 1.0 / (np.sqrt(q) * q)
    a0, a1, a2 = a[:, 0], a[:, 1], a[:, 2]
    b0, b1, b2 = b[:, 0], b[:, 1], b[:, 2]
    c0, c1, c2 = c[:, 0], c[:, 1], c[:, 2]
    Ta = np.zeros((M, 3, 2, 2))
    Ta[:, 0, 0, 0] = a0 * inv
    Ta[:, 0, 0, 1] = (-a0 * a1 * b1 - a0 * a2 * b2 + a1 ** 2 * b0 + a2 ** 2 * b0) * ipn
    Ta[:, 0, 1, 1] = (a0 * b0 * (c1 * a2 - c2 * a1) - b1 * (c2 * q + a0 * (c0 * a2 - c2 * a0))
                      + b2 * (c1 * q + a0 * (c0 * a1 - c1 * a0))) * ipn
    Ta[:, 1, 0, 0] = a1 * inv
    Ta[:, 1, 0, 1] = (a0 ** 2 * b1 - a0 * a1 * b0 - a1 * a2 * b2 + a2 ** 2 * b1) * ipn
    Ta[:, 1, 1, 1] = (-a1 * b1 * (c0 * a2 - c2 * a0) + b0 * (c2 * q + a1 * (c1 * a2 - c2 * a1))
                      - b2 * (c0 * q + a1 * (-c0 * a1 + c1 * a0))) * ipn
    Ta[:, 2, 0, 0] = a2 * inv
    Ta[:, 2, 0, 1] = (a0 ** 2 * b2 - a0 * a2 * b0 + a1 ** 2 * b2 - a1 * a2 * b1) * ipn
    Ta[:, 2, 1, 1] = (a2 * b2 * (c0 * a1 - c1 * a0) - b0 * (c1 * q + a2 * (-c1 * a2 + c2 * a1))
                      + b1 * (c0 * q + a2 * (-c0 * a2 + c2 * a0))) * ipn
    Tb = np.zeros((M, 3, 2, 2))
    Tb[:, 0, 0, 1] = a0 * inv
    Tb[:, 0, 1, 1] = a1 * c2 * inv - a2 * c1 * inv
    Tb[:, 1, 0, 1] = a1 * inv
    Tb[:, 1, 1, 1] = -a0 * c2 * inv + a2 * c0 * inv
    Tb[:, 2, 0, 1] = a2 * inv
    Tb[:, 2, 1, 1] = a0 * c1 * inv - a1 * c0 * inv
    Tc = np.zeros((M, 3, 2, 2))
    Tc[:, 0, 1, 1] = -a1 * b2 * inv + a2 * b1 * inv
    Tc[:, 1, 1, 1] = a0 * b2 * inv - a2 * b0 * inv
    Tc[:, 2, 1, 1] = -a0 * b1 * inv + a1 * b0 * inv
    return Ta, Tb, Tc


def _grad_project_onto_plane(x, p, n):
    """transformationGrad.cuh:70-88 (Matrix33 results)."""
    na = (p * n).sum(1)
    inv_dot = 1.0 / (x * n).sum(1)
    xts = na * inv_dot
    outer = lambda u, v: u[:, :, None] * v[:, None, :]
    gX = xts[:, None, None] * np.eye(3) - outer(x, n) * (na * inv_dot ** 2)[:, None, None]
    gP = outer(x, n) * inv_dot[:, None, None]
    gN = outer(x, p) * inv_dot[:, None, None] - outer(x, x) * (na * inv_dot ** 2)[:, None, None]
    return gX, gP, gN


def _grad_covariance(invJ, sd):
    """ewaSplattingGrad.cuh:101-132: TV[k, l, i, j] and TS[k, i, j] as the reference
    writes them (a, b, c, d = invJ(0,0), invJ(0,1), invJ(1,0), invJ(1,1))."""
    a, b, c, d = invJ[:, 0, 0], invJ[:, 0, 1], invJ[:, 1, 0], invJ[:, 1, 1]
    s0, s1 = sd[:, 0], sd[:, 1]
    det = a * d - b * c
    ic = 1.0 / det ** 3
    M = len(a)
    TV = np.zeros((M, 2, 2, 2, 2))
    TV[:, 0, 0, 0, 0] = -2 * d * (b ** 2 * s1 ** 2 + d ** 2 * s0 ** 2) * ic
    TV[:, 0, 0, 0, 1] = (-b * s1 ** 2 * det + 2 * d * (a * b * s1 ** 2 + c * d * s0 ** 2)) * ic
    TV[:, 0, 0, 1, 0] = TV[:, 0, 0, 0, 1]
    TV[:, 0, 0, 1, 1] = 2 * (a * s1 ** 2 * det - d * (a ** 2 * s1 ** 2 + c ** 2 * s0 ** 2)) * ic
    TV[:, 0, 1, 0, 0] = 2 * a * b * d * s1 ** 2 * ic + 2 * c * d ** 2 * s0 ** 2 * ic
    TV[:, 0, 1, 0, 1] = -(a * s1 ** 2 * det + 2 * c * (a * b * s1 ** 2 + c * d * s0 ** 2)) * ic
    TV[:, 0, 1, 1, 0] = TV[:, 0, 1, 0, 1]
    TV[:, 0, 1, 1, 1] = 2 * c * (a ** 2 * s1 ** 2 + c ** 2 * s0 ** 2) * ic
    TV[:, 1, 0, 0, 0] = 2 * b * (b ** 2 * s1 ** 2 + d ** 2 * s0 ** 2) * ic
    TV[:, 1, 0, 0, 1] = -(2 * b * (a * b * s1 ** 2 + c * d * s0 ** 2) + d * s0 ** 2 * det) * ic
    TV[:, 1, 0, 1, 0] = TV[:, 1, 0, 0, 1]
    TV[:, 1, 0, 1, 1] = 2 * a ** 2 * b * s1 ** 2 * ic + 2 * a * c * d * s0 ** 2 * ic
    TV[:, 1, 1, 0, 0] = 2 * (-a * (b ** 2 * s1 ** 2 + d ** 2 * s0 ** 2) + d * s0 ** 2 * det) * ic
    TV[:, 1, 1, 0, 1] = (2 * a * (a * b * s1 ** 2 + c * d * s0 ** 2) - c * s0 ** 2 * det) * ic
    TV[:, 1, 1, 1, 0] = TV[:, 1, 1, 0, 1]
    TV[:, 1, 1, 1, 1] = -2 * a * (a ** 2 * s1 ** 2 + c ** 2 * s0 ** 2) * ic
    isq = 1.0 / det ** 2
    TS = np.zeros((M, 2, 2, 2))
    TS[:, 0, 0, 0] = 2 * d ** 2 * s0 * isq
    TS[:, 0, 0, 1] = -2 * c * d * s0 * isq
    TS[:, 0, 1, 0] = TS[:, 0, 0, 1]
    TS[:, 0, 1, 1] = 2 * c ** 2 * s0 * isq
    TS[:, 1, 0, 0] = 2 * b ** 2 * s1 * isq
    TS[:, 1, 0, 1] = -2 * a * b * s1 * isq
    TS[:, 1, 1, 0] = TS[:, 1, 0, 1]
    TS[:, 1, 1, 1] = 2 * a ** 2 * s1 * isq
    return TV, TS


def ewa_gradient(px, py, W, H, f, p, n, sd, smoothing=0.0):
    """The reference's (unclamped weight, d w / d point, d w / d normal, d w / d std-dev)
    of each slot, gradEWAWeight (ewaSplattingGrad.cuh:163-206)."""
    p, n, sd = (np.asarray(a, np.float64) for a in (p, n, sd))
    px, py = np.asarray(px, np.float64), np.asarray(py, np.float64)
    s, x0, x1, xt0, xt1, invJ, V = _forward(px, py, W, H, f, p, n, sd, smoothing)
    # gradScreenspace: D[i, j] = d s_j / d p_i
    D = np.zeros((len(p), 3, 2))
    D[:, 0, 0] = -f / p[:, 2]
    D[:, 1, 1] = f / p[:, 2]
    D[:, 2, 0] = p[:, 0] * f / p[:, 2] ** 2
    D[:, 2, 1] = -p[:, 1] * f / p[:, 2] ** 2
    # gradJacobian
    T0, T1, TN = _grad_construct_inv_jacobian(xt0, xt1, n)
    gX0, gP0, gN0 = _grad_project_onto_plane(x0, p, n)
    gX1, gP1, gN1 = _grad_project_onto_plane(x1, p, n)
    G = np.zeros((2, 3))
    G[0, 0] = -1.0 / f
    G[1, 1] = 1.0 / f
    tr = lambda A: np.transpose(A, (0, 2, 1))
    t322_m = lambda T, Mx: np.einsum("mjkl,mij->mikl", T, Mx)          # T322 * M33 / M23
    t322_g = lambda T: np.einsum("mjkl,ij->mikl", T, G)
    dInvJ_dScreen = t322_g(t322_m(T0, tr(gX0))) + t322_g(t322_m(T1, tr(gX1)))
    eye = np.eye(3)
    dInvJ_dP = t322_m(T0, tr(gP0 - eye)) + t322_m(T1, tr(gP1 - eye))
    dInvJ_dN = TN + t322_m(T0, tr(gN0)) + t322_m(T1, tr(gN1))
    # gradCovariance
    TV, TS = _grad_covariance(invJ, sd)
    # gradGaussian
    x = np.stack([px, py], 1) - s
    iV = np.linalg.inv(V)
    e = np.einsum("mi,mij,mj->m", x, iV, x)
    detV = np.linalg.det(V)
    norm = 2.0 * np.pi * np.sqrt(detV)
    gauss = np.exp(-0.5 * e) / norm
    idet = np.abs(np.linalg.det(invJ))
    w = gauss / idet
    dw_ds = np.einsum("mij,mj->mi", iV, x) * w[:, None]
    Jk = np.linalg.inv(invJ)
    dw_dInvJ = tr(Jk) * (-(gauss / idet))[:, None, None]
    Vkf = 1.0 / (norm * idet)
    Vkg = np.exp(-0.5 * e)
    gVf = tr(iV) * (-1.0 / (4.0 * np.pi * np.sqrt(detV) * idet))[:, None, None]
    a_ = np.einsum("mji,mj->mi", iV, x)                                   # iV^T x
    b_ = np.einsum("mij,mj->mi", iV, x)                                   # iV x
    gVg = a_[:, :, None] * b_[:, None, :] * (np.exp(-0.5 * e) * 0.5)[:, None, None]
    dw_dV = gVf * Vkg[:, None, None] + gVg * Vkf[:, None, None]
    # chain (linalg.cuh contractions)
    gInvJ = np.einsum("mij,mklij->mkl", dw_dV, TV) + dw_dInvJ
    g_n = np.einsum("mkl,mikl->mi", gInvJ, dInvJ_dN)
    screen_to_p = np.einsum("mkl,mlij->mkij", D, dInvJ_dScreen)            # T222 * M32
    g_p = np.einsum("mkl,mikl->mi", gInvJ, dInvJ_dP + screen_to_p) + np.einsum("mj,mij->mi", dw_ds, D)
    g_s = np.einsum("mij,mkij->mk", dw_dV, TS)
    return w, g_p, g_n, g_s
