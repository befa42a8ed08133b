"""Float64 numpy restatement of the registration pose step and ZNCC — TEST
INFRASTRUCTURE ONLY (see oracle/__init__.py).

pose_compose      M = (prefix @ T(w, v) @ suffix)[:3], T = [exp([w]x), exp([w]x) v; 0 1]:
                  the pose of src/nanodrr/registration/registration.py:64-68
                  (convert(rot, xyz, "so3_log"), src/nanodrr/geometry/se3.py:62-69)
                  folded with world_to_voxel as in src/nanodrr/drr/backends.py:84-86.
pose_compose_bwd  its adjoint, with dR/dw_k from the closed form of Gallego &
                  Yezzi, "A compact formula for the derivative of a 3-D rotation
                  in exponential coordinates" (J. Math. Imaging Vis. 2015):
                  dR/dw_k = (w_k [w]x + [w x (I - R) e_k]x) R / |w|^2,
                  and the first-order series [e_k]x + ([e_k]x K + K [e_k]x)/2 below
                  |w| = 1e-7.
zncc              zero-normalised cross-correlation of src/nanodrr/metrics/ncc.py:8-37
                  (population variance + eps inside the square root) and its
                  gradient with respect to the first image.
"""

from __future__ import annotations

import numpy as np

from .geometry import hat, so3_exp


def make_T(w, v) -> np.ndarray:
    R = so3_exp(w)
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = R @ np.asarray(v, np.float64)
    return T


def pose_compose(prefix, suffix, rot, xyz) -> np.ndarray:
    """(B, 3, 4) float64."""
    prefix, suffix = np.asarray(prefix, np.float64), np.asarray(suffix, np.float64)
    rot, xyz = np.asarray(rot, np.float64), np.asarray(xyz, np.float64)
    return np.stack([(prefix[b] @ make_T(rot[b], xyz[b]) @ suffix[b])[:3] for b in range(len(prefix))])


def drot(w) -> np.ndarray:
    """(3, 3, 3): dR/dw_k for k = 0, 1, 2."""
    w = np.asarray(w, np.float64)
    th2 = float(w @ w)
    out = np.empty((3, 3, 3))
    if th2 < 1e-14:
        K = hat(w)
        for k in range(3):
            E = hat(np.eye(3)[k])
            out[k] = E + 0.5 * (E @ K + K @ E)
        return out
    R = so3_exp(w)
    I_R = np.eye(3) - R
    for k in range(3):
        out[k] = (w[k] * hat(w) + hat(np.cross(w, I_R[:, k]))) @ R / th2
    return out


def pose_compose_bwd(prefix, suffix, rot, xyz, gM):
    """(grot (B, 3), gxyz (B, 3)) = dL/d(rot, xyz) given gM = dL/dM (B, 3, 4)."""
    prefix, suffix = np.asarray(prefix, np.float64), np.asarray(suffix, np.float64)
    rot, xyz, gM = (np.asarray(a, np.float64) for a in (rot, xyz, gM))
    grot = np.zeros((len(prefix), 3))
    gxyz = np.zeros((len(prefix), 3))
    for b in range(len(prefix)):
        # dL/dT for the 4x4 T, M = P[:3] T S
        gT = prefix[b][:3].T @ gM[b] @ suffix[b].T
        R = so3_exp(rot[b])
        dR = drot(rot[b])
        # T[:3,:3] = R, T[:3,3] = R v
        gxyz[b] = R.T @ gT[:3, 3]
        for k in range(3):
            grot[b, k] = np.sum(gT[:3, :3] * dR[k]) + gT[:3, 3] @ (dR[k] @ xyz[b])
    return grot, gxyz


def zncc(x, y, eps: float = 1e-7):
    """(score (B,), grad_x (B, N)) over the flattened non-batch dims."""
    x = np.asarray(x, np.float64).reshape(len(x), -1)
    y = np.asarray(y, np.float64).reshape(len(y), -1)
    n = x.shape[1]
    dx = x - x.mean(axis=1, keepdims=True)
    dy = y - y.mean(axis=1, keepdims=True)
    sx = np.sqrt((dx * dx).mean(axis=1, keepdims=True) + eps)
    sy = np.sqrt((dy * dy).mean(axis=1, keepdims=True) + eps)
    xh, yh = dx / sx, dy / sy
    score = (xh * yh).mean(axis=1)
    # score = C / (sx sy) with C = mean(dx dy): dC/dx_j = dy_j / n (dy sums to zero) and
    # dsx/dx_j = dx_j / (n sx), so dscore/dx_j = (yh_j - score xh_j) / (n sx).
    grad = (yh - score[:, None] * xh) / (n * sx)
    return score, grad


def patch_zncc(x, y, patch: int, eps: float = 1e-7):
    """(score (B,), grad_x (B, H, W)): the mean over every patch x patch window (stride 1) of
    (B, H, W) images of the windows' zncc (the patch form of ncc.py:21-24,83-88: each unfolded
    window standardised on its own), and its gradient, the windows' gradients scattered back
    and averaged."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    B, H, W = x.shape
    ho, wo = H - patch + 1, W - patch + 1
    score = np.zeros(B)
    grad = np.zeros((B, H, W))
    for b in range(B):
        wx = np.stack([x[b, i:i + patch, j:j + patch].ravel() for i in range(ho) for j in range(wo)])
        wy = np.stack([y[b, i:i + patch, j:j + patch].ravel() for i in range(ho) for j in range(wo)])
        s, g = zncc(wx, wy, eps)
        score[b] = s.mean()
        for k, (i, j) in enumerate((i, j) for i in range(ho) for j in range(wo)):
            grad[b, i:i + patch, j:j + patch] += g[k].reshape(patch, patch) / (ho * wo)
    return score, grad
