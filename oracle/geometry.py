"""Float64 numpy camera and pose helpers — TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates src/nanodrr/camera/intrinsics.py:53-68, extrinsics.py:7-57 and the
SE(3) exponential of src/nanodrr/geometry/se3.py:179-221 (Rodrigues and the
left Jacobian V, here with exact small-angle series).
"""

from __future__ import annotations

import numpy as np

ORIENTATION = {
    "AP": np.array([[-1, 0, 0, 0], [0, 0, -1, 0], [0, -1, 0, 0], [0, 0, 0, 1]], np.float64),
    "PA": np.array([[-1, 0, 0, 0], [0, 0, 1, 0], [0, -1, 0, 0], [0, 0, 0, 1]], np.float64),
    None: np.array([[-1, 0, 0, 0], [0, -1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float64),
}


def make_k_inv(sdd, delx, dely, x0, y0, height, width) -> np.ndarray:
    fx, fy = sdd / delx, sdd / dely
    cx, cy = x0 / delx + width / 2.0, y0 / dely + height / 2.0
    return np.array([[[1 / fx, 0, -cx / fx], [0, 1 / fy, -cy / fy], [0, 0, 1]]], np.float64)


def _rx(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def _ry(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def _rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def euler_zxy_deg(angles) -> np.ndarray:
    """Intrinsic ZXY: R = Rz(a) Rx(b) Ry(c), angles in degrees, (B, 3) -> (B, 3, 3)."""
    a = np.deg2rad(np.asarray(angles, np.float64))
    return np.stack([_rz(r[0]) @ _rx(r[1]) @ _ry(r[2]) for r in a])


def make_rt_inv(rotation, translation, orientation="AP", isocenter=None) -> np.ndarray:
    R = euler_zxy_deg(rotation)
    t = np.einsum("bij,bj->bi", R, np.asarray(translation, np.float64))
    if isocenter is not None:
        t = t + np.asarray(isocenter, np.float64)
    pose = np.tile(np.eye(4), (R.shape[0], 1, 1))
    pose[:, :3, :3] = R
    pose[:, :3, 3] = t
    return pose @ ORIENTATION[orientation]


def hat(w) -> np.ndarray:
    x, y, z = w
    return np.array([[0, -z, y], [z, 0, -x], [-y, x, 0]], np.float64)


def so3_exp(w) -> np.ndarray:
    w = np.asarray(w, np.float64)
    th = np.linalg.norm(w)
    K = hat(w)
    if th < 1e-4:
        A = 1 - th**2 / 6 + th**4 / 120
        B = 0.5 - th**2 / 24 + th**4 / 720
    else:
        A = np.sin(th) / th
        B = (1 - np.cos(th)) / th**2
    return np.eye(3) + A * K + B * K @ K


def se3_exp(w, u) -> np.ndarray:
    w = np.asarray(w, np.float64)
    th = np.linalg.norm(w)
    K = hat(w)
    if th < 1e-4:
        B = 0.5 - th**2 / 24 + th**4 / 720
        Cc = 1 / 6 - th**2 / 120 + th**4 / 5040
    else:
        B = (1 - np.cos(th)) / th**2
        Cc = (th - np.sin(th)) / th**3
    V = np.eye(3) + B * K + Cc * K @ K
    T = np.eye(4)
    T[:3, :3] = so3_exp(w)
    T[:3, 3] = V @ np.asarray(u, np.float64)
    return T
