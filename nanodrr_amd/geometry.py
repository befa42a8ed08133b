"""Rigid-body geometry: homogeneous transforms, rotation parameterisations and
the SE(3) exponential/logarithm (torch, differentiable, any dtype).

Follows nanodrr's geometry package (src/nanodrr/geometry/transform.py:5-28,
src/nanodrr/geometry/se3.py:8-258) with the rotation conversions written here
instead of taken from roma. Two deliberate differences:
  * `se3_exp` / `_se3_V_matrix` use Taylor series below a small-angle threshold,
    so V is exact at theta = 0 (the reference clamps theta inside sqrt but
    divides by the unclamped theta^2, se3.py:211-219, which is NaN at 0);
  * `so3_exp` (Rodrigues) likewise switches to series near 0.
"""

from __future__ import annotations

import math
from enum import Enum

import torch

# ---------------------------------------------------------------------------
# Homogeneous points (transform.py:5-28)
# ---------------------------------------------------------------------------


def transform_point(xform: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
    """Apply (B, ..., N+1, N+1) homogeneous transforms to (B, ..., N) points (dehomogenised)."""
    vh = torch.cat([v, torch.ones_like(v[..., :1])], dim=-1)
    out = torch.einsum("b...ij,b...j->b...i", xform, vh)
    return out[..., :-1] / out[..., -1:]


# ---------------------------------------------------------------------------
# Rotations
# ---------------------------------------------------------------------------


def matmul3(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """a @ b for batches of small matrices (3 x 3, 4 x 4, 3 x 1) as broadcast products and one
    sum: elementwise kernels on the GPU instead of a library GEMM launch per product."""
    return (a[..., :, :, None] * b[..., None, :, :]).sum(-2)


def hat(v: torch.Tensor) -> torch.Tensor:
    """Skew-symmetric matrix [v]x of (..., 3) vectors (se3.py:224-242)."""
    x, y, z = v.unbind(-1)
    o = torch.zeros_like(x)
    return torch.stack([o, -z, y, z, o, -x, -y, x, o], dim=-1).reshape(*v.shape[:-1], 3, 3)


def _axis_rot(axis: str, a: torch.Tensor) -> torch.Tensor:
    c, s = torch.cos(a), torch.sin(a)
    o, i = torch.zeros_like(a), torch.ones_like(a)
    if axis == "X":
        m = [i, o, o, o, c, -s, o, s, c]
    elif axis == "Y":
        m = [c, o, s, o, i, o, -s, o, c]
    else:
        m = [c, -s, o, s, c, o, o, o, i]
    return torch.stack(m, dim=-1).reshape(*a.shape, 3, 3)


def euler_to_rotmat(convention: str, angles: torch.Tensor, degrees: bool = False) -> torch.Tensor:
    """Rotation matrices from (..., 3) Euler angles.

    Upper-case conventions ("ZXY") are intrinsic: R = R_Z(a0) R_X(a1) R_Y(a2);
    lower-case ("zxy") are extrinsic: R = R_y(a2) R_x(a1) R_z(a0). This is the
    convention of scipy's Rotation.from_euler and of roma.euler_to_rotmat.
    """
    if len(convention) != 3 or not (convention.isupper() or convention.islower()):
        raise ValueError(f"convention must be 3 axis letters, all upper or all lower case: {convention!r}")
    if set(convention.upper()) - set("XYZ"):
        raise ValueError(f"convention letters must come from X, Y, Z: {convention!r}")
    if degrees:
        angles = angles * (math.pi / 180.0)
    mats = [_axis_rot(ax, angles[..., k]) for k, ax in enumerate(convention.upper())]
    if convention.isupper():
        return matmul3(matmul3(mats[0], mats[1]), mats[2])
    return matmul3(matmul3(mats[2], mats[1]), mats[0])


def rotmat_to_euler_zxy(R: torch.Tensor, degrees: bool = False) -> torch.Tensor:
    """Inverse of euler_to_rotmat("ZXY", .) with the middle angle in [-90, 90] degrees."""
    b = torch.asin(R[..., 2, 1].clamp(-1.0, 1.0))
    a = torch.atan2(-R[..., 0, 1], R[..., 1, 1])
    c = torch.atan2(-R[..., 2, 0], R[..., 2, 2])
    out = torch.stack([a, b, c], dim=-1)
    return out * (180.0 / math.pi) if degrees else out


def _series_coeffs(theta2: torch.Tensor):
    """A = sin t / t, B = (1 - cos t) / t^2, Cc = (t - sin t) / t^3, exact or by Taylor series near 0."""
    small = theta2 < 1e-6
    t2 = torch.where(small, torch.ones_like(theta2), theta2)
    t = t2.sqrt()
    A = torch.where(small, 1.0 - theta2 / 6.0 + theta2 * theta2 / 120.0, torch.sin(t) / t)
    B = torch.where(small, 0.5 - theta2 / 24.0 + theta2 * theta2 / 720.0, (1.0 - torch.cos(t)) / t2)
    Cc = torch.where(small, 1.0 / 6.0 - theta2 / 120.0 + theta2 * theta2 / 5040.0, (t - torch.sin(t)) / (t2 * t))
    return A, B, Cc


def so3_exp(rotvec: torch.Tensor) -> torch.Tensor:
    """Rodrigues: R = I + A [w]x + B [w]x^2 for (..., 3) rotation vectors."""
    K = hat(rotvec)
    A, B, _ = _series_coeffs((rotvec * rotvec).sum(-1))
    eye = torch.eye(3, dtype=rotvec.dtype, device=rotvec.device)
    return eye + A[..., None, None] * K + B[..., None, None] * matmul3(K, K)


def so3_log(R: torch.Tensor) -> torch.Tensor:
    """Rotation vector of (..., 3, 3) rotations (angle in [0, pi])."""
    tr = R[..., 0, 0] + R[..., 1, 1] + R[..., 2, 2]
    cos = (tr - 1.0) / 2.0
    w = torch.stack([R[..., 2, 1] - R[..., 1, 2], R[..., 0, 2] - R[..., 2, 0], R[..., 1, 0] - R[..., 0, 1]], -1)
    theta = torch.atan2(0.5 * w.norm(dim=-1), cos)  # |w| = 2 sin(theta): accurate at 0 and pi
    sin = torch.sin(theta)
    near0 = theta < 1e-4
    near_pi = theta > math.pi - 1e-3
    scale = torch.where(near0, 0.5 + theta * theta / 12.0, theta / (2.0 * torch.where(near0, torch.ones_like(sin), sin)))
    rv = w * scale[..., None]
    if near_pi.any():
        # theta ~ pi: sym(R) - cos(theta) I = (1 - cos(theta)) a a^T gives the axis up to sign
        eye = torch.eye(3, dtype=R.dtype, device=R.device)
        S = ((R + R.transpose(-1, -2)) / 2.0 - cos[..., None, None] * eye) / (1.0 - cos)[..., None, None]
        diag = torch.diagonal(S, dim1=-2, dim2=-1).clamp(min=0.0)
        k = diag.argmax(-1, keepdim=True)
        col = torch.gather(S, -1, k[..., None, :].expand(*S.shape[:-1], 1)).squeeze(-1)
        axis = col / col.norm(dim=-1, keepdim=True).clamp(min=1e-30)
        sign = torch.sign((axis * w).sum(-1, keepdim=True))
        sign = torch.where(sign == 0, torch.ones_like(sign), sign)
        rv = torch.where(near_pi[..., None], axis * sign * theta[..., None], rv)
    return rv


def unitquat_to_rotmat(q: torch.Tensor) -> torch.Tensor:
    """(..., 4) unit quaternions in XYZW order to rotation matrices."""
    x, y, z, w = q.unbind(-1)
    m = [
        1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
        2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
        2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y),
    ]
    return torch.stack(m, dim=-1).reshape(*q.shape[:-1], 3, 3)


def special_procrustes(M: torch.Tensor) -> torch.Tensor:
    """Closest rotation (det +1) to each (..., 3, 3) matrix in Frobenius norm."""
    U, _, Vh = torch.linalg.svd(M)
    d = torch.det(U @ Vh)
    D = torch.diag_embed(torch.stack([torch.ones_like(d), torch.ones_like(d), d], -1))
    return U @ D @ Vh


def _vec10_to_symmetric4x4(vec: torch.Tensor) -> torch.Tensor:
    A = torch.zeros(*vec.shape[:-1], 4, 4, dtype=vec.dtype, device=vec.device)
    i, j = torch.triu_indices(4, 4)
    A[..., i, j] = vec
    A[..., j, i] = vec
    return A


def quaternion_adjugate_to_quaternion(rotation: torch.Tensor) -> torch.Tensor:
    """10-D quaternion-adjugate (upper triangle of a symmetric 4x4) to a WXYZ unit quaternion:
    the column of largest norm, normalised (se3.py:135-155; arXiv:2205.09116)."""
    A = _vec10_to_symmetric4x4(rotation)
    norms = A.norm(dim=-2)
    k = norms.argmax(-1)
    col = torch.gather(A, -1, k[..., None, None].expand(*A.shape[:-2], 4, 1)).squeeze(-1)
    return col / norms.amax(-1, keepdim=True)


def rotmat_geodesic_distance(R1: torch.Tensor, R2: torch.Tensor, clamping: float = 1.0) -> torch.Tensor:
    """Angle (radians) of R1^T R2."""
    tr = (R1 * R2).sum(dim=(-1, -2))
    return torch.acos(((tr - 1.0) / 2.0).clamp(-clamping, clamping))


# ---------------------------------------------------------------------------
# SE(3)
# ---------------------------------------------------------------------------


class Parameterization(str, Enum):
    EULER = "euler"
    QUATERNION = "quaternion"
    QUATERNION_ADJUGATE = "quaternion_adjugate"
    ROTATION_9D = "rotation_9d"
    SO3_LOG = "so3_log"
    SE3_LOG = "se3_log"

    @property
    def dim(self) -> int:
        return {"euler": 3, "quaternion": 4, "quaternion_adjugate": 10, "rotation_9d": 9,
                "so3_log": 3, "se3_log": 3}[self.value]


def make_se3(R: torch.Tensor, t: torch.Tensor) -> torch.Tensor:
    """(..., 4, 4) homogeneous matrices from rotations and translations (se3.py:118-132)."""
    top = torch.cat([R, t[..., :, None]], dim=-1)
    bottom = torch.zeros(*R.shape[:-2], 1, 4, dtype=R.dtype, device=R.device)
    bottom[..., 0, 3] = 1.0
    return torch.cat([top, bottom], dim=-2)


def se3_V(rotvec: torch.Tensor) -> torch.Tensor:
    """V = I + B [w]x + Cc [w]x^2, the left Jacobian of SO(3) (series near 0)."""
    K = hat(rotvec)
    _, B, Cc = _series_coeffs((rotvec * rotvec).sum(-1))
    eye = torch.eye(3, dtype=rotvec.dtype, device=rotvec.device)
    return eye + B[..., None, None] * K + Cc[..., None, None] * matmul3(K, K)


def se3_exp(log_rotation: torch.Tensor, log_translation: torch.Tensor) -> torch.Tensor:
    """SE(3) exponential of (rotation vector, translation log) pairs (se3.py:179-194)."""
    R = so3_exp(log_rotation)
    t = matmul3(se3_V(log_rotation), log_translation[..., :, None])[..., 0]
    return make_se3(R, t)


def se3_log_map(matrix: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """SE(3) logarithm: (rotation vector, translation log) (se3.py:158-176)."""
    R, t = matrix[..., :3, :3], matrix[..., :3, 3]
    w = so3_log(R)
    u = torch.linalg.solve(se3_V(w), t[..., :, None])[..., 0]
    return w, u


def rotation_to_matrix(rotation, parameterization, convention=None, degrees=False) -> torch.Tensor:
    """(..., 3, 3) rotation matrices from any supported parameterisation (se3.py:72-115)."""
    p = Parameterization(parameterization)
    if p == Parameterization.EULER:
        if convention is None:
            raise ValueError("convention must be specified for Euler angles (e.g., 'XYZ', 'ZYX')")
        return euler_to_rotmat(convention, rotation, degrees=degrees)
    if p == Parameterization.QUATERNION:
        return unitquat_to_rotmat(rotation)
    if p == Parameterization.QUATERNION_ADJUGATE:
        q = quaternion_adjugate_to_quaternion(rotation)  # WXYZ
        return unitquat_to_rotmat(torch.cat([q[..., 1:], q[..., :1]], dim=-1))
    if p == Parameterization.ROTATION_9D:
        return special_procrustes(rotation.reshape(*rotation.shape[:-1], 3, 3))
    return so3_exp(rotation)


def convert(rotation, translation, parameterization, convention=None, degrees=True, isocenter=None) -> torch.Tensor:
    """Camera pose (..., 4, 4) from a rotation parameterisation and camera centre (se3.py:28-69):
    t = R @ translation (+ isocenter); for se3_log the pair is exponentiated directly."""
    p = Parameterization(parameterization)
    if p == Parameterization.SE3_LOG:
        return se3_exp(rotation, translation)
    R = rotation_to_matrix(rotation, p, convention, degrees)
    t = matmul3(R, translation[..., :, None])[..., 0]
    if isocenter is not None:
        t = t + isocenter
    return make_se3(R, t)
