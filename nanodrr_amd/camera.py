"""Cone-beam camera: inverse intrinsics and camera-to-world extrinsics.

Mirrors src/nanodrr/camera/intrinsics.py:5-68 and
src/nanodrr/camera/extrinsics.py:7-110 (same arguments, conventions and
errors); the ZXY Euler conversions come from nanodrr_amd.geometry.
"""

from __future__ import annotations

import torch

from .geometry import convert, rotmat_to_euler_zxy

# Fixed re-orientation of the camera frame per acquisition (extrinsics.py:7-31):
# the camera looks along +z; AP puts the source anterior of the patient.
_ORIENTATION_MATRICES = {
    "AP": [[-1, 0, 0, 0], [0, 0, -1, 0], [0, -1, 0, 0], [0, 0, 0, 1]],
    "PA": [[-1, 0, 0, 0], [0, 0, 1, 0], [0, -1, 0, 0], [0, 0, 0, 1]],
    None: [[-1, 0, 0, 0], [0, -1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]],
}


def make_k_inv(sdd, delx, dely, x0, y0, height, width, dtype=None, device=None) -> torch.Tensor:
    """(1, 3, 3) inverse of K = [[fx, 0, cx], [0, fy, cy], [0, 0, 1]] with
    f = sdd / spacing and c = offset / spacing + size / 2 (intrinsics.py:53-68)."""
    fx, fy = sdd / delx, sdd / dely
    cx, cy = x0 / delx + width / 2.0, y0 / dely + height / 2.0
    return torch.tensor(
        [[[1.0 / fx, 0.0, -cx / fx], [0.0, 1.0 / fy, -cy / fy], [0.0, 0.0, 1.0]]], dtype=dtype, device=device
    )


def _orientation(orientation, device, dtype) -> torch.Tensor:
    if orientation not in _ORIENTATION_MATRICES:
        raise ValueError(f"Unknown orientation {orientation!r}. Expected one of: 'AP', 'PA', or None.")
    return torch.tensor(_ORIENTATION_MATRICES[orientation], device=device, dtype=dtype)


def _iso(isocenter, ref: torch.Tensor) -> torch.Tensor:
    if isocenter is not None:
        return isocenter.to(device=ref.device, dtype=ref.dtype)
    return torch.zeros(3, device=ref.device, dtype=ref.dtype)


def make_rt_inv(rotation, translation, orientation="AP", isocenter=None) -> torch.Tensor:
    """(B, 4, 4) camera-to-world matrices: pose(ZXY Euler degrees, centre
    R @ translation + isocenter) @ reorientation (extrinsics.py:34-57)."""
    reorient = _orientation(orientation, rotation.device, rotation.dtype)
    pose = convert(rotation, translation, "euler", convention="ZXY", isocenter=_iso(isocenter, rotation))
    return pose @ reorient


def invert_rt_inv(extrinsic_inv, orientation="AP", isocenter=None):
    """Recover (rotation ZXY degrees, translation) from make_rt_inv's output (extrinsics.py:60-91)."""
    R_orient = _orientation(orientation, extrinsic_inv.device, extrinsic_inv.dtype)[:3, :3]
    iso = _iso(isocenter, extrinsic_inv).view(1, 3)
    R_pose = extrinsic_inv[..., :3, :3] @ R_orient.T
    t_pose = extrinsic_inv[..., :3, 3]
    translation = torch.einsum("bij,bj->bi", R_pose.transpose(-1, -2), t_pose - iso)
    rotation = rotmat_to_euler_zxy(R_pose, degrees=True)
    return rotation, translation
