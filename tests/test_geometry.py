"""Camera and SE(3) conventions: product (torch) vs scipy, the float64 oracle
and the reference's round-trip tests (tests/camera/test_extrinsics.py:11-67,
tests/geometry/test_se3.py:13-65, tests/geometry/test_transform.py:6-65)."""

from __future__ import annotations

import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from nanodrr_amd import geometry as G
from nanodrr_amd.camera import invert_rt_inv, make_k_inv, make_rt_inv
from oracle import geometry as og

ORIENTATIONS = ["AP", "PA", None]


def test_euler_zxy_matches_scipy_intrinsic():
    ang = torch.tensor([[30.0, -15.0, 45.0], [5.0, -3.0, 8.0], [-170.0, 80.0, 12.0]], dtype=torch.float64)
    R = G.euler_to_rotmat("ZXY", ang, degrees=True).numpy()
    ref = Rotation.from_euler("ZXY", ang.numpy(), degrees=True).as_matrix()
    np.testing.assert_allclose(R, ref, atol=1e-14)
    np.testing.assert_allclose(og.euler_zxy_deg(ang.numpy()), ref, atol=1e-14)


@pytest.mark.parametrize("conv", ["XYZ", "zyx", "YXZ"])
def test_euler_conventions_match_scipy(conv):
    ang = torch.tensor([[12.0, -40.0, 71.0]], dtype=torch.float64)
    R = G.euler_to_rotmat(conv, ang, degrees=True).numpy()
    np.testing.assert_allclose(R, Rotation.from_euler(conv, ang.numpy(), degrees=True).as_matrix(), atol=1e-14)


def test_make_k_inv_is_inverse_intrinsics():
    k_inv = make_k_inv(1020.0, 2.0, 1.5, 3.0, -4.0, 200, 180, dtype=torch.float64)[0].numpy()
    fx, fy = 1020 / 2.0, 1020 / 1.5
    cx, cy = 3.0 / 2.0 + 180 / 2, -4.0 / 1.5 + 200 / 2
    K = np.array([[fx, 0, cx], [0, fy, cy], [0, 0, 1]])
    np.testing.assert_allclose(k_inv @ K, np.eye(3), atol=1e-12)


@pytest.mark.parametrize("orientation", ORIENTATIONS)
def test_make_rt_inv_matches_oracle(orientation):
    rot = torch.tensor([[10.0, 20.0, -30.0], [5.0, -3.0, 8.0]], dtype=torch.float64)
    tr = torch.tensor([[50.0, 75.0, 300.0], [1.0, 80.0, -2.0]], dtype=torch.float64)
    iso = torch.tensor([10.0, -20.0, 5.0], dtype=torch.float64)
    got = make_rt_inv(rot, tr, orientation, iso).numpy()
    ref = og.make_rt_inv(rot.numpy(), tr.numpy(), orientation, iso.numpy())
    np.testing.assert_allclose(got, ref, atol=1e-12)


@pytest.mark.parametrize("orientation", ORIENTATIONS)
def test_rt_inv_roundtrip(orientation):
    rotation = torch.tensor([[30.0, -15.0, 45.0]])
    translation = torch.tensor([[100.0, -200.0, 500.0]])
    ext = make_rt_inv(rotation, translation, orientation=orientation)
    r, t = invert_rt_inv(ext, orientation=orientation)
    torch.testing.assert_close(r, rotation)
    torch.testing.assert_close(t, translation)


@pytest.mark.parametrize("orientation", ORIENTATIONS)
def test_rt_inv_roundtrip_with_isocenter(orientation):
    rotation = torch.tensor([[10.0, 20.0, -30.0]])
    translation = torch.tensor([[50.0, 75.0, 300.0]])
    iso = torch.tensor([10.0, -20.0, 5.0])
    ext = make_rt_inv(rotation, translation, orientation=orientation, isocenter=iso)
    r, t = invert_rt_inv(ext, orientation=orientation, isocenter=iso)
    torch.testing.assert_close(r, rotation)
    torch.testing.assert_close(t, translation)


@pytest.mark.parametrize("orientation", ORIENTATIONS)
def test_rt_inv_roundtrip_batch(orientation):
    torch.manual_seed(0)
    rotation = torch.rand(8, 3) * 360 - 180
    translation = torch.rand(8, 3) * 1000 - 500
    ext = make_rt_inv(rotation, translation, orientation=orientation)
    r, t = invert_rt_inv(ext, orientation=orientation)
    torch.testing.assert_close(make_rt_inv(r, t, orientation=orientation), ext, atol=1e-3, rtol=1e-3)
    torch.testing.assert_close(t, translation, atol=1e-3, rtol=1e-3)


def test_invalid_orientation():
    with pytest.raises(ValueError, match="Unknown orientation"):
        make_rt_inv(torch.zeros(1, 3), torch.zeros(1, 3), orientation="LAT")
    with pytest.raises(ValueError, match="Unknown orientation"):
        invert_rt_inv(torch.eye(4).unsqueeze(0), orientation="LAT")


def test_so3_exp_matches_matrix_exp_and_scipy():
    torch.manual_seed(1)
    w = torch.cat([torch.randn(6, 3, dtype=torch.float64), torch.tensor([[0.0, 0.0, 0.0], [1e-9, -2e-9, 0.0]],
                                                                          dtype=torch.float64)])
    R = G.so3_exp(w)
    np.testing.assert_allclose(R.numpy(), torch.linalg.matrix_exp(G.hat(w)).numpy(), atol=1e-13)
    np.testing.assert_allclose(R.numpy(), Rotation.from_rotvec(w.numpy()).as_matrix(), atol=1e-13)


def test_se3_exp_matches_matrix_exp_including_zero():
    """exp of the 4x4 twist [[w]x, u; 0, 0]; finite and exact at w = 0 (unlike se3.py:211-219)."""
    torch.manual_seed(2)
    w = torch.cat([torch.randn(4, 3, dtype=torch.float64) * 0.7, torch.zeros(1, 3, dtype=torch.float64),
                   torch.full((1, 3), 1e-7, dtype=torch.float64)])
    u = torch.randn(6, 3, dtype=torch.float64) * 10
    T = G.se3_exp(w, u)
    twist = torch.zeros(6, 4, 4, dtype=torch.float64)
    twist[:, :3, :3] = G.hat(w)
    twist[:, :3, 3] = u
    np.testing.assert_allclose(T.numpy(), torch.linalg.matrix_exp(twist).numpy(), atol=1e-11)
    for k in range(6):
        np.testing.assert_allclose(T[k].numpy(), og.se3_exp(w[k].numpy(), u[k].numpy()), atol=1e-12)


def test_se3_log_exp_roundtrip():
    torch.manual_seed(3)
    B = 4
    R = torch.linalg.matrix_exp(G.hat(torch.randn(B, 3) * 0.1))
    t = torch.randn(B, 3)
    T = G.make_se3(R, t)
    w, u = G.se3_log_map(T)
    torch.testing.assert_close(G.convert(w, u, G.Parameterization.SE3_LOG), T, atol=1e-4, rtol=1e-4)


def test_so3_log_near_pi():
    w = torch.tensor([[np.pi - 1e-6, 0.0, 0.0], [0.0, 2.0, 1.0]], dtype=torch.float64)
    w[1] *= (np.pi - 1e-5) / w[1].norm()
    R = G.so3_exp(w)
    np.testing.assert_allclose(G.so3_exp(G.so3_log(R)).numpy(), R.numpy(), atol=1e-8)


def test_rotation_parameterisations_are_orthonormal():
    torch.manual_seed(4)
    q = torch.randn(5, 4)
    q = q / q.norm(dim=-1, keepdim=True)
    for rot, p in ((torch.randn(5, 3) * 45, "euler"), (q, "quaternion"), (torch.randn(5, 9), "rotation_9d"),
                   (torch.randn(5, 3), "so3_log")):
        R = G.rotation_to_matrix(rot, p, convention="ZYX", degrees=True)
        torch.testing.assert_close(R.transpose(-1, -2) @ R, torch.eye(3).expand(5, 3, 3), atol=1e-5, rtol=1e-5)
        torch.testing.assert_close(torch.det(R), torch.ones(5), atol=1e-5, rtol=1e-5)
    qd = q.double() / q.double().norm(dim=-1, keepdim=True)
    np.testing.assert_allclose(G.unitquat_to_rotmat(qd).numpy(), Rotation.from_quat(qd.numpy()).as_matrix(), atol=1e-12)


def test_quaternion_adjugate_unit_norm():
    torch.manual_seed(5)
    q = torch.randn(5, 4)
    q = q / q.norm(dim=-1, keepdim=True)
    A = torch.zeros(5, 4, 4)
    scales = torch.linspace(1.0, 2.0, steps=4)
    for i in range(4):
        A[:, :, i] = q * scales[i]
    A = 0.5 * (A + A.transpose(-1, -2))
    i, j = torch.triu_indices(4, 4)
    n = G.quaternion_adjugate_to_quaternion(A[:, i, j]).norm(dim=-1)
    torch.testing.assert_close(n, torch.ones_like(n), atol=1e-5, rtol=1e-5)


def test_transform_point_identity_translation_composition():
    pts = torch.randn(2, 5, 3)
    torch.testing.assert_close(G.transform_point(torch.eye(4).expand(2, 4, 4), pts), pts)
    A = torch.eye(4)
    A[:3, :3] = torch.tensor([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    A[:3, 3] = torch.tensor([1.0, 0.0, 0.0])
    Bm = torch.eye(4)
    Bm[:3, 3] = torch.tensor([0.0, 2.0, -1.0])
    p = torch.randn(1, 10, 3)
    torch.testing.assert_close(G.transform_point(Bm[None], G.transform_point(A[None], p)),
                               G.transform_point((Bm @ A)[None], p))
    torch.testing.assert_close(G.transform_point(Bm[None], p), p + Bm[:3, 3])
