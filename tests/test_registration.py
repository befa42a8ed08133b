"""CPU checks of the registration row: the oracle's pose step and ZNCC pinned
against independent mathematics, and the host-side metrics and validation.

The reference ships no registration tests; the pose step follows
src/nanodrr/registration/registration.py:64-68 and the NCC
src/nanodrr/metrics/ncc.py:8-37.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation
from scipy.stats import pearsonr

from oracle import geometry as og
from oracle import registration as oreg


def _problem(B=5, seed=0, scale=0.5):
    rng = np.random.default_rng(seed)
    prefix = np.tile(np.eye(4), (B, 1, 1))
    suffix = np.tile(np.eye(4), (B, 1, 1))
    prefix[:, :3, :] = rng.normal(size=(B, 3, 4)) * 3
    suffix[:, :3, :3] = Rotation.random(B, random_state=seed).as_matrix()
    suffix[:, :3, 3] = rng.normal(size=(B, 3)) * 20
    rot = rng.normal(size=(B, 3)) * scale
    rot[0] = 0.0
    rot[1] = [1e-9, -2e-9, 5e-10]
    xyz = rng.normal(size=(B, 3)) * 10
    return prefix, suffix, rot, xyz


def _torch_compose(prefix, suffix, rot, xyz):
    """Independent float64 torch statement: R = matrix_exp([w]x)."""
    x, y, z = rot.unbind(-1)
    o = torch.zeros_like(x)
    K = torch.stack([o, -z, y, z, o, -x, -y, x, o], dim=-1).reshape(-1, 3, 3)
    R = torch.linalg.matrix_exp(K)
    top = torch.cat([R, (R @ xyz[:, :, None])], dim=-1)
    bottom = torch.tensor([0.0, 0, 0, 1], dtype=torch.float64).expand(len(rot), 1, 4)
    T = torch.cat([top, bottom], dim=1)
    return (prefix @ T @ suffix)[:, :3]


def test_pose_compose_matches_scipy_rotvec():
    prefix, suffix, rot, xyz = _problem()
    M = oreg.pose_compose(prefix, suffix, rot, xyz)
    for b in range(len(rot)):
        T = np.eye(4)
        R = Rotation.from_rotvec(rot[b]).as_matrix()
        T[:3, :3] = R
        T[:3, 3] = R @ xyz[b]
        np.testing.assert_allclose(M[b], (prefix[b] @ T @ suffix[b])[:3], rtol=0, atol=1e-12)


def test_pose_compose_matches_reference_pose_form():
    """rt_inv @ pivot @ T @ pivot_inv then world_to_voxel (registration.py:64-68, backends.py:84-86)."""
    from nanodrr_amd.geometry import convert

    rng = np.random.default_rng(3)
    B = 4
    rt_inv = og.make_rt_inv(rng.uniform(-20, 20, (B, 3)), rng.uniform(-30, 30, (B, 3)) + [0, 600, 0], "AP")
    w2v = np.eye(4)
    w2v[:3, 3] = [127.5, 127.5, 127.5]
    c = rng.normal(size=(B, 3)) * 5
    pivot = np.tile(np.eye(4), (B, 1, 1))
    pivot[:, :3, 3] = c
    pivot_inv = np.tile(np.eye(4), (B, 1, 1))
    pivot_inv[:, :3, 3] = -c
    rot, xyz = rng.normal(size=(B, 3)) * 0.3, rng.normal(size=(B, 3)) * 8
    T = convert(torch.from_numpy(rot), torch.from_numpy(xyz), "so3_log").numpy()
    ref = (w2v @ rt_inv @ pivot @ T @ pivot_inv)[:, :3]
    M = oreg.pose_compose(w2v @ rt_inv @ pivot, pivot_inv, rot, xyz)
    np.testing.assert_allclose(M, ref, rtol=0, atol=1e-10)


def test_pose_compose_bwd_matches_torch_autograd():
    prefix, suffix, rot, xyz = _problem(seed=1)
    gM = np.random.default_rng(7).normal(size=(len(rot), 3, 4))
    grot, gxyz = oreg.pose_compose_bwd(prefix, suffix, rot, xyz, gM)
    r = torch.tensor(rot, requires_grad=True)
    v = torch.tensor(xyz, requires_grad=True)
    M = _torch_compose(torch.from_numpy(prefix), torch.from_numpy(suffix), r, v)
    (M * torch.from_numpy(gM)).sum().backward()
    np.testing.assert_allclose(grot, r.grad.numpy(), rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(gxyz, v.grad.numpy(), rtol=1e-9, atol=1e-9)


def test_pose_compose_bwd_matches_finite_differences():
    prefix, suffix, rot, xyz = _problem(seed=2, scale=1.0)
    gM = np.random.default_rng(8).normal(size=(len(rot), 3, 4))
    grot, gxyz = oreg.pose_compose_bwd(prefix, suffix, rot, xyz, gM)
    h = 1e-6
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        for arr, g in ((rot, grot), (xyz, gxyz)):
            plus, minus = arr.copy(), arr.copy()
            plus += e
            minus -= e
            args_p = (prefix, suffix, plus, xyz) if arr is rot else (prefix, suffix, rot, plus)
            args_m = (prefix, suffix, minus, xyz) if arr is rot else (prefix, suffix, rot, minus)
            fd = ((oreg.pose_compose(*args_p) - oreg.pose_compose(*args_m)) * gM).sum(axis=(1, 2)) / (2 * h)
            np.testing.assert_allclose(g[:, k], fd, rtol=1e-6, atol=1e-6)


def test_drot_at_zero_is_generator():
    for k in range(3):
        np.testing.assert_array_equal(oreg.drot(np.zeros(3))[k], og.hat(np.eye(3)[k]))


def test_zncc_matches_pearson():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(3, 64 * 48))
    y = 0.4 * x + rng.normal(size=x.shape)
    score, _ = oreg.zncc(x, y, eps=0.0)
    for b in range(3):
        assert abs(score[b] - pearsonr(x[b], y[b]).statistic) < 1e-12


def test_zncc_affine_invariance_and_extremes():
    rng = np.random.default_rng(1)
    x = rng.uniform(size=(2, 500))
    s_pos, _ = oreg.zncc(x, 3.5 * x + 2.0, eps=0.0)
    s_neg, _ = oreg.zncc(x, -0.2 * x + 7.0, eps=0.0)
    np.testing.assert_allclose(s_pos, 1.0, atol=1e-13)
    np.testing.assert_allclose(s_neg, -1.0, atol=1e-13)


@pytest.mark.parametrize("eps", [0.0, 1e-7, 1e-2])
def test_zncc_grad_matches_finite_differences(eps):
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 40))
    y = rng.normal(size=(2, 40)) + x
    _, g = oreg.zncc(x, y, eps)
    h = 1e-6
    for i in (0, 7, 39):
        xp, xm = x.copy(), x.copy()
        xp[:, i] += h
        xm[:, i] -= h
        fd = (oreg.zncc(xp, y, eps)[0] - oreg.zncc(xm, y, eps)[0]) / (2 * h)
        np.testing.assert_allclose(g[:, i], fd, rtol=1e-6, atol=1e-9)
    # shift invariance: the gradient sums to zero
    np.testing.assert_allclose(g.sum(axis=1), 0.0, atol=1e-12)


def test_oracle_patch_zncc_is_pinned():
    """The window ZNCC of the oracle: one window covering the image is the whole-image ZNCC;
    the score is the mean of a brute-force loop over windows; the gradient matches central
    finite differences."""
    rng = np.random.default_rng(3)
    x = rng.random((2, 7, 6))
    y = rng.random((2, 7, 6)) + 0.5 * x
    s_full, g_full = oreg.patch_zncc(x[:, :6, :6], y[:, :6, :6], 6, 1e-7)
    s_img, g_img = oreg.zncc(x[:, :6, :6], y[:, :6, :6], 1e-7)
    np.testing.assert_allclose(s_full, s_img, rtol=1e-13)
    np.testing.assert_allclose(g_full.reshape(2, -1), g_img, rtol=1e-12, atol=1e-15)
    s, g = oreg.patch_zncc(x, y, 3, 1e-4)
    for b in range(2):
        vals = []
        for i in range(5):
            for j in range(4):
                vals.append(pearson_eps(x[b, i:i + 3, j:j + 3].ravel(), y[b, i:i + 3, j:j + 3].ravel(), 1e-4))
        assert s[b] == pytest.approx(np.mean(vals), rel=1e-12)
    h = 1e-6
    for (b, i, j) in ((0, 0, 0), (1, 3, 2), (0, 6, 5), (1, 2, 5)):
        xp, xm = x.copy(), x.copy()
        xp[b, i, j] += h
        xm[b, i, j] -= h
        fd = (oreg.patch_zncc(xp, y, 3, 1e-4)[0][b] - oreg.patch_zncc(xm, y, 3, 1e-4)[0][b]) / (2 * h)
        assert g[b, i, j] == pytest.approx(fd, rel=1e-6, abs=1e-9)


def pearson_eps(a, b, eps):
    """ZNCC of two vectors with eps added to each population variance, from numpy's covariance."""
    c = np.cov(a, b, bias=True)
    return c[0, 1] / np.sqrt((c[0, 0] + eps) * (c[1, 1] + eps))


def test_double_geodesic_known_answers():
    from nanodrr_amd.metrics import DoubleGeodesicSE3

    p1 = torch.eye(4, dtype=torch.float64)[None].repeat(2, 1, 1)
    p2 = p1.clone()
    p2[0, :3, :3] = torch.from_numpy(Rotation.from_rotvec([0, 0, 0.3]).as_matrix())
    p2[1, :3, 3] = torch.tensor([3.0, 4.0, 0.0])
    rot, xyz, dg = DoubleGeodesicSE3(sdd=1000.0, eps=0.0)(p1, p2)
    np.testing.assert_allclose(rot.numpy(), [500 * 0.3, 0.0], atol=1e-9)
    np.testing.assert_allclose(xyz.numpy(), [0.0, 5.0], atol=1e-12)
    np.testing.assert_allclose(dg.numpy(), [150.0, 5.0], atol=1e-9)


def test_mtre_known_answer():
    from nanodrr_amd.metrics import mtre

    a = torch.eye(4, dtype=torch.float64)[None]
    b = a.clone()
    b[0, :3, 3] = torch.tensor([1.0, 2.0, 2.0])  # camera-to-world offset: every fiducial moves by 3 mm
    fid = torch.randn(16, 3, dtype=torch.float64) * 50
    np.testing.assert_allclose(mtre(a, b, fid).numpy(), [3.0], atol=1e-12)
    np.testing.assert_allclose(mtre(a, a, fid).numpy(), [0.0], atol=1e-12)


def test_pose_compose_rejects_host_tensors():
    from nanodrr_amd.registration import pose_compose

    eye = torch.eye(4, dtype=torch.float64)[None]
    with pytest.raises(ValueError, match="device tensor"):
        pose_compose(eye, eye, torch.zeros(1, 3), torch.zeros(1, 3))


def test_zncc_rejects_host_tensors():
    from nanodrr_amd.metrics import zncc

    with pytest.raises(ValueError, match="GPU"):
        zncc(torch.rand(2, 1, 4, 4), torch.rand(2, 1, 4, 4))


def test_settled_iteration_is_the_start_of_the_final_run_below_1mm():
    from nanodrr_amd.posenet import _settled

    trace = [(5, [2.0, 0.5, 0.9]), (10, [0.8, 0.6, 0.8]), (15, [1.2, 0.4, 0.7]), (20, [0.9, 0.3, 0.6]),
             (25, [0.7, 1.1, 0.5])]
    assert _settled(trace, 0) == 20    # dipped below at 10, rose at 15, below from 20 on
    assert _settled(trace, 1) is None  # ends above 1 mm
    assert _settled(trace, 2) == 5     # below from the first check
    assert _settled([], 0) is None


def test_register_rejects_unknown_adjoint_and_graph_mode():
    """Argument errors come back before anything touches a device."""
    from nanodrr_amd.registration import register

    with pytest.raises(ValueError, match="adjoint"):
        register(None, None, adjoint="finite-differences")
    with pytest.raises(ValueError, match="explicit"):
        register(None, None, graph=True, adjoint="autograd")
    with pytest.raises(ValueError, match="iters_per_graph"):
        register(None, None, graph=True, iters_per_graph=0)
