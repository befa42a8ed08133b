"""Pins the CPU oracle before it is trusted as the checker of the HIP kernels.

Each check is chosen so that a plausible mistake in oracle.raymarch (a
dropped lerp term, a swapped axis, a sign error in the adjoint, a wrong
label rounding) fails one of them.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import raymarch as orm
from oracle import torch_ref
from tests.helpers import oracle_inputs, random_volume, ref_camera_np, rel_to_max


def _identity_M(shape):
    """Camera space = world space, voxel (i, j, k) at world (i, j, k) - centre."""
    D, H, W = shape
    M = np.zeros((1, 3, 4))
    M[0, :, :3] = np.eye(3)
    M[0, :, 3] = [(W - 1) / 2, (H - 1) / 2, (D - 1) / 2]
    return M


def test_unit_impulse_integral_equals_one():
    """Known answer of tests/drr/test_renderer.py:52-91: a ray through a lone
    unit voxel integrates to 1 (S = 1001, ray from x = -1.5 to 1.5 mm)."""
    vol = np.zeros((3, 3, 3))
    vol[1, 1, 1] = 1.0
    src = np.array([[[-1.5, 0.0, 0.0]]])
    tgt = np.array([[[1.5, 0.0, 0.0]]])
    S = 1001
    out = orm.forward(vol, None, 1, _identity_M(vol.shape), src, tgt, orm.step_size(src, tgt, S), S)
    assert out.shape == (1, 1, 1)
    assert out[0, 0, 0] == pytest.approx(1.0, rel=1e-3, abs=1e-3)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("spacing", [1.0, 1.5])
def test_constant_volume_chord_length(axis, spacing):
    """A constant volume c interpolates to c on [0, n-1] with linear ramps to 0
    at -1 and n, so an axis-aligned interior ray integrates to c * n * spacing."""
    shape = (18, 20, 22)  # D, H, W
    c = 0.7
    vol = np.full(shape, c)
    n = shape[::-1][axis]  # voxels along the ray's axis (x -> W, y -> H, z -> D)
    D, H, W = shape
    M = np.zeros((1, 3, 4))
    M[0, :, :3] = np.eye(3) / spacing
    M[0, :, 3] = [(W - 1) / 2, (H - 1) / 2, (D - 1) / 2]
    a = np.zeros(3)
    a[axis] = 1.0
    off = np.array([0.3, -1.2, 2.4]) * spacing
    off[axis] = 0.0
    src = (off - a * (n / 2 + 6) * spacing)[None, None]
    tgt = (off + a * (n / 2 + 6) * spacing)[None, None]
    S = 8001
    out = orm.forward(vol, None, 1, M, src, tgt, orm.step_size(src, tgt, S), S)
    assert out[0, 0, 0] == pytest.approx(c * n * spacing, rel=2e-4)


@pytest.mark.parametrize("orthographic", [False, True])
@pytest.mark.parametrize("n_classes", [1, 3])
def test_forward_matches_grid_sample_float64(orthographic, n_classes):
    """Independent implementation: ATen grid_sample (bilinear/nearest,
    align_corners=False) in float64 through the reference's grid coordinates."""
    image, label = random_volume(32, n_classes, seed=0)
    k_inv, rt_inv, sdd, h, w = ref_camera_np()
    inp = oracle_inputs(image, label, n_classes, k_inv, rt_inv, sdd, h, w, 200, orthographic)
    out = orm.forward(**inp)
    ref = torch_ref.march_voxel_M(
        image.double(), label.double(), n_classes, torch.from_numpy(inp["M"]), torch.from_numpy(inp["src"]),
        torch.from_numpy(inp["tgt"]), torch.from_numpy(inp["step"]), 200).numpy()
    assert out.shape == ref.shape == (1, n_classes, h * w)
    assert rel_to_max(ref, out) < 1e-11


@pytest.mark.parametrize("orthographic", [False, True])
@pytest.mark.parametrize("n_classes", [1, 3])
def test_backward_matches_autograd_float64(orthographic, n_classes):
    """The restated adjoint equals torch autograd of the grid_sample march for
    every input: the pose affine M, ray endpoints, step and volume."""
    image, label = random_volume(32, n_classes, seed=1)
    k_inv, rt_inv, sdd, h, w = ref_camera_np()
    inp = oracle_inputs(image, label, n_classes, k_inv, rt_inv, sdd, h, w, 200, orthographic)
    rng = np.random.default_rng(1)
    go = rng.standard_normal((1, n_classes, h * w))
    got = orm.backward(go, **inp)

    leaves = {k: torch.from_numpy(inp[k]).clone().requires_grad_(True) for k in ("M", "src", "tgt", "step")}
    img = image.double().requires_grad_(True)
    out = torch_ref.march_voxel_M(img, label.double(), n_classes, leaves["M"], leaves["src"], leaves["tgt"],
                                  leaves["step"], 200)
    (out * torch.from_numpy(go)).sum().backward()

    assert rel_to_max(leaves["M"].grad, got["gM"]) < 1e-9
    gsrc = got["gsrc"] if orthographic else got["gsrc"].sum(axis=1, keepdims=True)
    assert rel_to_max(leaves["src"].grad, gsrc) < 1e-9
    assert rel_to_max(leaves["tgt"].grad, got["gtgt"]) < 1e-9
    assert rel_to_max(leaves["step"].grad, got["gstep"]) < 1e-9
    assert rel_to_max(img.grad.reshape(got["gvol"].shape), got["gvol"]) < 1e-9


def _smooth_volume(n=24):
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64),) * 3, indexing="ij")
    c = (n - 1) / 2
    return (np.exp(-((x - c - 2) ** 2 + (y - c) ** 2 + (z - c + 1) ** 2) / 30.0)
            + 0.5 * np.exp(-((x - c + 4) ** 2 + (y - c - 3) ** 2 + (z - c) ** 2) / 12.0))


def test_pose_gradient_central_differences():
    """d(sum(go * out))/dM by central differences (no torch involved)."""
    vol = _smooth_volume()
    k_inv, rt_inv, sdd, h, w = ref_camera_np(12, 12)
    tgt = orm.make_tgt(k_inv * np.array([[[2.0, 1, 1], [1, 2.0, 1], [1, 1, 1]]]), sdd, h, w)
    src = orm.make_src(False, tgt, sdd)
    S = 160
    step = orm.step_size(src, tgt, S)
    w2v = np.eye(4)
    w2v[:3, 3] = 11.5
    M = orm.camera_to_voxel(w2v, rt_inv)
    go = np.random.default_rng(3).standard_normal((1, 1, h * w))
    g = orm.backward(go, vol, None, 1, M, src, tgt, step, S, need_gvol=False)["gM"]
    fd = np.zeros_like(M)
    for i in range(3):
        for j in range(4):
            hstep = 1e-6 if j < 3 else 1e-4
            Mp, Mm = M.copy(), M.copy()
            Mp[0, i, j] += hstep
            Mm[0, i, j] -= hstep
            fp = (go * orm.forward(vol, None, 1, Mp, src, tgt, step, S)).sum()
            fm = (go * orm.forward(vol, None, 1, Mm, src, tgt, step, S)).sum()
            fd[0, i, j] = (fp - fm) / (2 * hstep)
    assert rel_to_max(fd, g) < 2e-3


def test_volume_gradient_adjoint_identity():
    """The march is linear in the volume: <go, A v> = <A^T go, v>."""
    image, _ = random_volume(20, 1, seed=4)
    k_inv, rt_inv, sdd, h, w = ref_camera_np()
    inp = oracle_inputs(image, None, 1, k_inv, rt_inv, sdd, h, w, 150)
    go = np.random.default_rng(4).standard_normal((1, 1, h * w))
    lhs = (go * orm.forward(**inp)).sum()
    rhs = (orm.backward(go, **inp)["gvol"] * inp["vol"]).sum()
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_channel_sum_equals_single_channel():
    """With every label in range, the per-class channels sum to the C = 1 render."""
    image, label = random_volume(32, 4, seed=5)
    k_inv, rt_inv, sdd, h, w = ref_camera_np()
    multi = orm.forward(**oracle_inputs(image, label, 4, k_inv, rt_inv, sdd, h, w, 200))
    single = orm.forward(**oracle_inputs(image, label, 1, k_inv, rt_inv, sdd, h, w, 200))
    assert rel_to_max(single[:, 0], multi.sum(axis=1)) < 1e-13


def test_label_rounding_is_half_to_even():
    """A ray along x at y = z = 0.5 (a tie) must read the label of row 0 (even),
    as grid_sample's nearest mode and _fused.py:28-32 do."""
    lab = np.zeros((3, 3, 5), np.int64)
    lab[0, 0, :] = 1  # (z=0, y=0) row: class 1; row (z=1,y=1) etc. class 0
    vol = np.ones((3, 3, 5))
    M = np.zeros((1, 3, 4))
    M[0, :, :3] = np.eye(3)
    src = np.array([[[0.0, 0.5, 0.5]]])
    tgt = np.array([[[4.0, 0.5, 0.5]]])
    out = orm.forward(vol, lab, 2, M, src, tgt, orm.step_size(src, tgt, 5), 5)
    assert out[0, 0, 0] == 0.0 and out[0, 1, 0] > 0.0


def test_inside_count_closed_form_matches_brute_force():
    """nanodrr_amd.roofline.inside_samples (closed form) vs the oracle's
    sample-by-sample count, on the reference camera and 4 random poses."""
    from nanodrr_amd.phantom import detector, sample_poses
    from nanodrr_amd.roofline import inside_samples

    rt, _, _ = sample_poses(4, seed=7, dtype=torch.float64)
    k_inv, sdd = detector(1020.0, 1.0, 40, 40, dtype=torch.float64)
    shape = (64, 64, 64)
    w2v = torch.eye(4, dtype=torch.float64) / 4.0
    w2v[3, 3] = 1.0
    w2v[:3, 3] = 31.5
    M = (w2v @ rt)[:, :3, :]
    tgt = torch_ref.make_tgt(k_inv, sdd, 40, 40)
    src = torch.zeros(1, 1, 3, dtype=torch.float64)
    got = inside_samples(M, src, tgt, 256, shape).numpy()
    ref = orm.inside_count(M.numpy(), src.numpy(), tgt.numpy(), 256, shape)
    assert ref.sum() > 0
    assert np.abs(got - ref).max() <= 1
    assert abs(int(got.sum()) - int(ref.sum())) <= max(2, ref.sum() // 10000)


def test_coefficient_records_reproduce_the_corner_interpolant():
    """Exact (float64) coefficient records evaluate to the zero-padded trilinear interpolant and its
    gradient everywhere, inside and around the volume; float32 records differ by rounding only."""
    rng = np.random.default_rng(11)
    vol = rng.standard_normal((5, 6, 7))
    coef = orm.coef_records(vol, np.float64)
    assert coef.shape == (6, 7, 8, 8)
    p = rng.uniform(-1.5, 7.5, size=(4000, 3))
    val, gx, gy, gz = orm._coef_eval(coef, p)
    corners, f, _ = orm._corners(vol, p)
    np.testing.assert_allclose(val, orm._lerp(corners, f), atol=1e-13)
    h = 1e-6  # the gradient, against central differences of the corner form
    for axis, g in enumerate((gx, gy, gz)):
        e = np.zeros(3)
        e[axis] = h
        fd = (orm._lerp(*orm._corners(vol, p + e)[:2]) - orm._lerp(*orm._corners(vol, p - e)[:2])) / (2 * h)
        cellwise = np.all(np.floor(p + e) == np.floor(p - e), axis=1)  # no cell face in between
        np.testing.assert_allclose(g[cellwise], fd[cellwise], atol=1e-7)
    c32 = orm.coef_records(vol, np.float32).astype(np.float64)
    assert np.abs(c32 - coef).max() < 1e-5 * np.abs(vol).max()


def test_oracle_forward_and_backward_from_records_equal_the_corner_form():
    image, label = random_volume(16, 3, 2)
    k_inv, rt_inv, sdd, h, w = ref_camera_np(8, 8)
    inp = oracle_inputs(image, label, 3, k_inv, rt_inv, sdd, h, w, 64)
    coef = orm.coef_records(inp["vol"], np.float64)
    np.testing.assert_allclose(orm.forward(**inp, coef=coef), orm.forward(**inp), rtol=1e-12, atol=1e-14)
    go = np.random.default_rng(0).standard_normal((1, 3, 64))
    a = orm.backward(go, **inp, need_gvol=False, coef=coef)
    b = orm.backward(go, **inp, need_gvol=False)
    for k in ("gM", "gsrc", "gtgt", "gstep"):
        np.testing.assert_allclose(a[k], b[k], rtol=1e-10, atol=1e-12)
