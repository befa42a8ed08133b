"""Host-side checks of Subject, DRR and the image filters (no GPU).

hu_to_mu against the known answers of its definition (preprocess.py:9-50), the
derived-tensor cache (mu image, invalidation on parameter change, in-place
edits and buffer replacement), DRR's state_dict and ray cache
(src/nanodrr/drr/drr.py:9-106), and the Sobel / Gaussian filters of the
gradient NCC against scipy.ndimage (ncc.py:91-137).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch
from scipy import ndimage

from nanodrr_amd import Subject
from nanodrr_amd._memo import Derived
from nanodrr_amd.camera import make_k_inv
from nanodrr_amd.drr import DRR
from nanodrr_amd.metrics import gaussian_taps, sobel
from nanodrr_amd.renderer import _make_tgt
from nanodrr_amd.subject import MU_BONE, MU_WATER, hu_to_mu

REFERENCE_BUFFERS = {"_image_hu", "label", "world_to_grid", "isocenter", "voxel_to_world", "world_to_voxel",
                     "voxel_to_grid"}


@pytest.mark.parametrize("mu_water, mu_bone, hu_bone", [(MU_WATER, MU_BONE, 1000.0), (0.02, 0.05, 1500.0)])
def test_hu_to_mu_known_answers(mu_water, mu_bone, hu_bone):
    hu = torch.tensor([-3000.0, -1000.0, -500.0, 0.0, hu_bone / 2, hu_bone, 2 * hu_bone], dtype=torch.float64)
    mu = hu_to_mu(hu, mu_water, mu_bone, hu_bone).numpy()
    # air (and anything below it) has no attenuation; water has mu_water; bone has mu_bone
    assert mu[0] == 0.0 and mu[1] == 0.0
    assert mu[3] == pytest.approx(mu_water, rel=1e-15)
    assert mu[5] == pytest.approx(mu_bone, rel=1e-15)
    # linear on each side of water: the midpoints are the means of the end points
    assert mu[2] == pytest.approx(mu_water / 2, rel=1e-15)
    assert mu[4] == pytest.approx((mu_water + mu_bone) / 2, rel=1e-15)
    assert mu[6] == pytest.approx(2 * mu_bone - mu_water, rel=1e-15)
    # continuous at water
    eps = 1e-9
    lo, hi = hu_to_mu(torch.tensor([-eps, eps], dtype=torch.float64), mu_water, mu_bone, hu_bone).numpy()
    assert hi - lo == pytest.approx(eps * (mu_water / 1000 + (mu_bone - mu_water) / hu_bone), rel=1e-6)


def test_hu_to_mu_clamps_negative_results_to_zero():
    """Parameters that make a branch negative (mu_water < 0 below water, mu_bone < 0 above it)
    give zero attenuation, never a negative one."""
    hu = torch.linspace(-1000, 1000, 11, dtype=torch.float64)
    assert (hu_to_mu(hu, -0.01, 0.05, 1000.0) >= 0).all()
    assert (hu_to_mu(hu, 0.02, -0.05, 1000.0) >= 0).all()


def _hu_subject(**mu):
    g = torch.Generator().manual_seed(0)
    hu = (torch.rand(1, 1, 6, 5, 4, generator=g) * 3000 - 1000).float()
    lab = torch.zeros_like(hu)
    v2w = torch.eye(4)
    return Subject(hu, lab, v2w, torch.linalg.inv(v2w), Subject._make_voxel_to_grid(hu.shape), torch.zeros(3),
                   **mu)


def test_subject_buffers_match_the_reference_state_dict():
    s = _hu_subject()
    assert set(s.state_dict()) == REFERENCE_BUFFERS
    assert s.n_classes == 1
    with pytest.raises(TypeError, match="mu_air"):
        _hu_subject(mu_air=0.0)


def test_subject_mu_image_follows_its_parameters():
    s = _hu_subject()
    a = s.image
    assert s.image is a  # cached
    torch.testing.assert_close(a, hu_to_mu(s._image_hu), rtol=0, atol=0)
    s.mu_water = MU_WATER  # same value: nothing changes
    assert s.image is a
    s.mu_water = 0.03
    b = s.image
    assert b is not a
    torch.testing.assert_close(b, hu_to_mu(s._image_hu, 0.03), rtol=0, atol=0)
    # set_mu overrides without touching the stored parameters; a parameter change ends it
    s.set_mu(mu_bone=0.08)
    torch.testing.assert_close(s.image, hu_to_mu(s._image_hu, 0.03, 0.08), rtol=0, atol=0)
    assert s.mu_bone == MU_BONE
    s.hu_bone = 1200.0
    torch.testing.assert_close(s.image, hu_to_mu(s._image_hu, 0.03, MU_BONE, 1200.0), rtol=0, atol=0)
    # an in-place edit of the HU volume, and a replaced buffer, are seen
    s._image_hu.add_(10.0)
    torch.testing.assert_close(s.image, hu_to_mu(s._image_hu, 0.03, MU_BONE, 1200.0), rtol=0, atol=0)
    s._image_hu = s._image_hu * 0.5
    torch.testing.assert_close(s.image, hu_to_mu(s._image_hu, 0.03, MU_BONE, 1200.0), rtol=0, atol=0)
    s.convert_to_mu = False
    assert s.image is s._image_hu


def test_subject_move_drops_derived_tensors():
    s = _hu_subject()
    _ = s.image
    assert "mu" in s._derived
    s.to(torch.float64)
    assert "mu" not in s._derived
    assert s.image.dtype == torch.float64


def test_derived_is_keyed_on_identity_and_version_not_address():
    d = Derived()
    calls = []

    def make(t):
        calls.append(1)
        return t * 2

    a = torch.ones(1000)
    d.get("x", a, None, lambda: make(a))
    d.get("x", a, None, lambda: make(a))
    assert len(calls) == 1
    a.mul_(3)  # version bump
    assert d.get("x", a, None, lambda: make(a))[0] == 6
    assert len(calls) == 2
    d.get("x", a, "other", lambda: make(a))  # extra key
    assert len(calls) == 3
    b = torch.ones(1000)  # a different tensor, whatever its address
    assert d.get("x", b, "other", lambda: make(b))[0] == 2
    assert len(calls) == 4


def test_drr_state_dict_and_rays():
    k_inv = make_k_inv(1020.0, 2.0, 2.0, 0.0, 0.0, 20, 30)
    drr = DRR(k_inv, torch.tensor([1020.0]), 20, 30)
    assert set(drr.state_dict()) == {"k_inv", "sdd"}
    tgt = drr.tgt
    assert drr.tgt is tgt and tgt.shape == (1, 600, 3)
    torch.testing.assert_close(tgt, _make_tgt(k_inv, torch.tensor([1020.0]), 20, 30, "cpu", torch.float32))
    assert drr.src.shape == (1, 1, 3) and (drr.src == 0).all()
    drr.height = 10  # an intrinsic changes: the rays follow
    assert drr.tgt.shape == (1, 300, 3)
    drr.sdd.mul_(0.5)  # in place
    torch.testing.assert_close(drr.tgt, _make_tgt(k_inv, torch.tensor([510.0]), 10, 30, "cpu", torch.float32))
    drr.orthographic = True
    assert drr.src.shape == (1, 300, 3)
    torch.testing.assert_close(drr.src[..., 2], drr.tgt[..., 2] - 510.0)


def test_drr_rays_carry_intrinsic_gradients():
    k_inv = make_k_inv(1020.0, 2.0, 2.0, 0.0, 0.0, 4, 4).requires_grad_(True)
    drr = DRR(k_inv, torch.tensor([1020.0]), 4, 4)
    for _ in range(2):  # a fresh graph per use: nothing stale is cached
        drr.tgt.sum().backward()
    assert k_inv.grad is not None and torch.isfinite(k_inv.grad).all()


def test_drr_from_carm_intrinsics():
    drr = DRR.from_carm_intrinsics(1020.0, 0.5, 0.75, 3.0, -2.0, 8, 6, dtype=torch.float64)
    torch.testing.assert_close(drr.k_inv, make_k_inv(1020.0, 0.5, 0.75, 3.0, -2.0, 8, 6).double())
    assert drr.sdd.dtype == torch.float64 and drr.sdd.item() == 1020.0


def test_gaussian_taps():
    for sigma in (0.5, 1.0, 2.3):
        w = gaussian_taps(sigma).double().numpy()
        assert len(w) == (int(6 * sigma + 1) | 1)
        assert w.sum() == pytest.approx(1.0, rel=1e-6)
        np.testing.assert_allclose(w, w[::-1], rtol=1e-6)
        c = len(w) // 2
        assert w[c + 1] / w[c] == pytest.approx(np.exp(-0.5 / sigma**2), rel=1e-5)


@pytest.mark.parametrize("sigma", [0.0, 1.0, 1.5])
def test_sobel_matches_scipy(sigma):
    """Gaussian blur with mirrored borders (torch's "reflect" pad is scipy's "mirror"), then the
    Sobel stencils (cross-correlation, zeros beyond the border)."""
    img = torch.rand(2, 1, 17, 13, generator=torch.Generator().manual_seed(1))
    got = sobel(img, sigma).double().numpy()
    for b in range(2):
        x = img[b, 0].double().numpy()
        if sigma > 0:
            taps = gaussian_taps(sigma).double().numpy()
            x = ndimage.correlate1d(ndimage.correlate1d(x, taps, axis=1, mode="mirror"), taps, axis=0, mode="mirror")
        kx = np.array([[1, 0, -1], [2, 0, -2], [1, 0, -1]], np.float64)
        ky = np.array([[1, 2, 1], [0, 0, 0], [-1, -2, -1]], np.float64)
        np.testing.assert_allclose(got[b, 0], ndimage.correlate(x, kx, mode="constant"), atol=1e-5)
        np.testing.assert_allclose(got[b, 1], ndimage.correlate(x, ky, mode="constant"), atol=1e-5)
