"""The NCC family on the native ZNCC kernels against the float64 oracle (GPU).

Whole-image NCC per channel, patch NCC (every p x p window, ncc.py:83-88),
the multiscale sum (ncc.py:40-62) and the gradient NCC (Sobel responses,
ncc.py:65-137), with their gradients; plus two Subject/DRR regressions that
need the kernels: a render on a packed layout after the HU -> mu parameters
change, and DRR.forward against render().
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from nanodrr_amd import render
from nanodrr_amd.metrics import (GradientNormalizedCrossCorrelation2d, MultiscaleNormalizedCrossCorrelation2d,
                                 NormalizedCrossCorrelation2d, patch_zncc, sobel)
from oracle import registration as oreg

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _images(B, C, H, W, seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.rand(B, C, H, W, generator=g)
    y = 0.6 * x + 0.4 * torch.rand(B, C, H, W, generator=g)
    return x, y


def _oracle_ncc(x, y, eps, patch=None):
    """(score (B,), grad (B, C, H, W)) of the channel mean of (window) ZNCCs, float64."""
    xn, yn = x.double().numpy(), y.double().numpy()
    B, C, H, W = xn.shape
    scores, grads = np.zeros((B, C)), np.zeros((B, C, H, W))
    for c in range(C):
        if patch is None:
            s, g = oreg.zncc(xn[:, c], yn[:, c], eps)
            g = g.reshape(B, H, W)
        else:
            s, g = oreg.patch_zncc(xn[:, c], yn[:, c], patch, eps)
        scores[:, c], grads[:, c] = s, g / C
    return scores.mean(1), grads


@pytest.mark.parametrize("C, patch", [(1, None), (3, None), (1, 1), (1, 5), (2, 7), (1, 24)])
def test_ncc_and_gradient_against_oracle(C, patch):
    x, y = _images(3, C, 24, 19 if patch != 24 else 24, seed=C + (patch or 0))
    eps = 1e-7 if patch is None else 1e-4
    ncc = NormalizedCrossCorrelation2d(patch_size=patch, eps=eps)
    xd = x.to(DEV).requires_grad_(True)
    score = ncc(xd, y.to(DEV))
    score.sum().backward()
    ref, gref = _oracle_ncc(x, y, eps, patch)
    np.testing.assert_allclose(score.detach().cpu().double().numpy(), ref, rtol=0, atol=2e-6)
    g = xd.grad.cpu().double().numpy()
    assert np.abs(g - gref).max() <= 1e-4 * np.abs(gref).max()  # a 1 x 1 window: both exactly zero


def test_patch_zncc_at_registration_size_and_determinism():
    """200 x 200 images, 13 x 13 windows (188^2 windows per image): score and gradient vs the
    oracle on two images, bitwise repeatable."""
    x, y = _images(2, 1, 200, 200, seed=9)
    xd, yd = x[:, 0].to(DEV), y[:, 0].to(DEV)
    s1 = patch_zncc(xd, yd, 13, 1e-4)
    s2 = patch_zncc(xd, yd, 13, 1e-4)
    assert torch.equal(s1, s2)
    ref, _ = oreg.patch_zncc(x[:, 0].double().numpy(), y[:, 0].double().numpy(), 13, 1e-4)
    np.testing.assert_allclose(s1.cpu().double().numpy(), ref, rtol=0, atol=2e-6)


def test_multiscale_is_the_weighted_sum():
    x, y = _images(2, 1, 20, 20, seed=4)
    ms = MultiscaleNormalizedCrossCorrelation2d([None, 5, 9], [0.5, 0.3, 0.2], eps=1e-4)
    got = ms(x.to(DEV), y.to(DEV)).cpu().double().numpy()
    ref = 0.5 * _oracle_ncc(x, y, 1e-4)[0] + 0.3 * _oracle_ncc(x, y, 1e-4, 5)[0] + 0.2 * _oracle_ncc(x, y, 1e-4, 9)[0]
    np.testing.assert_allclose(got, ref, rtol=0, atol=3e-6)
    with pytest.raises(ValueError, match="weight"):
        MultiscaleNormalizedCrossCorrelation2d([None, 5], [1.0])


@pytest.mark.parametrize("patch", [None, 5])
def test_gradient_ncc_is_the_ncc_of_the_sobel_responses(patch):
    x, y = _images(2, 1, 21, 18, seed=5)
    gncc = GradientNormalizedCrossCorrelation2d(patch_size=patch, sigma=1.0, eps=1e-4)
    xd = x.to(DEV).requires_grad_(True)
    score = gncc(xd, y.to(DEV))
    sx, sy = sobel(x, 1.0), sobel(y, 1.0)  # CPU filter, checked against scipy in test_subject.py
    ref, _ = _oracle_ncc(sx, sy, 1e-4, patch)
    np.testing.assert_allclose(score.detach().cpu().double().numpy(), ref, rtol=0, atol=3e-6)
    score.sum().backward()
    # the gradient flows through the filter into the image: check it by central differences
    xf = x.double().clone()
    k = (0, 0, 10, 9)
    h = 1e-3
    xp, xm = xf.clone(), xf.clone()
    xp[k] += h
    xm[k] -= h
    fd = (_oracle_ncc(sobel(xp, 1.0).double(), sy.double(), 1e-4, patch)[0][0]
          - _oracle_ncc(sobel(xm, 1.0).double(), sy.double(), 1e-4, patch)[0][0]) / (2 * h)
    assert xd.grad[k].item() == pytest.approx(fd, rel=2e-2, abs=1e-6)


def test_ncc_rejects_bad_inputs():
    x, y = _images(1, 1, 8, 8, seed=0)
    with pytest.raises(ValueError, match="same size"):
        NormalizedCrossCorrelation2d()(x.to(DEV), y[..., :7].to(DEV))
    with pytest.raises(ValueError, match="patch size"):
        patch_zncc(x[:, 0].to(DEV), y[:, 0].to(DEV), 9)
    with pytest.raises(ValueError, match="GPU"):
        patch_zncc(x[:, 0], y[:, 0], 3)


def _hu_phantom_subject(device):
    from nanodrr_amd import Subject
    from nanodrr_amd.phantom import make_phantom

    ph = make_phantom((48, 40, 44), (2.0, 2.0, 2.0), seed=3)
    return Subject.from_volume(ph.hu, spacing=ph.spacing, convert_to_mu=True).to(device)


def test_packed_render_follows_a_change_of_mu_parameters():
    """The quad copy is rebuilt when the mu image is: after each change the quad render equals a
    linear-layout render of the new image (the kernels read the volume itself there)."""
    from nanodrr_amd import renderer
    from nanodrr_amd.camera import make_k_inv, make_rt_inv

    subject = _hu_phantom_subject(DEV)
    k_inv = make_k_inv(1020.0, 2.0, 2.0, 0.0, 0.0, 32, 32).to(DEV)
    rt_inv = make_rt_inv(torch.tensor([[4.0, -6.0, 3.0]]), torch.tensor([[2.0, 600.0, -1.0]]), "AP",
                         subject.isocenter.cpu()).to(DEV)
    sdd = torch.tensor([1020.0], device=DEV)
    old = renderer.VOLUME_LAYOUT
    outs = []
    try:
        for mu_water in (0.0192, 0.025, 0.013, 0.025):
            subject.mu_water = mu_water
            renderer.VOLUME_LAYOUT = "quad"
            quad = render(subject, k_inv, rt_inv, sdd, 32, 32, n_samples=128)
            renderer.VOLUME_LAYOUT = "linear"
            lin = render(subject, k_inv, rt_inv, sdd, 32, 32, n_samples=128)
            torch.testing.assert_close(quad, lin, rtol=2e-6, atol=1e-6)
            outs.append(quad)
    finally:
        renderer.VOLUME_LAYOUT = old
    assert not torch.allclose(outs[0], outs[1])
    torch.testing.assert_close(outs[1], outs[3], rtol=0, atol=0)


def test_drr_module_renders_like_render():
    from nanodrr_amd.drr import DRR
    from tests.helpers import random_subject
    from nanodrr_amd.camera import make_rt_inv

    subject = random_subject(n_classes=3, device=DEV)
    drr = DRR.from_carm_intrinsics(100.0, 1.0, 1.0, 0.0, 0.0, 16, 16, device=DEV)
    rt_inv = make_rt_inv(torch.tensor([[5.0, -3.0, 8.0]]), torch.tensor([[1.0, 80.0, -2.0]]), "AP",
                         torch.zeros(3)).to(DEV)
    a = drr(subject, rt_inv, n_samples=200)
    b = render(subject, drr.k_inv, rt_inv, drr.sdd, 16, 16, n_samples=200)
    torch.testing.assert_close(a, b, rtol=0, atol=0)
