"""GPU parity of the one-launch ray and camera set-up: make_rays against the
torch statement of nanodrr's _make_tgt/_make_src/step
(src/nanodrr/drr/renderer.py:82,101-132) and the float64 oracle, and
camera_to_voxel (with its adjoint) against float64 torch
(src/nanodrr/drr/backends.py:96)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import raymarch as orm

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _intrinsics(B, seed):
    from nanodrr_amd import make_k_inv

    g = torch.Generator().manual_seed(seed)
    ks = []
    for _ in range(B):
        sdd = 900.0 + 200.0 * float(torch.rand(1, generator=g))
        dx, dy = 0.5 + float(torch.rand(1, generator=g)), 0.5 + float(torch.rand(1, generator=g))
        x0, y0 = 10.0 * float(torch.rand(1, generator=g)) - 5.0, 10.0 * float(torch.rand(1, generator=g)) - 5.0
        ks.append(make_k_inv(sdd, dx, dy, x0, y0, 48, 40))
    return torch.cat(ks)


@pytest.mark.parametrize("orthographic", [False, True])
@pytest.mark.parametrize("bk,bs", [(1, 1), (3, 1), (1, 3), (3, 3)])
def test_make_rays_matches_torch_and_oracle(orthographic, bk, bs):
    from nanodrr_amd._ops import make_rays
    from nanodrr_amd.renderer import _make_src, _make_tgt

    k_inv = _intrinsics(bk, 7)
    sdd = torch.tensor([1020.0, 950.0, 1100.0][:bs])
    tgt, src, step = make_rays(k_inv.to(DEV), sdd.to(DEV), 48, 40, orthographic, 300)
    B = max(bk, bs)
    assert tgt.shape == (B, 48 * 40, 3) and step.shape == (B, 48 * 40)
    assert src.shape == ((B, 48 * 40, 3) if orthographic else (B, 1, 3))
    t_ref = _make_tgt(k_inv.double(), sdd.double(), 48, 40, "cpu", torch.float64)
    s_ref = _make_src(orthographic, t_ref, sdd.double())
    st_ref = (t_ref - s_ref).norm(dim=-1) / 299.0
    torch.testing.assert_close(tgt.cpu().double(), t_ref.expand(B, -1, -1), rtol=2e-7, atol=1e-4)
    torch.testing.assert_close(src.cpu().double(), s_ref.expand(B, -1, -1), rtol=2e-7, atol=1e-4)
    torch.testing.assert_close(step.cpu().double(), st_ref.expand(B, -1), rtol=4e-7, atol=0)
    # the float64 oracle of the rays (one camera)
    ot = orm.make_tgt(k_inv[:1].double().numpy(), sdd[:1].double().numpy(), 48, 40)
    np.testing.assert_allclose(tgt[:1].cpu().double().numpy(), ot, rtol=2e-7, atol=1e-4)


def test_camera_to_voxel_and_adjoint_match_float64():
    from nanodrr_amd._ops import camera_to_voxel
    from nanodrr_amd.phantom import sample_poses

    rt, _, _ = sample_poses(9, seed=3, rot_deg=40.0, trans_mm=50.0)
    w2v = torch.eye(4, dtype=torch.float64)
    w2v[:3, :3] = torch.diag(torch.tensor([1 / 0.8, 1 / 0.8, 1.0], dtype=torch.float64))
    w2v[:3, 3] = torch.tensor([255.5, 255.5, 199.5], dtype=torch.float64)
    r = rt.to(DEV).requires_grad_(True)
    M = camera_to_voxel(w2v.to(DEV), r)
    ref = (w2v @ rt.double())[:, :3]
    # FP64 products rounded once to fp32
    torch.testing.assert_close(M.detach().cpu().double(), ref, rtol=6e-8, atol=1e-12)
    gM = torch.randn(9, 3, 4, generator=torch.Generator().manual_seed(1))
    (M * gM.to(DEV)).sum().backward()
    r64 = rt.double().requires_grad_(True)
    ((w2v @ r64)[:, :3] * gM.double()).sum().backward()
    torch.testing.assert_close(r.grad.cpu().double(), r64.grad, rtol=6e-8, atol=1e-12)


def test_render_uses_native_setup_and_matches_torch_setup():
    """render() with the native set-up equals render() with tgt/src given from the torch statement."""
    import nanodrr_amd.renderer as renderer
    from nanodrr_amd import render
    from nanodrr_amd.phantom import detector, make_phantom, phantom_subject, sample_poses

    subject = phantom_subject(make_phantom((48, 48, 48), (5.0, 5.0, 5.0), seed=2), 1, torch.float32, DEV)
    rt_inv, _, _ = sample_poses(3, seed=2, isocenter=subject.isocenter)
    k_inv, sdd = detector(1020.0, 5.0, 40, 40)
    k_inv, sdd = k_inv.to(DEV), sdd.to(DEV)
    rt = rt_inv.to(DEV).requires_grad_(True)
    out = render(subject, k_inv, rt, sdd, 40, 40, n_samples=128)
    out.square().sum().backward()
    tgt = renderer._make_tgt(k_inv, sdd, 40, 40, DEV, torch.float32)
    src = renderer._make_src(False, tgt, sdd)
    rt2 = rt_inv.to(DEV).requires_grad_(True)
    out2 = render(subject, k_inv, rt2, sdd, 40, 40, n_samples=128, src=src, tgt=tgt)
    out2.square().sum().backward()
    assert (out - out2).abs().max() / out2.abs().max() < 1e-6
    assert (rt.grad - rt2.grad)[:, :3].abs().max() / rt2.grad[:, :3].abs().max() < 1e-5
