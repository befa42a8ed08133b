"""GPU parity of the single-pass adjoint (fused_raymarch_jac + fused_raymarch_bwd_jac):
the image is the plain forward's bit for bit, and the pose and camera
gradients formed from the stored per-ray sums match the float64 golden
vectors and the two-pass adjoint (the counterpart of nanodrr's
src/nanodrr/drr/_fused.py:477-531) on the same inputs."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import raymarch as orm
from tests.helpers import rel_to_max

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
GOLDEN = Path(__file__).resolve().parent / "golden" / "ref_camera_32.npz"
LAYOUTS = ["linear", "stencil", "quad"]


def _golden_inputs(name, ortho, C, layout):
    from nanodrr_amd._ops import pack_labels, pack_volume

    g = np.load(GOLDEN)
    vol = torch.from_numpy(g["volume"]).to(DEV)
    packed = None if layout == "linear" else pack_volume(vol, layout)
    lab = torch.from_numpy(g["label3"]).to(DEV) if C > 1 else torch.empty(0, dtype=torch.int8, device=DEV)
    if C > 1:
        lab = pack_labels(lab, layout)
    D, H, W = vol.shape
    w2v = np.eye(4)
    w2v[:3, 3] = [(W - 1) / 2, (H - 1) / 2, (D - 1) / 2]
    tgt = orm.make_tgt(g["k_inv"], g["sdd"], 16, 16)
    src = orm.make_src(ortho, tgt, g["sdd"])
    step = orm.step_size(src, tgt, 200)
    M = orm.camera_to_voxel(w2v, g["rt_inv"])
    geo = [torch.tensor(v, dtype=torch.float32, device=DEV) for v in (M, src, tgt, step)]
    go = torch.tensor(g[f"{name}_go"], dtype=torch.float32, device=DEV)
    return g, vol, lab, packed, geo, go


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("name,ortho,C", [("persp_c1", False, 1), ("ortho_c1", True, 1),
                                          ("persp_c3", False, 3), ("ortho_c3", True, 3)])
def test_single_pass_matches_golden_and_two_pass(name, ortho, C, layout):
    from nanodrr_amd._ops import fused_raymarch, fused_raymarch_bwd, fused_raymarch_bwd_jac, fused_raymarch_jac

    g, vol, lab, packed, (M, src, tgt, step), go = _golden_inputs(name, ortho, C, layout)
    out_plain = fused_raymarch(vol, lab, M, src, tgt, step, 200, C, 16, packed, layout)
    out, jac = fused_raymarch_jac(vol, lab, M, src, tgt, step, 200, C, 16, 7, packed, layout)
    assert torch.equal(out, out_plain)  # same march, same accumulation order
    gM, gsrc, gtgt, gstep = fused_raymarch_bwd_jac(go, jac, M, src, tgt, step, 16, True)
    _, gM2, gsrc2, gtgt2, gstep2 = fused_raymarch_bwd(go, vol, lab, M, src, tgt, step, 200, C, 16, False, True,
                                                      packed, layout)
    # against the two-pass adjoint: the same sums, the factor go * step moved out of them
    for a, b in ((gM, gM2), (gsrc, gsrc2), (gtgt, gtgt2), (gstep, gstep2)):
        assert rel_to_max(b.cpu(), a.cpu()) < 1e-5
    # against the float64 golden vectors (tolerances of the two-pass golden test)
    tol = 1e-4 if C == 1 else 2e-3
    assert rel_to_max(g[f"{name}_gM"], gM.cpu()) < tol
    gs = g[f"{name}_gsrc"] if ortho else g[f"{name}_gsrc"].sum(1, keepdims=True)
    gs_got = gsrc.cpu() if ortho else gsrc.sum(1, keepdim=True).cpu()
    assert rel_to_max(gs, gs_got) < tol
    assert rel_to_max(g[f"{name}_gtgt"], gtgt.cpu()) < tol
    assert rel_to_max(g[f"{name}_gstep"], gstep.cpu()) < tol


@pytest.mark.parametrize("C", [1, 2, 4, 6])
@pytest.mark.parametrize("layout", LAYOUTS)
def test_single_pass_matches_two_pass_on_phantom(C, layout):
    """64^3 phantom, 4 poses, 64 x 64 detector at 4 mm, S = 160 and S = 600 (compensated image sums)."""
    import nanodrr_amd.renderer as renderer
    from nanodrr_amd import render
    from nanodrr_amd.phantom import detector, make_phantom, phantom_subject, sample_poses

    ph = make_phantom((64, 64, 64), (4.0, 4.0, 4.0), seed=C)
    subject = phantom_subject(ph, C, torch.bfloat16 if layout == "stencil" else torch.float32, DEV)
    rt_inv, _, _ = sample_poses(4, seed=C, isocenter=subject.isocenter)
    k_inv, sdd = detector(1020.0, 4.0, 64, 64)
    k_inv, sdd, rt_inv = k_inv.to(DEV), sdd.to(DEV), rt_inv.to(DEV)
    wts = torch.randn(4, C, 64, 64, generator=torch.Generator().manual_seed(C)).to(DEV)
    old_layout, old_adjoint = renderer.VOLUME_LAYOUT, renderer.ADJOINT
    try:
        renderer.VOLUME_LAYOUT = layout
        for S in (160, 600):
            res = {}
            for adjoint in ("single_pass", "two_pass"):
                renderer.ADJOINT = adjoint
                rt = rt_inv.clone().requires_grad_(True)
                sd = sdd.clone().requires_grad_(True)
                out = render(subject, k_inv, rt, sd, 64, 64, n_samples=S)
                (out * wts).sum().backward()
                res[adjoint] = (out.detach(), rt.grad, sd.grad)
            assert torch.equal(res["single_pass"][0], res["two_pass"][0])
            assert rel_to_max(res["two_pass"][1][:, :3].cpu(), res["single_pass"][1][:, :3].cpu()) < 1e-5
            # d/d sdd is one scalar summed over every ray with cancellation: the two fp32
            # adjoints group its terms differently (chunks, per-class factoring)
            assert rel_to_max(res["two_pass"][2].cpu(), res["single_pass"][2].cpu()) < 5e-5
    finally:
        renderer.VOLUME_LAYOUT, renderer.ADJOINT = old_layout, old_adjoint


def test_single_pass_is_deterministic_and_validates():
    from nanodrr_amd._ops import fused_raymarch_bwd_jac, fused_raymarch_jac

    g, vol, lab, packed, (M, src, tgt, step), go = _golden_inputs("persp_c3", False, 3, "stencil")
    out, jac = fused_raymarch_jac(vol, lab, M, src, tgt, step, 200, 3, 16, 6, packed, "stencil")
    a = fused_raymarch_bwd_jac(go, jac, M, src, tgt, step, 16, False)[0]
    b = fused_raymarch_bwd_jac(go, jac, M, src, tgt, step, 16, False)[0]
    assert torch.equal(a, b)
    with pytest.raises(ValueError, match="7-plane"):
        fused_raymarch_bwd_jac(go, jac, M, src, tgt, step, 16, True)
    with pytest.raises(ValueError, match="jac_planes"):
        fused_raymarch_jac(vol, lab, M, src, tgt, step, 200, 3, 16, 5, packed, "stencil")


def test_single_pass_refuses_volume_gradients():
    """render() takes the two-pass adjoint whenever the volume requires grad."""
    import nanodrr_amd.renderer as renderer
    from nanodrr_amd import make_k_inv, make_rt_inv, render
    from tests.helpers import random_subject

    subject = random_subject(device=DEV)
    subject._image_hu = subject._image_hu.detach().requires_grad_(True)
    k_inv = make_k_inv(100.0, 1.0, 1.0, 0.0, 0.0, 16, 16, device=DEV)
    rt_inv = make_rt_inv(torch.tensor([[5.0, -3.0, 8.0]]), torch.tensor([[1.0, 80.0, -2.0]]), "AP",
                         torch.zeros(3)).to(DEV).requires_grad_(True)
    assert renderer.ADJOINT == "single_pass"
    out = render(subject, k_inv, rt_inv, torch.tensor([100.0], device=DEV), 16, 16, n_samples=100)
    out.sum().backward()
    assert subject._image_hu.grad is not None and rt_inv.grad is not None
