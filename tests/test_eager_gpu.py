"""The eager entry points (nanodrr_amd._ops.eager and the module-level eager
pose step and ZNCC) call the op implementations directly, skipping the
torch.library dispatch; they must give exactly what the registered custom ops
give, outputs and gradients, bit for bit."""

from __future__ import annotations

import pytest
import torch

from tests.helpers import random_subject

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _geometry(B=3, size=24, n_classes=2, layout="linear", seed=0):
    from nanodrr_amd.camera import make_k_inv
    from nanodrr_amd.renderer import _make_src, _make_tgt

    subject = random_subject(size, n_classes, seed, DEV)
    g = torch.Generator().manual_seed(seed)
    k_inv = make_k_inv(100.0, 1.5, 1.5, 0.0, 0.0, 12, 10).to(DEV)
    sdd = torch.tensor([100.0], device=DEV)
    tgt = _make_tgt(k_inv, sdd, 12, 10, DEV, torch.float32).contiguous()
    src = _make_src(False, tgt, sdd).contiguous()
    M = torch.zeros(B, 3, 4)
    M[:, :3, :3] = torch.eye(3) * 0.2 + 0.02 * torch.randn(B, 3, 3, generator=g)
    M[:, :, 3] = torch.tensor([size / 2, size / 2, -size]) + torch.randn(B, 3, generator=g)
    step = ((tgt - src).norm(dim=-1) / 63.0).contiguous()
    vol = subject.image[0, 0].contiguous()
    lab = subject.packed_labels(layout)
    packed = None if layout == "linear" else subject.packed_volume(layout)
    return vol, lab, M.to(DEV), src, tgt, step, packed, n_classes


def _run(op, vol, lab, M, src, tgt, step, packed, C, layout, jac, grads_for):
    ins = {"vol": vol, "M": M, "src": src, "tgt": tgt, "step": step}
    leaves = {k: v.detach().clone().requires_grad_(k in grads_for) for k, v in ins.items()}
    args = (leaves["vol"], lab, leaves["M"], leaves["src"], leaves["tgt"], leaves["step"], 64, C, 10)
    out = op(*args, 7, packed, layout)[0] if jac else op(*args, packed, layout)
    w = torch.randn(out.shape, generator=torch.Generator().manual_seed(5)).to(DEV)
    (out * w).sum().backward()
    return out.detach(), {k: leaves[k].grad for k in grads_for}


@pytest.mark.parametrize("layout", ["linear", "quad"])
def test_eager_raymarch_matches_custom_op(layout):
    from nanodrr_amd import _ops

    geo = _geometry(layout=layout)
    # two-pass op: volume, pose and camera gradients
    a = _run(_ops.eager.fused_raymarch, *geo, layout, False, ("vol", "M", "src", "tgt", "step"))
    b = _run(_ops.fused_raymarch, *geo, layout, False, ("vol", "M", "src", "tgt", "step"))
    assert torch.equal(a[0], b[0])
    for k in a[1]:
        if k == "vol":  # hardware atomics: the order of the additions varies
            torch.testing.assert_close(a[1][k], b[1][k], rtol=1e-5, atol=1e-6)
        else:
            assert torch.equal(a[1][k], b[1][k]), k
    # single-pass op: pose and camera gradients
    a = _run(_ops.eager.fused_raymarch_jac, *geo, layout, True, ("M", "src", "tgt", "step"))
    b = _run(_ops.fused_raymarch_jac, *geo, layout, True, ("M", "src", "tgt", "step"))
    assert torch.equal(a[0], b[0])
    for k in a[1]:
        assert torch.equal(a[1][k], b[1][k]), k


def test_eager_setup_ops_match_custom_ops():
    from nanodrr_amd import _ops
    from nanodrr_amd.camera import make_k_inv
    from nanodrr_amd.phantom import sample_poses

    k_inv = make_k_inv(1000.0, 1.0, 1.2, 3.0, -2.0, 20, 16).to(DEV)
    sdd = torch.tensor([1000.0, 1050.0], device=DEV)
    for x, y in zip(_ops.eager.make_rays(k_inv, sdd, 20, 16, False, 100), _ops.make_rays(k_inv, sdd, 20, 16, False, 100)):
        assert torch.equal(x, y)
    rt, _, _ = sample_poses(4, seed=1)
    w2v = torch.eye(4, dtype=torch.float64, device=DEV)
    w2v[:3, 3] = 12.5
    gM = torch.randn(4, 3, 4, generator=torch.Generator().manual_seed(2)).to(DEV)
    res = []
    for op in (_ops.eager.camera_to_voxel, _ops.camera_to_voxel):
        r = rt.to(DEV).requires_grad_(True)
        M = op(w2v, r)
        (M * gM).sum().backward()
        res.append((M.detach(), r.grad))
    assert torch.equal(res[0][0], res[1][0]) and torch.equal(res[0][1], res[1][1])


def test_eager_pose_step_and_zncc_match_custom_ops():
    from nanodrr_amd.metrics import _zncc, _zncc_eager
    from nanodrr_amd.registration import _pose_compose_eager, pose_compose

    g = torch.Generator().manual_seed(3)
    prefix = (torch.eye(4, dtype=torch.float64) + 0.1 * torch.randn(3, 4, 4, generator=g, dtype=torch.float64))
    suffix = (torch.eye(4, dtype=torch.float64) + 0.1 * torch.randn(3, 4, 4, generator=g, dtype=torch.float64))
    prefix, suffix = prefix.to(DEV).contiguous(), suffix.to(DEV).contiguous()
    rot0, xyz0 = 0.3 * torch.randn(3, 3, generator=g), 5 * torch.randn(3, 3, generator=g)
    gM = torch.randn(3, 3, 4, generator=g).to(DEV)
    res = []
    for op in (_pose_compose_eager, pose_compose):
        rot, xyz = rot0.to(DEV).requires_grad_(True), xyz0.to(DEV).requires_grad_(True)
        M = op(prefix, suffix, rot, xyz)
        (M * gM).sum().backward()
        res.append((M.detach(), rot.grad, xyz.grad))
    for a, b in zip(*res):
        assert torch.equal(a, b)
    x0 = torch.rand(2, 1, 40, 30, generator=g)
    y = (x0 + 0.3 * torch.rand(2, 1, 40, 30, generator=g)).to(DEV)
    res = []
    for op in (_zncc_eager, _zncc):
        x = x0.to(DEV).requires_grad_(True)
        s = op(x, y, 1e-7)[0]
        s.backward(torch.tensor([1.0, -2.0], device=DEV))
        res.append((s.detach(), x.grad))
    for a, b in zip(*res):
        assert torch.equal(a, b)
