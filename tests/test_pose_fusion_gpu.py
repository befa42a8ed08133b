"""The pose-fused march (nanodrr_amd::fused_raymarch_pose) composes the camera ->
voxel-index affine M = (world_to_voxel @ rt_inv)[:3] inside the march kernels
and forms the rt_inv gradient inside the pose reduction. Both steps are the same
FP64 arithmetic with the same single rounding as camera_to_voxel and its
adjoint, so it must give exactly what camera_to_voxel followed by
fused_raymarch_jac gives: image, jac planes, rt_inv and camera gradients, bit
for bit."""

from __future__ import annotations

import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
H, W = 48, 40


def _setup(C, layout, B=4, S=160):
    from nanodrr_amd._ops import eager
    from nanodrr_amd.phantom import detector, make_phantom, phantom_subject, sample_poses

    ph = make_phantom((64, 64, 64), (4.0, 4.0, 4.0), seed=C)
    subject = phantom_subject(ph, C, torch.bfloat16 if layout == "stencil" else torch.float32, DEV)
    rt_inv = sample_poses(B, seed=C + 10, isocenter=subject.isocenter)[0].to(DEV)
    k_inv, sdd = detector(1020.0, 4.0, H, W)
    tgt, src, step = eager.make_rays(k_inv.to(DEV), sdd.to(DEV), H, W, False, S)
    vol = subject.image[0, 0].contiguous()
    lab = subject.packed_labels(layout) if C > 1 else vol.new_empty(0, dtype=torch.int8)
    packed = None if layout == "linear" else subject.packed_volume(layout)
    w = torch.randn(B, C, H * W, generator=torch.Generator().manual_seed(C)).to(DEV)
    return subject, vol, lab, packed, rt_inv, src, tgt, step, S, w


def _leaves(rt_inv, src, tgt, step, gcam):
    return (rt_inv.clone().requires_grad_(True), *(t.clone().requires_grad_(gcam) for t in (src, tgt, step)))


@pytest.mark.parametrize("gcam", [False, True])
@pytest.mark.parametrize("layout", ["linear", "quad", "stencil"])
@pytest.mark.parametrize("C", [1, 3])
def test_pose_fused_march_matches_composed(C, layout, gcam):
    from nanodrr_amd import _ops

    subject, vol, lab, packed, rt_inv, src, tgt, step, S, w = _setup(C, layout)
    w2v = subject.world_to_voxel_f64()
    planes = 7 if gcam else 6
    res = []
    for path in ("composed", "op", "eager"):
        rt, s, t, st = _leaves(rt_inv, src, tgt, step, gcam)
        if path == "composed":
            M = _ops.camera_to_voxel(w2v, rt)
            out, jac = _ops.fused_raymarch_jac(vol, lab, M, s, t, st, S, C, W, planes, packed, layout)
        else:
            op = _ops.fused_raymarch_pose if path == "op" else _ops.eager.fused_raymarch_pose
            out, jac = op(vol, lab, w2v, rt, s, t, st, S, C, W, planes, packed, layout)
        (out * w).sum().backward()
        res.append((out.detach(), jac, rt.grad) + ((s.grad, t.grad, st.grad) if gcam else ()))
    for other in res[1:]:
        for i, (a, b) in enumerate(zip(res[0], other)):
            assert torch.equal(a, b), i
    # all four rows of the rt_inv gradient are formed (the bottom one through world_to_voxel's
    # translation column)
    assert torch.count_nonzero(res[0][2]) > 12 * rt_inv.shape[0]


@pytest.mark.parametrize("layout", ["linear", "quad"])
def test_pose_fused_march_without_jac_matches_plain(layout):
    """jac_planes = 0: the forward alone, equal to camera_to_voxel then fused_raymarch."""
    from nanodrr_amd import _ops

    subject, vol, lab, packed, rt_inv, src, tgt, step, S, _ = _setup(2, layout)
    w2v = subject.world_to_voxel_f64()
    M = _ops.camera_to_voxel(w2v, rt_inv)
    ref = _ops.fused_raymarch(vol, lab, M, src, tgt, step, S, 2, W, packed, layout)
    out, jac = _ops.eager.fused_raymarch_pose(vol, lab, w2v, rt_inv, src, tgt, step, S, 2, W, 0, packed, layout)
    assert jac.numel() == 0 and torch.equal(out, ref)


def test_pose_fused_backward_with_shared_gradient():
    """One (C, N) upstream gradient broadcast over the poses (batch stride 0) is read in place."""
    from nanodrr_amd import _ops

    subject, vol, lab, packed, rt_inv, src, tgt, step, S, w = _setup(2, "quad")
    w2v = subject.world_to_voxel_f64()
    M = _ops.camera_to_voxel(w2v, rt_inv)
    _, jac = _ops.fused_raymarch_pose(vol, lab, w2v, rt_inv, src, tgt, step, S, 2, W, 7, packed, "quad")
    go = w[:1].expand(4, -1, -1)
    a = _ops.fused_raymarch_pose_bwd(go, jac, w2v, rt_inv, src, tgt, step, W, True)
    b = _ops.fused_raymarch_bwd_jac(go, jac, M, src, tgt, step, W, True)
    assert torch.equal(a[0], _ops.camera_to_voxel_bwd(w2v, b[0]))
    for x, y in zip(a[1:], b[1:]):
        assert torch.equal(x, y)


@pytest.mark.parametrize("requires_grad", [False, True])
def test_render_routes_through_pose_fusion(requires_grad):
    """render() (single-pass adjoint, volume without grad) equals camera_to_voxel then march()."""
    import nanodrr_amd.renderer as renderer
    from nanodrr_amd import _ops, render
    from nanodrr_amd.phantom import detector

    subject, _, _, _, rt_inv, _, _, _, _, w = _setup(3, "quad")
    k_inv, sdd = detector(1020.0, 4.0, H, W)
    k_inv, sdd = k_inv.to(DEV), sdd.to(DEV)
    old = renderer.VOLUME_LAYOUT, renderer.ADJOINT
    try:
        renderer.VOLUME_LAYOUT, renderer.ADJOINT = "quad", "single_pass"
        rt = rt_inv.clone().requires_grad_(requires_grad)
        out = render(subject, k_inv, rt, sdd, H, W, n_samples=200)
        rt_ref = rt_inv.clone().requires_grad_(requires_grad)
        tgt, src, step = _ops.eager.make_rays(k_inv, sdd, H, W, False, 200)
        M = _ops.eager.camera_to_voxel(subject.world_to_voxel_f64(), rt_ref)
        ref = renderer.march(subject, M, src, tgt, step, 200, H, W)
        assert torch.equal(out.detach(), ref.detach())
        if requires_grad:
            wts = w.reshape(out.shape)
            (out * wts).sum().backward()
            (ref * wts).sum().backward()
            assert torch.equal(rt.grad, rt_ref.grad)
    finally:
        renderer.VOLUME_LAYOUT, renderer.ADJOINT = old


def test_pose_fused_march_validates():
    from nanodrr_amd import _ops

    subject, vol, lab, packed, rt_inv, src, tgt, step, S, _ = _setup(1, "linear")
    w2v = subject.world_to_voxel_f64()
    with pytest.raises(ValueError, match="jac_planes"):
        _ops.fused_raymarch_pose(vol, lab, w2v, rt_inv, src, tgt, step, S, 1, W, 5, None, "linear")
    with pytest.raises(ValueError, match="world_to_voxel"):
        _ops.fused_raymarch_pose(vol, lab, w2v.float(), rt_inv, src, tgt, step, S, 1, W, 6, None, "linear")
    with pytest.raises(ValueError, match="rt_inv"):
        _ops.fused_raymarch_pose(vol, lab, w2v, rt_inv[:, :3].contiguous(), src, tgt, step, S, 1, W, 6, None,
                                 "linear")
    _, jac = _ops.fused_raymarch_pose(vol, lab, w2v, rt_inv, src, tgt, step, S, 1, W, 6, None, "linear")
    with pytest.raises(ValueError, match="7-plane"):
        _ops.fused_raymarch_pose_bwd(torch.ones(4, 1, H * W, device=DEV), jac, w2v, rt_inv, src, tgt, step, W, True)
