"""The per-pose sample-lattice offset of the single-pass forward (the `lattice`
argument of nanodrr_amd::fused_raymarch_jac / fused_raymarch_pose,
ndrr_raymarch_fwd_jac's lattice_offset): sample s of pose b at
t = (s + lattice[b]) / (S - 1). It must render what marching copies of the rays
shifted by lattice[b] steps renders (to fp32 rounding: the kernels place the
samples from the unshifted rays in double-float arithmetic), give the same pose
gradient through the unchanged contraction, and leave every bit unchanged when
the offsets are zero."""

from __future__ import annotations

import pytest
import torch

from tests.helpers import rel_to_max

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
H, W, S = 48, 40, 160


def _setup(C, layout, B=4):
    from nanodrr_amd import _ops
    from nanodrr_amd.phantom import detector, make_phantom, phantom_subject, sample_poses

    ph = make_phantom((64, 64, 64), (4.0, 4.0, 4.0), seed=C + 20)
    subject = phantom_subject(ph, C, torch.bfloat16 if layout == "stencil" else torch.float32, DEV)
    rt_inv = sample_poses(B, seed=C + 30, isocenter=subject.isocenter)[0].to(DEV)
    k_inv, sdd = detector(1020.0, 4.0, H, W)
    tgt, src, step = _ops.eager.make_rays(k_inv.to(DEV), sdd.to(DEV), H, W, False, S)
    M = _ops.camera_to_voxel(subject.world_to_voxel_f64(), rt_inv)
    vol = subject.image[0, 0].contiguous()
    lab = subject.packed_labels(layout) if C > 1 else vol.new_empty(0, dtype=torch.int8)
    packed = None if layout == "linear" else subject.packed_volume(layout)
    go = torch.randn(B, C, H * W, generator=torch.Generator().manual_seed(C)).to(DEV)
    return subject, vol, lab, packed, M, rt_inv, src, tgt, step, go


def _shifted(src, tgt, u):
    """src and tgt moved by u[b] steps along each ray, formed in float64 and rounded once."""
    d = (tgt.double() - src.double()) / float(S - 1)
    ub = u.double().reshape(-1, 1, 1)
    return (src.double() + ub * d).float().contiguous(), (tgt.double() + ub * d).float().contiguous()


@pytest.mark.parametrize("layout", ["linear", "quad", "stencil"])
@pytest.mark.parametrize("C", [1, 3])
def test_lattice_offset_matches_shifted_rays(C, layout):
    from nanodrr_amd import _ops

    _, vol, lab, packed, M, _, src, tgt, step, go = _setup(C, layout)
    u = torch.tensor([0.5, -0.25, 0.37, -0.5], device=DEV)
    out, jac = _ops.fused_raymarch_jac(vol, lab, M, src, tgt, step, S, C, W, 6, packed, layout, u)
    gM = _ops.fused_raymarch_bwd_jac(go, jac, M, src, tgt, step, W, False)[0]
    s2, t2 = _shifted(src, tgt, u)
    ref, jac_ref = _ops.fused_raymarch_jac(vol, lab, M, s2, t2, step, S, C, W, 6, packed, layout)
    gM_ref = _ops.fused_raymarch_bwd_jac(go, jac_ref, M, s2, t2, step, W, False)[0]
    assert rel_to_max(ref.cpu(), out.cpu()) < 1e-5
    # the pose gradient sums every ray's terms; a sample within ~1 ulp of a cell face may take the
    # neighbouring cell's derivative in one of the two placements
    assert rel_to_max(gM_ref.cpu(), gM.cpu()) < 1e-4
    # a real shift: the offset moves the samples by up to half a step
    assert rel_to_max(out.cpu(), fused_unshifted(vol, lab, M, src, tgt, step, C, packed, layout).cpu()) > 1e-4


def fused_unshifted(vol, lab, M, src, tgt, step, C, packed, layout):
    from nanodrr_amd import _ops

    return _ops.fused_raymarch_jac(vol, lab, M, src, tgt, step, S, C, W, 6, packed, layout)[0]


@pytest.mark.parametrize("layout", ["linear", "quad"])
def test_zero_lattice_offset_changes_no_bit(layout):
    from nanodrr_amd import _ops

    subject, vol, lab, packed, M, rt_inv, src, tgt, step, _ = _setup(2, layout)
    zero = torch.zeros(4, device=DEV)
    a = _ops.fused_raymarch_jac(vol, lab, M, src, tgt, step, S, 2, W, 7, packed, layout)
    b = _ops.fused_raymarch_jac(vol, lab, M, src, tgt, step, S, 2, W, 7, packed, layout, zero)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
    w2v = subject.world_to_voxel_f64()
    c = _ops.fused_raymarch_pose(vol, lab, w2v, rt_inv, src, tgt, step, S, 2, W, 0, packed, layout, zero)
    assert torch.equal(a[0], c[0])


def test_lattice_offset_with_camera_gradients():
    """The camera gradients of the shifted march are those of the unshifted rays: sample s sits at
    src + t (tgt - src) with the shifted t, so d/dsrc = (1 - t) and d/dtgt = t."""
    from nanodrr_amd import _ops

    _, vol, lab, packed, M, _, src, tgt, step, go = _setup(1, "quad")
    u = torch.tensor([0.5, -0.25, 0.37, -0.5], device=DEV)
    tg = tgt.expand(4, -1, -1).contiguous()
    res = []
    for shifted in (False, True):
        t_leaf = tg.clone().requires_grad_(True)
        if shifted:
            d = (t_leaf - src) / float(S - 1)
            ub = u.reshape(-1, 1, 1)
            out = _ops.fused_raymarch_jac(vol, lab, M, src + ub * d, t_leaf + ub * d, step, S, 1, W, 7, packed,
                                          "quad")[0]
        else:
            out = _ops.fused_raymarch_jac(vol, lab, M, src, t_leaf, step, S, 1, W, 7, packed, "quad", u)[0]
        (out * go).sum().backward()
        res.append(t_leaf.grad)
    # per-ray endpoint gradients: allow the rare cell-face flip of one sample
    scale = res[1].norm(dim=-1)
    err = (res[0] - res[1]).norm(dim=-1) / scale.clamp_min(1e-3 * scale.max())
    assert (err < 1e-4).float().mean() > 0.999


def test_lattice_offset_validates():
    from nanodrr_amd import _ops

    _, vol, lab, packed, M, _, src, tgt, step, _ = _setup(1, "linear")
    with pytest.raises(ValueError, match="lattice"):
        _ops.fused_raymarch_jac(vol, lab, M, src, tgt, step, S, 1, W, 6, None, "linear", torch.zeros(3, device=DEV))
    with pytest.raises(ValueError, match="lattice"):
        _ops.fused_raymarch_jac(vol, lab, M, src, tgt, step, S, 1, W, 6, None, "linear",
                                torch.zeros(4, device=DEV, dtype=torch.float64))


@pytest.mark.parametrize("C", [1, 3])
def test_lattice_offset_against_float64_oracle(C):
    """The kernels' offset lattice against the float64 oracle marching the rays shifted by
    offset_b steps in float64: the same sample positions, so the images agree to fp32 rounding."""
    import numpy as np

    from nanodrr_amd import _ops
    from nanodrr_amd.phantom import make_phantom
    from oracle import raymarch as orm

    _, vol, lab, packed, M, _, src, tgt, step, _ = _setup(C, "linear")
    ph = make_phantom((64, 64, 64), (4.0, 4.0, 4.0), seed=C + 20)
    vol64 = ph.mu.double().numpy()
    lab64 = ph.labels.long().clamp(max=C - 1).numpy() if C > 1 else None
    u = torch.tensor([0.5, -0.25, 0.37, -0.5], device=DEV)
    out, _ = _ops.fused_raymarch_jac(vol, lab, M, src, tgt, step, S, C, W, 6, None, "linear", u)
    rays = np.sort(np.random.default_rng(C).choice(H * W, size=400, replace=False))
    d = (tgt.double() - src.double()) / float(S - 1)
    ub = u.double().reshape(-1, 1, 1)
    s64 = (src.double() + ub * d)[:, rays].cpu().numpy()
    t64 = (tgt.double() + ub * d)[:, rays].cpu().numpy()
    st64 = step.double().expand(4, -1)[:, rays].cpu().numpy()
    ref = orm.forward(vol64, lab64, C, M.double().cpu().numpy(), s64, t64, st64, S)
    got = out[:, :, rays].double().cpu().numpy()
    assert np.abs(ref - got).max() / np.abs(ref).max() < 1e-5
