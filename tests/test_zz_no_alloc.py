"""The allocation-free contraction (ndrr_raymarch_bwd_jac_pose_noalloc) against
ndrr_raymarch_bwd_jac_pose with a caller workspace, through the C ABI.

Both reduce the same partials in the same order, so results must agree bit for
bit. Two calls in a row, the second on a smaller input, show that the fixed
scratch array's leftover contents from the first call do not leak into the
second."""

from __future__ import annotations

import pytest
import torch

from nanodrr_amd import _lib
from nanodrr_amd._ops import _stream, fused_raymarch_pose, make_rays
from nanodrr_amd.camera import make_k_inv, make_rt_inv
from tests.helpers import random_subject

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _case(batch, det, n_classes, seed):
    subject = random_subject(size=32, n_classes=n_classes, seed=seed, device=DEV)
    g = torch.Generator().manual_seed(seed)
    rot = (torch.rand(batch, 3, generator=g) * 2 - 1) * 10.0
    xyz = (torch.rand(batch, 3, generator=g) * 2 - 1) * 3.0 + torch.tensor([0.0, 80.0, 0.0])
    rt_inv = make_rt_inv(rot, xyz, "AP", torch.zeros(3)).to(DEV).contiguous()
    k_inv = make_k_inv(100.0, 0.5, 0.5, 0.0, 0.0, det, det).to(DEV)
    sdd = torch.tensor([100.0], device=DEV)
    tgt, src, step = make_rays(k_inv, sdd, det, det, False, 200)
    w2v = subject.world_to_voxel_f64()
    layout = "quad"
    packed = subject.packed_volume(layout)
    lab = subject.packed_labels(layout) if n_classes > 1 else subject.label_i8.new_empty(0)
    _, jac = fused_raymarch_pose(subject.image[0, 0], lab, w2v, rt_inv, src, tgt, step, 200, n_classes, det, 7,
                                 packed, layout)
    go = torch.randn(batch, n_classes, det * det, generator=g).to(DEV)
    return go, jac, w2v, rt_inv, src, tgt, step, det


def _contract(noalloc, go, jac, w2v, rt_inv, src, tgt, step, width):
    lib = _lib.lib()
    B, planes, C, N = jac.shape
    g_rt = torch.full((B, 4, 4), float("nan"), device=DEV)
    gsrc = torch.full((B, N, 3), float("nan"), device=DEV)
    gtgt = torch.full((B, N, 3), float("nan"), device=DEV)
    gstep = torch.full((B, N), float("nan"), device=DEV)
    strides = (0, 0, 0, 0)  # src (1, 1, 3), tgt (1, N, 3), step (1, N): shared by every pose
    head = (go.data_ptr(), C * N, jac.data_ptr(), planes, w2v.data_ptr(), rt_inv.data_ptr(), src.data_ptr(),
            tgt.data_ptr(), step.data_ptr())
    tail = (g_rt.data_ptr(), gsrc.data_ptr(), gtgt.data_ptr(), gstep.data_ptr(), 1, B, N, C, width, *strides,
            _stream(DEV))
    if noalloc:
        rc = lib.ndrr_raymarch_bwd_jac_pose_noalloc(*head, *tail)
        _lib.check(rc, "ndrr_raymarch_bwd_jac_pose_noalloc")
    else:
        ws = torch.empty(lib.ndrr_bwd_workspace_floats(B, N, width), device=DEV)
        rc = lib.ndrr_raymarch_bwd_jac_pose(*head, ws.data_ptr(), *tail)
        _lib.check(rc, "ndrr_raymarch_bwd_jac_pose")
    torch.cuda.synchronize()
    return g_rt, gsrc, gtgt, gstep


def test_noalloc_contraction_matches_workspace_entry_point():
    for batch, det, n_classes, seed in ((4, 64, 3, 0), (2, 24, 1, 1)):  # the second call is smaller
        print(f"noalloc case B={batch} det={det} C={n_classes}", flush=True)
        args = _case(batch, det, n_classes, seed)
        ref = _contract(False, *args)
        got = _contract(True, *args)
        for r, o in zip(ref, got):
            assert torch.isfinite(r).all()
            assert torch.equal(r, o)


def test_noalloc_rejects_work_beyond_its_scratch_before_launch():
    lib = _lib.lib()
    B, N = 2000, 40000  # 2000 * ceil(40000 / 768) * 12 floats > the 2^20 of the array
    assert B * -(-N // 768) * 12 > (1 << 20)
    rc = lib.ndrr_raymarch_bwd_jac_pose_noalloc(1, 0, 1, 7, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, B, N, 1, 200,
                                                0, 0, 0, 0, None)
    assert rc == -5
