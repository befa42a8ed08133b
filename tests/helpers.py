"""Shared fixtures: the reference test geometry and random subjects.

Mirrors the helpers of nanodrr's tests (tests/drr/test_fused.py:11-47,
tests/drr/test_renderer.py:8-49) with numpy float64 twins for the oracle.
"""

from __future__ import annotations

import numpy as np
import torch

from oracle import geometry as og
from oracle import raymarch as orm


def random_volume(size: int = 32, n_classes: int = 1, seed: int = 0):
    """torch.rand volume and floor-random labels, as make_random_subject (test_fused.py:11-34)."""
    torch.manual_seed(seed)
    image = torch.rand(1, 1, size, size, size, dtype=torch.float32)
    if n_classes > 1:
        label = (torch.rand_like(image) * n_classes).floor().clamp(max=n_classes - 1)
    else:
        label = torch.zeros_like(image)
    return image, label


def random_subject(size=32, n_classes=1, seed=0, device="cpu"):
    """Product Subject: 1 mm voxels centred at the origin, convert_to_mu=False."""
    from nanodrr_amd import Subject

    image, label = random_volume(size, n_classes, seed)
    v2w = torch.eye(4)
    v2w[:3, 3] = -(size - 1) / 2
    w2v = torch.linalg.inv(v2w)
    s = Subject(image, label, v2w, w2v, Subject._make_voxel_to_grid(image.shape), torch.zeros(3),
                max_label=n_classes - 1, convert_to_mu=False)
    return s.to(device)


def ref_camera_np(height=16, width=16):
    """The reference test camera (test_fused.py:37-47) in float64 numpy."""
    sdd = 100.0
    k_inv = og.make_k_inv(sdd, 1.0, 1.0, 0.0, 0.0, height, width)
    rt_inv = og.make_rt_inv([[5.0, -3.0, 8.0]], [[1.0, 80.0, -2.0]], "AP", np.zeros(3))
    return k_inv, rt_inv, np.array([sdd]), height, width


def oracle_inputs(image, label, n_classes, k_inv, rt_inv, sdd, height, width, n_samples, orthographic=False,
                  size=None):
    """numpy float64 inputs of oracle.raymarch for a volume centred at the origin with 1 mm voxels."""
    vol = image.reshape(image.shape[-3:]).double().numpy()
    D, H, W = vol.shape
    w2v = np.eye(4)
    w2v[:3, 3] = [(W - 1) / 2, (H - 1) / 2, (D - 1) / 2]
    tgt = orm.make_tgt(k_inv, sdd, height, width)
    src = orm.make_src(orthographic, tgt, sdd)
    step = orm.step_size(src, tgt, n_samples)
    M = orm.camera_to_voxel(w2v, rt_inv)
    lab = orm.label_classes(label.reshape(label.shape[-3:]).numpy()) if n_classes > 1 else None
    return dict(vol=vol, lab=lab, n_classes=n_classes, M=M, src=src, tgt=tgt, step=step, n_samples=n_samples)


def rel_to_max(ref, out) -> float:
    ref = np.asarray(ref, np.float64)
    out = np.asarray(out, np.float64)
    return float(np.abs(ref - out).max() / np.abs(ref).max())
