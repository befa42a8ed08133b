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


def coef_header(packed: torch.Tensor) -> tuple[float, float, int]:
    """(scale s, 1/s, record type code) from a coefficient-layout buffer's header."""
    hdr = packed[:16].cpu().view(torch.float32)
    return float(hdr[1]), float(hdr[2]), int(packed[:16].cpu().view(torch.int32)[3])


def coef_decode(packed: torch.Tensor, dtype, shape) -> np.ndarray:
    """(D+1, H+1, W+1, 8) float64 coefficients a coefficient-layout buffer holds, unscaled: fp32
    records as they are; bf16 storage holds c0 in bf16 and the rest in fp16; fp16 storage all in
    fp16; 16-bit slots hold the coefficient times s."""
    D, H, W = shape
    body = packed[256:].cpu()
    if dtype == torch.float32:
        return body.view(torch.float32).reshape(D + 1, H + 1, W + 1, 8).double().numpy()
    _, inv, _ = coef_header(packed)
    half = body.view(torch.float16).reshape(D + 1, H + 1, W + 1, 8).double()
    if dtype == torch.bfloat16:
        half[..., 0] = body.view(torch.bfloat16).reshape(D + 1, H + 1, W + 1, 8)[..., 0].double()
    return half.numpy() * inv


def coef_expected(vol: np.ndarray, dtype) -> tuple[torch.Tensor, float]:
    """(raw record slots as int16 (or fp32 bits for fp32 records), scale s) a coefficient pack of
    `vol` must produce: oracle.raymarch.coef_records in fp32, times s = the largest power of two with
    8 max|v| s <= 65504, each slot rounded to its type (c0 in bf16 for bf16 storage)."""
    coef = orm.coef_records(vol, np.float32)
    if dtype == torch.float32:
        return torch.from_numpy(coef).view(torch.int32), 1.0
    m = np.float32(np.abs(np.asarray(vol, np.float32)).max())
    s = 1.0 if m == 0 else float(2.0 ** np.floor(np.log2(np.float32(8188.0) / m)))
    scaled = torch.from_numpy(coef * np.float32(s))
    slots = scaled.to(torch.float16).view(torch.int16).clone()
    if dtype == torch.bfloat16:
        slots[..., 0] = scaled[..., 0].to(torch.bfloat16).view(torch.int16)
    return slots, s


LCOEF_TOP = np.float32(2.0 ** -7 - 2.0 ** -18)  # every labelled slot stays at or below this times 1


def lcoef_codes(lab: np.ndarray) -> np.ndarray:
    """(D+1, H+1, W+1, 8) 2-bit class codes of labelled coefficient records: corner j = x + 2y + 4z
    of cell (z0, y0, x0) in [-1, D-1] x ..., lab & 3 inside the volume, 0 outside."""
    lab = np.asarray(lab, np.int64)
    D, H, W = lab.shape
    P = np.zeros((D + 2, H + 2, W + 2), np.int64)
    P[1:-1, 1:-1, 1:-1] = lab & 3
    return np.stack([P[(j >> 2):(j >> 2) + D + 1, ((j >> 1) & 1):((j >> 1) & 1) + H + 1, (j & 1):(j & 1) + W + 1]
                     for j in range(8)], axis=-1)


def _as_int16(bits: torch.Tensor) -> torch.Tensor:
    """16-bit patterns held in int32 (0 .. 65535) as int16."""
    return ((bits + 32768) % 65536 - 32768).to(torch.int16)


def lcoef_scale(coef: np.ndarray) -> tuple[float, np.float32]:
    """(s, max|coefficient|) of labelled records: the largest power of two with max|coefficient| s
    <= 2^-7 - 2^-18, i.e. 2^floor(log2) of that fp32 quotient (clamped to [2^-126, 2^126]); 1 for an
    all-zero volume."""
    m = np.float32(np.abs(coef).max())
    if m == 0:
        return 1.0, m
    t = np.clip(np.float32(LCOEF_TOP / m), np.float32(2.0 ** -126), np.float32(2.0 ** 126))
    bits = np.array([t], np.float32).view(np.uint32) & np.uint32(0x7F800000)
    return float(bits.view(np.float32)[0]), m


def lcoef_expected(vol: np.ndarray, lab: np.ndarray, dtype) -> tuple[torch.Tensor, float, np.float32]:
    """(raw record slots as int16, scale s, max|coefficient|) a labelled coefficient pack of a not
    negative `vol` and class map `lab` must produce: oracle.raymarch.coef_records in fp32, times s
    (lcoef_scale), rounded to fp16 (c0 to bf16 for bf16 storage), with the corners' class codes
    (lcoef_codes) in the slots' top exponent bits: bits 13-14 of an fp16 slot, 14-15 of a bf16 c0."""
    coef = orm.coef_records(vol, np.float32)
    s, m = lcoef_scale(coef)
    scaled = torch.from_numpy(coef * np.float32(s))
    codes = torch.from_numpy(lcoef_codes(lab)).to(torch.int32)
    slots = (scaled.to(torch.float16).view(torch.int16).to(torch.int32) & 0x9FFF) | (codes << 13)
    if dtype == torch.bfloat16:
        c0 = scaled[..., 0].abs().to(torch.bfloat16).view(torch.int16).to(torch.int32) & 0xFFFF
        slots[..., 0] = c0 | (codes[..., 0] << 14)
    return _as_int16(slots), s, m


def quad_coef_records(vol: np.ndarray, dtype=np.float32) -> np.ndarray:
    """(D+1, H+1, W+2, 4) quad coefficient records of the zero-padded volume, one per quad record
    (z0, y0, x) in [-1, D-1] x [-1, H-1] x [-1, W]: [a, by, bz, byz] of the (y, z) quad at x, formed
    with the arithmetic of `dtype` in the order ndrr_pack_volume documents (by first)."""
    vol = np.asarray(vol, dtype)
    D, H, W = vol.shape
    P = np.zeros((D + 2, H + 2, W + 2), dtype)
    P[1:-1, 1:-1, 1:-1] = vol
    v = [P[a:a + D + 1, b:b + H + 1, :] for a in (0, 1) for b in (0, 1)]  # v[2a + b]
    by = v[1] - v[0]
    return np.stack([v[0], by, v[2] - v[0], (v[3] - v[2]) - by], axis=-1)


def qcoef_expected(vol: np.ndarray) -> tuple[torch.Tensor, float]:
    """(raw record slots as int16, scale s) of a quad coefficient pack of `vol`: quad_coef_records in
    fp32 times s (the coefficient layout's rule from max|v|), rounded to fp16."""
    rec = quad_coef_records(vol, np.float32)
    m = np.float32(np.abs(np.asarray(vol, np.float32)).max())
    s = 1.0 if m == 0 else float(2.0 ** np.floor(np.log2(np.float32(8188.0) / m)))
    return torch.from_numpy(rec * np.float32(s)).to(torch.float16).view(torch.int16), s


def qcoef_decode(packed: torch.Tensor, shape) -> np.ndarray:
    """(D+1, H+1, W+2, 4) float64 quad coefficients a quad coefficient buffer holds, unscaled."""
    D, H, W = shape
    _, inv, _ = coef_header(packed)
    return packed[256:].cpu().view(torch.float16).reshape(D + 1, H + 1, W + 2, 4).double().numpy() * inv


def qcoef_to_coef(q: np.ndarray) -> np.ndarray:
    """The coefficient records (oracle.raymarch.coef_records' layout, float64) of the interpolant
    lerp_x(Q(x0), Q(x0 + 1)) that quad coefficient records define: c0 = a0, cx = a1 - a0,
    cy = by0, cxy = by1 - by0, cz = bz0, cxz = bz1 - bz0, cyz = byz0, cxyz = byz1 - byz0."""
    q = np.asarray(q, np.float64)
    q0, q1 = q[:, :, :-1], q[:, :, 1:]
    a0, by0, bz0, byz0 = (q0[..., k] for k in range(4))
    a1, by1, bz1, byz1 = (q1[..., k] for k in range(4))
    return np.stack([a0, a1 - a0, by0, by1 - by0, bz0, bz1 - bz0, byz0, byz1 - byz0], axis=-1)


def lcoef_decode(packed: torch.Tensor, dtype, shape) -> tuple[np.ndarray, np.ndarray]:
    """((D+1, H+1, W+1, 8) float64 coefficients unscaled, (D+1, H+1, W+1, 8) class codes) that a
    labelled coefficient buffer holds: the code bits masked out of each slot before decoding."""
    D, H, W = shape
    raw = packed[256:].cpu().view(torch.int16).reshape(D + 1, H + 1, W + 1, 8).to(torch.int32) & 0xFFFF
    codes = (raw >> 13) & 3
    vals = raw & 0x9FFF
    if dtype == torch.bfloat16:
        codes[..., 0] = (raw[..., 0] >> 14) & 3
        vals[..., 0] = raw[..., 0] & 0x3FFF
    as16 = _as_int16(vals)
    half = as16.view(torch.float16).double()
    if dtype == torch.bfloat16:
        half[..., 0] = as16[..., 0].view(torch.bfloat16).double()
    _, inv, _ = coef_header(packed)
    return half.numpy() * inv, codes.numpy()
