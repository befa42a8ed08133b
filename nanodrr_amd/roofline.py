"""Algorithmic byte counts of the ray-march kernels (the roofline numerator).

A sample "touches" the volume when its position lies strictly inside
(-1, W) x (-1, H) x (-1, D) in voxel-index coordinates: then at least one of
its eight trilinear corners is a voxel. `inside_samples` counts those
samples per ray in closed form (a slab intersection per axis; float64, any
device), so the per-launch byte count needs no march.

Per touching sample the march reads the 8 corners (8 * e_v bytes, e_v the
volume element size) and, with a label map, one int8 label. Per ray it reads
the target (12 B), the step (4 B) and writes C outputs (4 C B); the backward
reads C upstream gradients (4 C B) and, with camera gradients, writes 28 B.
"""

from __future__ import annotations

import torch

_ELEM = {torch.float32: 4, torch.bfloat16: 2, torch.float16: 2}


def inside_samples(M: torch.Tensor, src: torch.Tensor, tgt: torch.Tensor, n_samples: int, shape) -> torch.Tensor:
    """(B, N) int64 count of samples s in [0, S) with M[b] @ [lerp(src, tgt, s/(S-1)), 1]
    strictly inside (-1, W) x (-1, H) x (-1, D)."""
    D, H, W = shape
    M = M.to(torch.float64)
    B = M.shape[0]
    src = src.to(torch.float64).expand(B, tgt.shape[1] if src.shape[1] > 1 else 1, 3)
    tgt = tgt.to(torch.float64).expand(B, -1, -1)
    p0 = torch.einsum("bij,bnj->bni", M[:, :, :3], src) + M[:, None, :, 3]
    d = torch.einsum("bij,bnj->bni", M[:, :, :3], tgt - src)
    p0 = p0.expand_as(d)
    lo = torch.full_like(d, -1.0)
    hi = torch.tensor([W, H, D], dtype=torch.float64, device=d.device).expand_as(d)
    nz = d != 0
    safe = torch.where(nz, d, torch.ones_like(d))
    a = (lo - p0) / safe
    b = (hi - p0) / safe
    t0 = torch.where(nz, torch.minimum(a, b), torch.where((p0 > lo) & (p0 < hi), -torch.inf, torch.inf))
    t1 = torch.where(nz, torch.maximum(a, b), torch.where((p0 > lo) & (p0 < hi), torch.inf, -torch.inf))
    t_lo = t0.amax(-1)
    t_hi = t1.amin(-1)
    s_m1 = float(n_samples - 1)
    # integer s strictly between t_lo * (S-1) and t_hi * (S-1), clipped to [0, S-1]
    first = torch.clamp(torch.floor(t_lo * s_m1) + 1, min=0.0)
    last = torch.clamp(torch.ceil(t_hi * s_m1) - 1, max=s_m1)
    return torch.clamp(last - first + 1, min=0.0).nan_to_num(0.0).to(torch.int64)


def march_bytes(n_inside: int, n_rays: int, n_poses: int, n_classes: int, vol_dtype: torch.dtype,
                backward: bool = False, need_gcam: bool = False, jac_planes: int = 0, need_gvol: bool = False) -> int:
    """Algorithmic HBM bytes of one forward (or two-pass backward) launch; a forward
    with jac_planes > 0 also writes the single-pass adjoint's per-ray sums. A backward
    with volume gradients adds, per touching sample, the fp32 read-modify-write of the
    eight corners' gradient entries (8 * 8 bytes)."""
    ev = _ELEM[vol_dtype]
    per_sample = 8 * ev + (1 if n_classes > 1 else 0) + (64 if backward and need_gvol else 0)
    per_ray = 12 + 4 + 4 * n_classes + (28 if (backward and need_gcam) else 0) + 4 * jac_planes * n_classes
    return n_inside * per_sample + n_rays * per_ray + 48 * n_poses


def contraction_bytes(n_rays: int, n_poses: int, n_classes: int, jac_planes: int = 6, need_gcam: bool = False) -> int:
    """Algorithmic HBM bytes of the single-pass backward: per ray the stored sums,
    the upstream gradient, tgt and step read (plus 28 bytes of camera gradients
    written), and M read per pose."""
    per_ray = 4 * jac_planes * n_classes + 4 * n_classes + 12 + 4 + (28 if need_gcam else 0)
    return n_rays * per_ray + 48 * n_poses
