"""Pure-PyTorch restatement of nanodrr's `render_torch` (grid_sample backend).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py). Restates
src/nanodrr/drr/backends.py:9-63 and the ray construction of
src/nanodrr/drr/renderer.py:71-132 with ATen's `F.grid_sample` (bilinear,
align_corners=False, zero padding) as the interpolation routine. Works in
any dtype: float64 gives an independent check of oracle.raymarch (a
different implementation of the same integral), float32 is the CPU
baseline that bench.py times. One deviation: labels outside [0, C) are
dropped (masked) instead of indexing out of range in `scatter_add_`.
"""

from __future__ import annotations

import torch
import torch.nn.functional as F


def make_tgt(k_inv, sdd, height, width):
    v, u = torch.meshgrid(
        torch.arange(height, dtype=k_inv.dtype) + 0.5, torch.arange(width, dtype=k_inv.dtype) + 0.5, indexing="ij"
    )
    uv1 = torch.stack([u, v, torch.ones_like(u)], dim=-1).reshape(height * width, 3)
    return sdd[:, None, None] * torch.einsum("bij,nj->bni", k_inv, uv1)


def make_src(orthographic, tgt, sdd):
    if orthographic:
        src = tgt.clone()
        src[..., 2] -= sdd[:, None]
        return src
    return torch.zeros(tgt.shape[0], 1, 3, dtype=tgt.dtype)


def voxel_to_grid(shape, dtype=torch.float64):
    """Voxel index -> [-1, 1] grid coordinate with align_corners=False (subject.py:218-242)."""
    D, H, W = shape
    size = torch.tensor([W, H, D], dtype=dtype)
    mat = torch.eye(4, dtype=dtype)
    mat[:3, :3] = torch.diag(2.0 / size)
    mat[:3, 3] = 1.0 / size - 1.0
    return mat


def _transform(xform, v):
    vh = torch.cat([v, torch.ones_like(v[..., :1])], dim=-1)
    out = torch.einsum("bij,bnj->bni", xform, vh)
    return out[..., :3] / out[..., 3:]


def march(image, label, n_classes, xform, src, tgt, step, n_samples):
    """(B, C, N) line integrals; `xform` (B, 4, 4) maps camera space to grid coordinates.

    image/label: (1, 1, D, H, W) tensors (label float, trunc'd to classes).
    """
    B = xform.shape[0]
    src = src.expand(B, -1, -1)
    tgt = tgt.expand(B, -1, -1)
    src_g = _transform(xform, src)
    tgt_g = _transform(xform, tgt)
    t = torch.linspace(0, 1, n_samples, dtype=src_g.dtype)
    pts = torch.lerp(src_g[:, None, :, None], tgt_g[:, None, :, None], t[None, :, None, None, None])
    img = F.grid_sample(image.expand(B, -1, -1, -1, -1), pts, mode="bilinear", align_corners=False)[:, 0, ..., 0]
    step = step.expand(B, -1)
    if n_classes == 1:
        return img.sum(dim=1, keepdim=True) * step[:, None, :]
    idx = F.grid_sample(label.expand(B, -1, -1, -1, -1).to(img.dtype), pts, mode="nearest",
                        align_corners=False)[:, 0, ..., 0].long()
    ok = (idx >= 0) & (idx < n_classes)
    out = torch.zeros(B, n_classes, img.shape[-1], dtype=img.dtype)
    out.scatter_add_(1, idx.clamp(0, n_classes - 1), torch.where(ok, img, torch.zeros_like(img)))
    return out * step[:, None, :]


def march_voxel_M(image, label, n_classes, M, src, tgt, step, n_samples):
    """Same as `march`, with the (B, 3, 4) camera -> voxel-index affine M as input
    (so autograd gives dL/dM, the quantity the fused adjoint produces)."""
    B = M.shape[0]
    bottom = torch.zeros(B, 1, 4, dtype=M.dtype)
    bottom[..., 3] = 1.0
    Mh = torch.cat([M, bottom], dim=1)
    xform = voxel_to_grid(image.shape[-3:], M.dtype) @ Mh
    return march(image, label, n_classes, xform, src, tgt, step, n_samples)


def render_torch(image, label, n_classes, world_to_grid, k_inv, rt_inv, sdd, height, width, n_samples,
                 orthographic=False):
    """Full restatement of render(..., backend="torch"): rays, step, march, reshape."""
    tgt = make_tgt(k_inv, sdd, height, width).to(rt_inv.dtype)
    src = make_src(orthographic, tgt, sdd).to(rt_inv.dtype)
    step = (tgt - src).norm(dim=-1) / float(n_samples - 1)
    xform = world_to_grid.to(rt_inv.dtype) @ rt_inv
    out = march(image.to(rt_inv.dtype), label, n_classes, xform, src, tgt, step, n_samples)
    return out.reshape(rt_inv.shape[0], n_classes, height, width)
