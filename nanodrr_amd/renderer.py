"""`render(...)`: differentiable DRRs through the native HIP ray march.

Same call signature, ray construction and errors as nanodrr's renderer
(src/nanodrr/drr/renderer.py:21-132) and fused backend
(src/nanodrr/drr/backends.py:66-117). Backends: "hip" (the native kernels)
and "auto" (= "hip"); "triton" is accepted as an alias of "hip" so callers of
the reference's fused backend switch unchanged. The reference's pure-PyTorch
`grid_sample` backend is not part of this package: "torch" raises.
"""

from __future__ import annotations

import torch

from . import _lib
from ._memo import Derived
from ._ops import eager
from .subject import _Same

_FUSED_DTYPES = (torch.float32, torch.float16, torch.bfloat16)
_BACKENDS = ("auto", "hip", "triton")

# Volume layout the kernels read: "stencil" (a cached copy with all eight
# corners of each cell: one 16-byte load per sample for 16-bit volumes),
# "quad" (a cached copy of each voxel's (y, z) corner quads: half the
# stencil's memory, which lets a 256^3 16-bit volume stay in the Infinity
# Cache), "coef" (a cached copy of each cell's interpolant coefficients, made
# from the unrounded volume: the gradient's edge differences rounded relative
# to their own size), "lcoef" (coefficient records that also carry the classes
# of their cell's corners, so the march loads no label map: 16-bit storage,
# 2 to 4 classes, a volume that is not negative), "qcoef" (the quad grid's
# records holding each (y, z) quad's bilinear coefficients in fp16, made from
# the unrounded volume: quad's footprint, fewer operations per sample than
# quad and more accurate than bf16 quad; 16-bit storage), "linear" (the (D, H, W)
# volume itself) or "auto": linear when the volume requires grad (a packed
# copy would be re-made after every update), else for 16-bit storage lcoef
# where it holds the subject (11-13% faster than coef at 4 classes, 1.7x at a
# 512x512x400 volume with 2 classes, where the label map's reads cost most) and coef
# otherwise (as fast as quad, with ~10x smaller gradient errors), and quad
# for fp32 (the fastest measured).
VOLUME_LAYOUT = "auto"
_LAYOUTS = ("linear", "quad", "stencil", "coef", "lcoef", "qcoef")

# How pose and camera gradients are formed when the volume needs none:
# "single_pass" (the forward march also stores per-ray sums and the backward
# contracts them with the upstream gradient, without marching the volume
# again) or "two_pass" (the backward marches the volume again, as nanodrr's
# kernel does). Volume gradients always take the two-pass adjoint.
ADJOINT = "single_pass"
_ADJOINTS = ("single_pass", "two_pass")


def fused_supported(subject, B: int, n_pixels: int) -> bool:
    """Hard limits of the kernels: one volume, 32-bit indexing, <= 8 classes (backends.py:66-69)."""
    return _supported(subject.image, subject.n_classes, B, n_pixels)


def _supported(image, C: int, B: int, n_pixels: int) -> bool:
    N = B * n_pixels
    return image.shape[0] == 1 and max(image.numel(), C * N, 3 * N) < 2**31 and 1 <= C <= _lib.MAX_CLASSES \
        and B <= 65535


_NO_LABELS: dict = {}  # an empty int8 tensor per device: the label argument of single-class marches


def _no_labels(device) -> torch.Tensor:
    t = _NO_LABELS.get(device)
    if t is None:
        t = _NO_LABELS[device] = torch.empty(0, dtype=torch.int8, device=device)
    return t


def _make_tgt(k_inv, sdd, height, width, device, dtype) -> torch.Tensor:
    """tgt[b, n] = sdd_b * K^-1_b [u + 1/2, v + 1/2, 1], n = v * width + u (renderer.py:101-117)."""
    v, u = torch.meshgrid(
        torch.arange(height, device=device, dtype=dtype) + 0.5,
        torch.arange(width, device=device, dtype=dtype) + 0.5,
        indexing="ij",
    )
    uv1 = torch.stack([u, v, torch.ones_like(u)], dim=-1).reshape(height * width, 3)
    return sdd[:, None, None] * torch.einsum("bij,nj->bni", k_inv, uv1)


def _make_src(orthographic: bool, tgt: torch.Tensor, sdd: torch.Tensor) -> torch.Tensor:
    """Perspective: one source at the camera origin per pose, (B, 1, 3).
    Orthographic: one source per ray, sdd behind its target (renderer.py:120-132)."""
    if orthographic:
        src = tgt.clone()
        src[..., 2] -= sdd[:, None]
        return src
    return torch.zeros(tgt.shape[0], 1, 3, device=tgt.device, dtype=tgt.dtype)


def render(
    subject,
    k_inv: torch.Tensor,
    rt_inv: torch.Tensor,
    sdd: torch.Tensor,
    height: int,
    width: int,
    n_samples: int = 500,
    orthographic: bool = False,
    src: torch.Tensor | None = None,
    tgt: torch.Tensor | None = None,
    backend: str = "auto",
) -> torch.Tensor:
    """Differentiable ray marching through `subject.image` (and its label map).

    Returns (B, C, H, W) fp32 line integrals, one channel per label class.
    Gradients flow to `rt_inv`, `k_inv`, `sdd`, `src`, `tgt` and, when it
    requires grad, the volume.
    """
    check_render_args(n_samples, backend)
    device, dtype = rt_inv.device, rt_inv.dtype
    if tgt is None and src is None and _native_rays(k_inv, sdd, dtype):
        tgt, src, step_size = _rays(k_inv, sdd, height, width, orthographic, n_samples)
    else:
        if tgt is None:
            tgt = _make_tgt(k_inv, sdd, height, width, device, dtype)
        if src is None:
            src = _make_src(orthographic, tgt, sdd)
        step_size = (tgt - src).norm(dim=-1) / float(n_samples - 1)
    return render_hip(subject, rt_inv, src, tgt, step_size, n_samples, height, width)


def check_render_args(n_samples: int, backend: str) -> None:
    """The argument errors of render (renderer.py:71-72,98), raised before anything runs."""
    if n_samples < 2:
        raise ValueError("n_samples must be at least 2")
    if backend not in _BACKENDS:
        if backend == "torch":
            raise ValueError(
                "backend='torch' (the grid_sample path) is not part of nanodrr_amd; "
                "use backend='hip' (or 'auto')"
            )
        raise ValueError(f"Unknown backend {backend!r}; expected 'auto', 'hip' or 'triton'")


_RAYS = Derived()  # the rays of the last detector rendered eagerly


def _rays(k_inv, sdd, height, width, orthographic, n_samples):
    """eager.make_rays, kept while k_inv and sdd are the same tensors at the same versions and the
    call is on the same stream (a fixed detector renders many poses: one launch and its host work
    fewer per render). Made afresh while a graph is captured or torch.compile traces (tensors made
    during a capture belong to the graph's memory)."""
    if torch.compiler.is_compiling() or torch.cuda.is_current_stream_capturing():
        return eager.make_rays(k_inv, sdd, height, width, orthographic, n_samples)
    key = (_Same(sdd), sdd._version, height, width, orthographic, n_samples,
           torch.cuda.current_stream(k_inv.device).cuda_stream)
    return _RAYS.get("rays", k_inv, key, lambda: eager.make_rays(k_inv, sdd, height, width, orthographic, n_samples))


def _native_rays(k_inv, sdd, dtype) -> bool:
    """Rays from one native launch: fp32 device intrinsics that need no gradient."""
    return (dtype == torch.float32 and k_inv.dtype == torch.float32 and sdd.dtype == torch.float32
            and k_inv.is_cuda and sdd.is_cuda and k_inv.dim() == 3 and sdd.dim() == 1
            and not (torch.is_grad_enabled() and (k_inv.requires_grad or sdd.requires_grad)))


def render_hip(subject, rt_inv, src, tgt, step_size, n_samples, height, width) -> torch.Tensor:
    """Native backend (the counterpart of render_fused, backends.py:72-117)."""
    image = subject.image
    if not (rt_inv.is_cuda and image.is_cuda):
        raise ValueError("the hip backend renders on a GPU: move the subject and poses to a cuda device")
    if ADJOINT not in _ADJOINTS:
        raise ValueError(f"ADJOINT must be one of {_ADJOINTS}, got {ADJOINT!r}")
    # Camera -> voxel-index affine, composed in float64 and rounded once: by the
    # march kernels themselves when no volume gradient is asked for and pose
    # gradients (if any) take the single-pass adjoint, else by one launch ahead.
    if rt_inv.dtype == torch.float32 and rt_inv.dim() == 3 and rt_inv.shape[1:] == (4, 4):
        if not image.requires_grad and (ADJOINT == "single_pass" or not torch.is_grad_enabled()):
            return march_pose(subject, rt_inv, src, tgt, step_size, n_samples, height, width)
        M = eager.camera_to_voxel(subject.world_to_voxel_f64(), rt_inv.contiguous())
    else:
        M = (subject.world_to_voxel.to(torch.float64) @ rt_inv.to(torch.float64))[:, :3, :].to(torch.float32)
    return march(subject, M, src, tgt, step_size, n_samples, height, width)


def volume_layout(image: torch.Tensor, layout: str | None = None) -> str:
    """The layout the kernels read for this volume: `layout` when given, else VOLUME_LAYOUT."""
    layout = VOLUME_LAYOUT if layout is None else layout
    if layout == "auto":
        if image.requires_grad or not packed_index_fits(*image.shape[-3:]):
            layout = "linear"
        elif image.dtype in (torch.bfloat16, torch.float16) and coef_fits(*image.shape[-3:]):
            layout = "coef"
        else:
            layout = "quad"
    if layout not in _LAYOUTS:
        raise ValueError(f"the volume layout must be 'auto' or one of {_LAYOUTS}, got {layout!r}")
    return layout


def coef_fits(D: int, H: int, W: int) -> bool:
    """Whether 16-bit coefficient records of a (D, H, W) volume fit the kernels' 32-bit buffer
    offsets (a 256-byte header and 16 bytes per cell, under 2 GiB)."""
    return 256 + (D + 1) * (H + 1) * (W + 1) * 16 < 2**31


def packed_index_fits(D: int, H: int, W: int) -> bool:
    """Whether the packed layouts can address this (D, H, W) volume: their record index is formed
    with 24-bit multiplies, so (D + 1)(H + 1) and W + 2 stay below 2^24, and the padded label
    map's with signed ones, so (H + 4)(W + 4) stays below 2^23."""
    return (D + 1) * (H + 1) < 2**24 and W + 2 < 2**24 and (H + 4) * (W + 4) < 2**23


def march_inputs(subject, M, src, tgt, step_size, n_samples, height, width, layout=None):
    """Checked kernel arguments of `march`: ((vol, lab, M, src, tgt, step, n_samples, C, width),
    packed volume or None, layout name, B, C). M is the (B|1, 3, 4) affine, or for `march_pose`
    the (B|1, 4, 4) rt_inv; either is expanded to B poses."""
    image = subject.image
    if not image.is_cuda:
        raise ValueError("the hip backend renders on a GPU: move the subject and poses to a cuda device")
    if image.dtype not in _FUSED_DTYPES:
        raise ValueError(f"volume dtype {image.dtype} unsupported; expected float32, float16 or bfloat16")
    C = subject.n_classes
    vol = image[0, 0]
    if not vol.is_contiguous():
        vol = vol.contiguous()
    M = M.to(torch.float32)
    B = max(M.shape[0], tgt.shape[0], src.shape[0])
    if M.shape[0] != B:
        M = M.expand(B, -1, -1)
    if not _supported(image, C, B, height * width):
        raise ValueError("inputs exceed the hip kernel's limits (one volume, 32-bit indexing, <= 8 classes)")
    auto = (VOLUME_LAYOUT if layout is None else layout) == "auto"
    layout = volume_layout(image, layout)
    # (traced graphs keep coef: whether labelled records fit is known only by packing them)
    if auto and layout == "coef" and 2 <= C <= 4 and not torch.compiler.is_compiling() \
            and subject.labelled_coef_fits():
        layout = "lcoef"
    packed = None if layout == "linear" else subject.packed_volume(layout)
    lab = subject.packed_labels(layout) if C > 1 else _no_labels(vol.device)
    args = (vol, lab, M.contiguous(), src.to(torch.float32).contiguous(), tgt.to(torch.float32).contiguous(),
            step_size.to(torch.float32).contiguous(), n_samples, C, width)
    return args, packed, layout, B, C


def march(subject, M, src, tgt, step_size, n_samples, height, width, offset=None, layout=None) -> torch.Tensor:
    """(B, C, H, W) line integrals of the rays src -> tgt (camera space) with the
    camera -> voxel-index affines M (B|1, 3, 4) through the subject's volume.
    `offset` (B,), in units of the sample step, slides each pose's sample lattice
    along its rays (sample s at t = (s + offset) / (n_samples - 1); the step is
    unchanged); the kernels take it directly unless the two-pass adjoint runs. `layout` overrides
    VOLUME_LAYOUT for this call."""
    args, packed, layout, B, C = march_inputs(subject, M, src, tgt, step_size, n_samples, height, width, layout)
    if ADJOINT not in _ADJOINTS:
        raise ValueError(f"ADJOINT must be one of {_ADJOINTS}, got {ADJOINT!r}")
    gcam = any(t.requires_grad for t in args[3:6])
    grad = torch.is_grad_enabled() and (subject.image.requires_grad or gcam or args[2].requires_grad)
    lattice = None
    if offset is not None:
        lattice = offset.detach().to(device=args[4].device, dtype=torch.float32).reshape(-1).expand(B).contiguous()
    if grad and ADJOINT == "single_pass" and not subject.image.requires_grad:
        out, _ = eager.fused_raymarch_jac(*args, 7 if gcam else 6, packed, layout, lattice)
    elif not grad and lattice is not None:
        out, _ = eager.fused_raymarch_jac(*args, 0, packed, layout, lattice)
    else:
        if lattice is not None:  # the two-pass adjoint: shifted copies of the rays
            vol, lab, M_, s, t, st = args[:6]
            d = (t - s) / float(n_samples - 1)
            u = lattice.reshape(-1, 1, 1)
            args = (vol, lab, M_, torch.addcmul(s, d, u), torch.addcmul(t, d, u), st) + args[6:]
        out = eager.fused_raymarch(*args, packed, layout)
    return out.reshape(B, C, height, width)


def march_pose(subject, rt_inv, src, tgt, step_size, n_samples, height, width) -> torch.Tensor:
    """`march` given the pose rt_inv (B|1, 4, 4) fp32: the kernels compose the camera -> voxel-index
    affine from it and the subject's float64 world_to_voxel, and the backward (the single-pass
    adjoint) gives the rt_inv gradient directly. No volume gradient."""
    args, packed, layout, B, C = march_inputs(subject, rt_inv, src, tgt, step_size, n_samples, height, width)
    vol, lab, rt, s, t, st, _, _, _ = args
    gcam = any(x.requires_grad for x in (s, t, st))
    planes = (7 if gcam else 6) if torch.is_grad_enabled() and (gcam or rt.requires_grad) else 0
    out, _ = eager.fused_raymarch_pose(vol, lab, subject.world_to_voxel_f64(), rt, s, t, st, n_samples, C, width,
                                       planes, packed, layout)
    return out.reshape(B, C, height, width)
