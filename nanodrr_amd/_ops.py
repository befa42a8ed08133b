"""PyTorch custom ops over the native ray march (``libnanodrr_amd.so``).

`nanodrr_amd::fused_raymarch` / `nanodrr_amd::fused_raymarch_bwd` keep the
schemas of nanodrr's Triton ops (src/nanodrr/drr/_fused.py:415-426 and
:477-491) so `torch.compile` treats the launches as opaque, with these
differences:
  * `lab` is the int8 label map (`Subject.label_i8`), not the float one;
  * `src`, `tgt` and `step` may keep a batch dimension of 1 (shared by every
    pose) instead of being expanded and copied per pose;
  * the output is always fp32 (an fp32 accumulation rounded to a bf16 output
    alone costs ~2^-9 relative error);
  * an optional `packed` copy of `vol` in another layout (`pack_volume`) is
    what the kernels read when given, with `layout` naming it; `vol` still
    fixes the shape, dtype and the layout of the volume gradient;
  * the launch goes through the C ABI of include/nanodrr_amd.h.
"""

from __future__ import annotations

from typing import Optional

import torch

from . import _lib

_DTYPE_CODE = {torch.float32: _lib.NDRR_F32, torch.bfloat16: _lib.NDRR_BF16, torch.float16: _lib.NDRR_F16}


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None or t.numel() == 0 else t.data_ptr()


def _stream(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _volume_arg(vol, packed, layout):
    """(pointer, layout code) the kernels read: the packed copy when given, else `vol`."""
    if packed is None or packed.numel() == 0:
        return vol.data_ptr(), _lib.LAYOUT_LINEAR
    code = _lib.LAYOUTS.get(layout)
    if code is None or code == _lib.LAYOUT_LINEAR:
        raise ValueError(f"layout must name the packed copy ('quad' or 'stencil'), got {layout!r}")
    D, H, W = vol.shape
    need = _lib.lib().ndrr_packed_bytes(D, H, W, _DTYPE_CODE[vol.dtype], code)
    if packed.dtype != torch.uint8 or packed.numel() != need or not packed.is_contiguous() \
            or packed.device != vol.device:
        raise ValueError(f"packed must be a contiguous uint8 tensor of {need} bytes on {vol.device}")
    return packed.data_ptr(), code


@torch.library.custom_op("nanodrr_amd::pack_volume", mutates_args=())
def pack_volume(vol: torch.Tensor, layout: str) -> torch.Tensor:
    """`layout` copy (uint8 bytes) of a (D, H, W) device volume (ndrr_pack_volume)."""
    if not vol.is_cuda or vol.dim() != 3 or vol.dtype not in _DTYPE_CODE:
        raise ValueError("pack_volume needs a (D, H, W) float32/bfloat16/float16 device tensor")
    if layout not in _lib.LAYOUTS:
        raise ValueError(f"unknown layout {layout!r}; expected one of {sorted(_lib.LAYOUTS)}")
    vol = vol.detach().contiguous()
    D, H, W = vol.shape
    lib = _lib.lib()
    code, dt = _lib.LAYOUTS[layout], _DTYPE_CODE[vol.dtype]
    out = torch.empty(lib.ndrr_packed_bytes(D, H, W, dt, code), device=vol.device, dtype=torch.uint8)
    with torch.cuda.device(vol.device):
        rc = lib.ndrr_pack_volume(vol.data_ptr(), dt, code, D, H, W, out.data_ptr(), _stream(vol.device))
    _lib.check(rc, "ndrr_pack_volume")
    return out


@pack_volume.register_fake
def _(vol, layout):
    D, H, W = vol.shape
    e = 4 if vol.dtype == torch.float32 else 2
    n = {"linear": D * H * W, "stencil": (D + 1) * (H + 1) * (W + 1) * 8, "quad": (D + 1) * (H + 1) * (W + 2) * 4}
    return torch.empty(n[layout] * e, device=vol.device, dtype=torch.uint8)


@torch.library.custom_op("nanodrr_amd::pack_labels", mutates_args=())
def pack_labels(lab: torch.Tensor, layout: str) -> torch.Tensor:
    """Labels in the form `layout`'s march reads (ndrr_pack_labels): the label quads
    for "quad", a copy of the (D, H, W) int8 map otherwise."""
    if not lab.is_cuda or lab.dim() != 3 or lab.dtype != torch.int8:
        raise ValueError("pack_labels needs a (D, H, W) int8 device tensor (Subject.label_i8)")
    if layout not in _lib.LAYOUTS:
        raise ValueError(f"unknown layout {layout!r}; expected one of {sorted(_lib.LAYOUTS)}")
    lab = lab.contiguous()
    D, H, W = lab.shape
    lib = _lib.lib()
    code = _lib.LAYOUTS[layout]
    out = torch.empty(lib.ndrr_packed_label_bytes(D, H, W, code), device=lab.device, dtype=torch.int8)
    with torch.cuda.device(lab.device):
        rc = lib.ndrr_pack_labels(lab.data_ptr(), code, D, H, W, out.data_ptr(), _stream(lab.device))
    _lib.check(rc, "ndrr_pack_labels")
    return out if layout == "quad" else out.view(D, H, W)


@pack_labels.register_fake
def _(lab, layout):
    D, H, W = lab.shape
    if layout == "quad":
        return torch.empty((D + 1) * (H + 1) * (W + 2) * 4, device=lab.device, dtype=torch.int8)
    return torch.empty(D, H, W, device=lab.device, dtype=torch.int8)


def _geometry(vol, lab, M, src, tgt, step, n_samples, n_classes, width, layout="linear"):
    """Validate shapes/devices and return the launch geometry (raises before any launch)."""
    if not vol.is_cuda:
        raise ValueError("nanodrr_amd renders on a GPU: the volume must be a device tensor")
    if vol.dim() != 3 or not vol.is_contiguous():
        raise ValueError("vol must be a contiguous (D, H, W) tensor")
    if vol.dtype not in _DTYPE_CODE:
        raise ValueError(f"vol dtype {vol.dtype} unsupported; expected float32, bfloat16 or float16")
    if n_samples < 2:
        raise ValueError("n_samples must be at least 2")
    if not 1 <= n_classes <= _lib.MAX_CLASSES:
        raise ValueError(f"n_classes must lie in [1, {_lib.MAX_CLASSES}], got {n_classes}")
    if M.dim() != 3 or M.shape[1:] != (3, 4) or M.dtype != torch.float32:
        raise ValueError("M must be a float32 (B, 3, 4) tensor")
    B = M.shape[0]
    if tgt.dim() != 3 or tgt.shape[2] != 3 or tgt.shape[0] not in (1, B):
        raise ValueError("tgt must be (B or 1, N, 3)")
    N = tgt.shape[1]
    if src.dim() != 3 or src.shape[2] != 3 or src.shape[0] not in (1, B) or src.shape[1] not in (1, N):
        raise ValueError("src must be (B or 1, 1 or N, 3)")
    if step.dim() != 2 or step.shape[1] != N or step.shape[0] not in (1, B):
        raise ValueError("step must be (B or 1, N)")
    for name, t in (("M", M), ("src", src), ("tgt", tgt), ("step", step)):
        if t.dtype != torch.float32 or not t.is_contiguous() or t.device != vol.device:
            raise ValueError(f"{name} must be a contiguous float32 tensor on {vol.device}")
    if n_classes > 1:
        if lab.dtype != torch.int8 or not lab.is_contiguous() or lab.device != vol.device:
            raise ValueError("lab must be a contiguous int8 tensor on the volume's device")
        if layout == "quad":
            need = _lib.lib().ndrr_packed_label_bytes(*vol.shape, _lib.LAYOUT_QUAD)
            if lab.numel() != need:
                raise ValueError(f"the quad layout reads label quads (pack_labels): {need} bytes, got {lab.numel()}")
        elif lab.shape != vol.shape:
            raise ValueError("lab must be the (D, H, W) int8 label map for this layout")
    if width < 0 or (width > 0 and N % width != 0):
        raise ValueError(f"width {width} must be 0 or divide the ray count {N}")
    D, H, W = vol.shape
    strides = (
        0 if src.shape[0] == 1 else src.shape[1] * 3,  # src batch stride
        0 if src.shape[1] == 1 else 3,                 # src ray stride
        0 if tgt.shape[0] == 1 else N * 3,             # tgt batch stride
        0 if step.shape[0] == 1 else N,                # step batch stride
    )
    return B, N, D, H, W, strides


@torch.library.custom_op("nanodrr_amd::fused_raymarch", mutates_args=())
def fused_raymarch(
    vol: torch.Tensor,
    lab: torch.Tensor,
    M: torch.Tensor,
    src: torch.Tensor,
    tgt: torch.Tensor,
    step: torch.Tensor,
    n_samples: int,
    n_classes: int,
    width: int,
    packed: Optional[torch.Tensor] = None,
    layout: str = "linear",
) -> torch.Tensor:
    """Line integrals (B, C, N) fp32 of the rays through `vol` (see ndrr_raymarch_fwd)."""
    B, N, D, H, W, strides = _geometry(vol, lab, M, src, tgt, step, n_samples, n_classes, width,
                                       layout if packed is not None else "linear")
    vptr, code = _volume_arg(vol, packed, layout)
    out = torch.empty(B, n_classes, N, device=vol.device, dtype=torch.float32)
    with torch.cuda.device(vol.device):
        rc = _lib.lib().ndrr_raymarch_fwd(
            vptr, _DTYPE_CODE[vol.dtype], code, _ptr(lab) if n_classes > 1 else None,
            M.data_ptr(), src.data_ptr(), tgt.data_ptr(), step.data_ptr(), out.data_ptr(),
            B, N, D, H, W, n_samples, n_classes, width, *strides, _stream(vol.device),
        )
    _lib.check(rc, "ndrr_raymarch_fwd")
    return out


@fused_raymarch.register_fake
def _(vol, lab, M, src, tgt, step, n_samples, n_classes, width, packed=None, layout="linear"):
    return torch.empty(M.shape[0], n_classes, tgt.shape[1], device=vol.device, dtype=torch.float32)


@torch.library.custom_op("nanodrr_amd::fused_raymarch_bwd", mutates_args=())
def fused_raymarch_bwd(
    go: torch.Tensor,
    vol: torch.Tensor,
    lab: torch.Tensor,
    M: torch.Tensor,
    src: torch.Tensor,
    tgt: torch.Tensor,
    step: torch.Tensor,
    n_samples: int,
    n_classes: int,
    width: int,
    need_gvol: bool,
    need_gcam: bool,
    packed: Optional[torch.Tensor] = None,
    layout: str = "linear",
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """Analytic adjoint (see ndrr_raymarch_bwd). Returns (gvol|empty, gM (B,3,4),
    gsrc (B,N,3)|empty, gtgt (B,N,3)|empty, gstep (B,N)|empty), all fp32."""
    if need_gvol and torch.are_deterministic_algorithms_enabled():
        raise RuntimeError(
            "Volume gradients on the hip backend use nondeterministic atomics; "
            "disable deterministic algorithms to request them"
        )
    B, N, D, H, W, strides = _geometry(vol, lab, M, src, tgt, step, n_samples, n_classes, width,
                                       layout if packed is not None else "linear")
    vptr, code = _volume_arg(vol, packed, layout)
    go = go.to(torch.float32).contiguous()
    if go.shape != (B, n_classes, N):
        raise ValueError(f"go must be (B, C, N) = {(B, n_classes, N)}, got {tuple(go.shape)}")
    dev = vol.device
    lib = _lib.lib()
    ws = torch.empty(lib.ndrr_bwd_workspace_floats(B, N, width), device=dev, dtype=torch.float32)
    gM = torch.empty(B, 3, 4, device=dev, dtype=torch.float32)
    # unused outputs are distinct empty tensors: custom-op outputs may not alias
    gvol = torch.zeros(vol.shape, device=dev, dtype=torch.float32) if need_gvol else gM.new_empty(0)
    gsrc = torch.empty(B, N, 3, device=dev, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    gtgt = torch.empty(B, N, 3, device=dev, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    gstep = torch.empty(B, N, device=dev, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    with torch.cuda.device(dev):
        rc = lib.ndrr_raymarch_bwd(
            go.data_ptr(), vptr, _DTYPE_CODE[vol.dtype], code, _ptr(lab) if n_classes > 1 else None,
            M.data_ptr(), src.data_ptr(), tgt.data_ptr(), step.data_ptr(),
            ws.data_ptr(), gM.data_ptr(), _ptr(gsrc), _ptr(gtgt), _ptr(gstep), _ptr(gvol),
            int(need_gcam), B, N, D, H, W, n_samples, n_classes, width, *strides, _stream(dev),
        )
    _lib.check(rc, "ndrr_raymarch_bwd")
    return gvol, gM, gsrc, gtgt, gstep


@fused_raymarch_bwd.register_fake
def _(go, vol, lab, M, src, tgt, step, n_samples, n_classes, width, need_gvol, need_gcam, packed=None,
      layout="linear"):
    B, N = M.shape[0], tgt.shape[1]
    gM = torch.empty(B, 3, 4, device=M.device, dtype=torch.float32)
    gvol = torch.empty(vol.shape, device=M.device, dtype=torch.float32) if need_gvol else gM.new_empty(0)
    gsrc = torch.empty(B, N, 3, device=M.device, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    gtgt = torch.empty(B, N, 3, device=M.device, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    gstep = torch.empty(B, N, device=M.device, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    return gvol, gM, gsrc, gtgt, gstep


def _setup_context(ctx, inputs, output):
    vol, lab, M, src, tgt, step, n_samples, n_classes, width, packed, layout = inputs
    ctx.save_for_backward(vol, lab, M, src, tgt, step, packed)
    ctx.n_samples, ctx.n_classes, ctx.width, ctx.layout = n_samples, n_classes, width, layout


def _reduce_to(g: torch.Tensor, like: torch.Tensor) -> torch.Tensor:
    """Sum a per-ray (B, N, ...) gradient over the dimensions `like` broadcasts."""
    dims = [i for i in range(2) if like.shape[i] == 1 and g.shape[i] != 1]
    if dims:
        g = g.sum(dim=dims, keepdim=True)
    return g.to(like.dtype)


def _backward(ctx, go):
    vol, lab, M, src, tgt, step, packed = ctx.saved_tensors
    need_gvol = ctx.needs_input_grad[0]
    need_gcam = any(ctx.needs_input_grad[i] for i in (3, 4, 5))
    gvol, gM, gsrc, gtgt, gstep = fused_raymarch_bwd(
        go, vol, lab, M, src, tgt, step, ctx.n_samples, ctx.n_classes, ctx.width, need_gvol, need_gcam, packed,
        ctx.layout,
    )
    return (
        gvol.to(vol.dtype) if need_gvol else None,
        None,  # lab: nearest-label routing is not differentiable
        gM.to(M.dtype),
        _reduce_to(gsrc, src) if need_gcam else None,
        _reduce_to(gtgt, tgt) if need_gcam else None,
        _reduce_to(gstep, step) if need_gcam else None,
        None,  # n_samples
        None,  # n_classes
        None,  # width
        None,  # packed: a re-layout of vol; its gradient is vol's
        None,  # layout
    )


torch.library.register_autograd("nanodrr_amd::fused_raymarch", _backward, setup_context=_setup_context)


@torch.library.custom_op("nanodrr_amd::labels_to_i8", mutates_args=())
def labels_to_i8(label: torch.Tensor, n_classes: int) -> torch.Tensor:
    """int8 label map for the kernels: trunc(label) if in [0, n_classes) else -1."""
    if not label.is_cuda:
        raise ValueError("labels_to_i8 runs on the GPU: pass a device tensor")
    src = label.detach().to(torch.float32).contiguous()
    dst = torch.empty(src.shape, device=src.device, dtype=torch.int8)
    with torch.cuda.device(src.device):
        rc = _lib.lib().ndrr_labels_to_i8(src.data_ptr(), dst.data_ptr(), src.numel(), n_classes,
                                          _stream(src.device))
    _lib.check(rc, "ndrr_labels_to_i8")
    return dst


@labels_to_i8.register_fake
def _(label, n_classes):
    return torch.empty(label.shape, device=label.device, dtype=torch.int8)
