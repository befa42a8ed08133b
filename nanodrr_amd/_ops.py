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

`nanodrr_amd::fused_raymarch_jac` / `nanodrr_amd::fused_raymarch_bwd_jac` are
the single-pass adjoint of the same pair (no counterpart in nanodrr): the
forward also stores per-ray sums from which the backward forms pose and camera
gradients without marching the volume again.
"""

from __future__ import annotations

import contextlib
import functools
from types import SimpleNamespace
from typing import Optional

import torch

from . import _lib

_DTYPE_CODE = {torch.float32: _lib.NDRR_F32, torch.bfloat16: _lib.NDRR_BF16, torch.float16: _lib.NDRR_F16}


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None or t.numel() == 0 else t.data_ptr()


def _index(device: torch.device) -> int:
    return device.index if device.index is not None else torch.cuda.current_device()


def _stream(device: torch.device) -> int:
    """Raw handle of torch's current stream on `device` (the capturing stream under graph capture),
    read without building a Stream object."""
    return torch._C._cuda_getCurrentRawStream(_index(device))


def _on(device: torch.device):
    """`torch.cuda.device(device)`, or nothing when `device` is already the current one."""
    idx = _index(device)
    return contextlib.nullcontext() if idx == torch._C._cuda_getDevice() else torch.cuda.device(idx)


@functools.lru_cache(maxsize=64)
def _packed_nbytes(D: int, H: int, W: int, dtype_code: int, layout_code: int) -> int:
    return _lib.lib().ndrr_packed_bytes(D, H, W, dtype_code, layout_code)


def _volume_arg(vol, packed, layout):
    """(pointer, layout code) the kernels read: the packed copy when given, else `vol`."""
    if packed is None or packed.numel() == 0:
        return vol.data_ptr(), _lib.LAYOUT_LINEAR
    code = _lib.LAYOUTS.get(layout)
    if code is None or code == _lib.LAYOUT_LINEAR:
        raise ValueError(f"layout must name the packed copy ('quad', 'stencil', 'coef', 'lcoef' or 'qcoef'), "
                         f"got {layout!r}")
    need = _packed_nbytes(*vol.shape, _DTYPE_CODE[vol.dtype], code)
    if packed.dtype != torch.uint8 or packed.numel() != need or not packed.is_contiguous() \
            or packed.device != vol.device:
        raise ValueError(f"packed must be a contiguous uint8 tensor of {need} bytes on {vol.device}")
    return packed.data_ptr(), code


@torch.library.custom_op("nanodrr_amd::pack_volume", mutates_args=())
def pack_volume(vol: torch.Tensor, layout: str) -> torch.Tensor:
    """`layout` copy (uint8 bytes) of a (D, H, W) device volume (ndrr_pack_volume). "qcoef" records are
    fp16 whatever `vol`'s type (pass the unrounded values of a 16-bit volume)."""
    if not vol.is_cuda or vol.dim() != 3 or vol.dtype not in _DTYPE_CODE:
        raise ValueError("pack_volume needs a (D, H, W) float32/bfloat16/float16 device tensor")
    if layout not in _lib.LAYOUTS:
        raise ValueError(f"unknown layout {layout!r}; expected one of {sorted(_lib.LAYOUTS)}")
    if layout == "lcoef":
        raise ValueError("labelled coefficient records are made from the volume and its labels (pack_coef_labels)")
    vol = vol.detach().contiguous()
    D, H, W = vol.shape
    lib = _lib.lib()
    code, dt = _lib.LAYOUTS[layout], _DTYPE_CODE[vol.dtype]
    out = torch.empty(lib.ndrr_packed_bytes(D, H, W, dt, code), device=vol.device, dtype=torch.uint8)
    with _on(vol.device):
        rc = lib.ndrr_pack_volume(vol.data_ptr(), dt, code, D, H, W, out.data_ptr(), _stream(vol.device))
    _lib.check(rc, "ndrr_pack_volume")
    return out


COEF_HEADER_BYTES = 256  # the coefficient layouts' header: max |v| (|coefficient|), scale s, 1/s, record type, ...


def _packed_bytes(D, H, W, layout, dtype):
    e = 4 if dtype == torch.float32 else 2
    n = {"linear": D * H * W, "stencil": (D + 1) * (H + 1) * (W + 1) * 8, "quad": (D + 1) * (H + 1) * (W + 2) * 4,
         "coef": (D + 1) * (H + 1) * (W + 1) * 8, "lcoef": (D + 1) * (H + 1) * (W + 1) * 8,
         "qcoef": (D + 1) * (H + 1) * (W + 2) * 4}[layout]
    if layout == "qcoef":
        e = 2  # fp16 records whatever the source
    return n * e + (COEF_HEADER_BYTES if layout in ("coef", "lcoef", "qcoef") else 0)


@pack_volume.register_fake
def _(vol, layout):
    return torch.empty(_packed_bytes(*vol.shape, layout, vol.dtype), device=vol.device, dtype=torch.uint8)


@torch.library.custom_op("nanodrr_amd::pack_coef", mutates_args=())
def pack_coef(vol: torch.Tensor, record_dtype: torch.dtype) -> torch.Tensor:
    """Coefficient records (the "coef" layout, uint8 bytes, after a 256-byte header) for `record_dtype`
    storage from a (D, H, W) device volume of any supported type (ndrr_pack_coef): each coefficient
    formed in fp32 and rounded once; 16-bit records scaled by a power of two, bf16 ones holding the
    value in bf16 and the seven edge differences in fp16."""
    if not vol.is_cuda or vol.dim() != 3 or vol.dtype not in _DTYPE_CODE or record_dtype not in _DTYPE_CODE:
        raise ValueError("pack_coef needs a (D, H, W) float32/bfloat16/float16 device tensor and such a record type")
    vol = vol.detach().contiguous()
    D, H, W = vol.shape
    lib = _lib.lib()
    rd = _DTYPE_CODE[record_dtype]
    out = torch.empty(lib.ndrr_packed_bytes(D, H, W, rd, _lib.LAYOUT_COEF), device=vol.device, dtype=torch.uint8)
    with _on(vol.device):
        rc = lib.ndrr_pack_coef(vol.data_ptr(), _DTYPE_CODE[vol.dtype], rd, D, H, W, out.data_ptr(),
                                _stream(vol.device))
    _lib.check(rc, "ndrr_pack_coef")
    return out


@pack_coef.register_fake
def _(vol, record_dtype):
    return torch.empty(_packed_bytes(*vol.shape, "coef", record_dtype), device=vol.device, dtype=torch.uint8)


LCOEF_ERRORS = {_lib.LCOEF_NEGATIVE: "the volume has negative values",
                _lib.LCOEF_CLASS: "a class lies outside [-1, n_classes), or a voxel is dropped (-1) with 4 classes",
                _lib.LCOEF_RANGE: "the largest coefficient is not finite or too large for the records' scale"}


def lcoef_status(packed: torch.Tensor) -> int:
    """Header word 4 of labelled coefficient records: the LCOEF_* bits of what their volume and labels
    break (0: valid). Reads the device header (synchronises with its stream)."""
    return int(packed[16:20].view(torch.int32).item())


@torch.library.custom_op("nanodrr_amd::pack_coef_labels", mutates_args=())
def pack_coef_labels(vol: torch.Tensor, lab: torch.Tensor, n_classes: int, record_dtype: torch.dtype) -> torch.Tensor:
    """Labelled coefficient records (the "lcoef" layout, uint8 bytes, ndrr_pack_coef_labels) of a not
    negative (D, H, W) device volume and its int8 class map (`Subject.label_i8`, n_classes <= 4) for
    bf16 or fp16 storage: the "coef" records at a scale that frees two bits per slot, which hold the
    classes of the cell's eight corners, so the march reads no label map. Raises ValueError when the
    volume or labels do not fit the format (the check reads the records' header back)."""
    if not vol.is_cuda or vol.dim() != 3 or vol.dtype not in _DTYPE_CODE:
        raise ValueError("pack_coef_labels needs a (D, H, W) float32/bfloat16/float16 device tensor")
    if record_dtype not in (torch.bfloat16, torch.float16):
        raise ValueError(f"labelled coefficient records are bfloat16 or float16, not {record_dtype}")
    if not 1 <= n_classes <= 4:
        raise ValueError(f"labelled coefficient records hold at most 4 classes, got {n_classes}")
    if lab.dtype != torch.int8 or lab.shape != vol.shape or lab.device != vol.device:
        raise ValueError("lab must be the (D, H, W) int8 class map on the volume's device")
    vol = vol.detach().contiguous()
    lab = lab.contiguous()
    D, H, W = vol.shape
    lib = _lib.lib()
    rd = _DTYPE_CODE[record_dtype]
    out = torch.empty(lib.ndrr_packed_bytes(D, H, W, rd, _lib.LAYOUT_LCOEF), device=vol.device, dtype=torch.uint8)
    with _on(vol.device):
        rc = lib.ndrr_pack_coef_labels(vol.data_ptr(), _DTYPE_CODE[vol.dtype], lab.data_ptr(), n_classes, rd, D, H, W,
                                       out.data_ptr(), _stream(vol.device))
    _lib.check(rc, "ndrr_pack_coef_labels")
    status = lcoef_status(out)
    if status:
        raise ValueError("labelled coefficient records cannot hold this volume: "
                         + "; ".join(m for bit, m in LCOEF_ERRORS.items() if status & bit))
    return out


@pack_coef_labels.register_fake
def _(vol, lab, n_classes, record_dtype):
    return torch.empty(_packed_bytes(*vol.shape, "lcoef", record_dtype), device=vol.device, dtype=torch.uint8)


@torch.library.custom_op("nanodrr_amd::pack_labels", mutates_args=())
def pack_labels(lab: torch.Tensor, layout: str) -> torch.Tensor:
    """Labels in the form `layout`'s march reads (ndrr_pack_labels): the label quads for "quad",
    the padded (D + 4, H + 4, W + 4) map for "stencil" and "coef" (flattened, like the quads), a
    copy of the (D, H, W) int8 map for "linear"."""
    if not lab.is_cuda or lab.dim() != 3 or lab.dtype != torch.int8:
        raise ValueError("pack_labels needs a (D, H, W) int8 device tensor (Subject.label_i8)")
    if layout not in _lib.LAYOUTS:
        raise ValueError(f"unknown layout {layout!r}; expected one of {sorted(_lib.LAYOUTS)}")
    if layout == "lcoef":  # the records carry the classes
        return lab.new_empty(0)
    lab = lab.contiguous()
    D, H, W = lab.shape
    lib = _lib.lib()
    code = _lib.LAYOUTS[layout]
    out = torch.empty(lib.ndrr_packed_label_bytes(D, H, W, code), device=lab.device, dtype=torch.int8)
    with _on(lab.device):
        rc = lib.ndrr_pack_labels(lab.data_ptr(), code, D, H, W, out.data_ptr(), _stream(lab.device))
    _lib.check(rc, "ndrr_pack_labels")
    return out.view(D, H, W) if layout == "linear" else out


@pack_labels.register_fake
def _(lab, layout):
    D, H, W = lab.shape
    if layout == "lcoef":
        return lab.new_empty(0)
    if layout in ("quad", "qcoef"):
        return torch.empty((D + 1) * (H + 1) * (W + 2) * 4, device=lab.device, dtype=torch.int8)
    if layout in ("stencil", "coef"):
        return torch.empty((D + 4) * (H + 4) * (W + 4), device=lab.device, dtype=torch.int8)
    return torch.empty(D, H, W, device=lab.device, dtype=torch.int8)


def _geometry(vol, lab, M, src, tgt, step, n_samples, n_classes, width, layout="linear", B=None):
    """Validate shapes/devices and return the launch geometry (raises before any launch). M may be
    None when the kernels compose it themselves; B is then given."""
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
    if M is not None:
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
        if t is None:
            continue
        if t.dtype != torch.float32 or not t.is_contiguous() or t.device != vol.device:
            raise ValueError(f"{name} must be a contiguous float32 tensor on {vol.device}")
    if n_classes > 1:
        if lab.dtype != torch.int8 or not lab.is_contiguous() or lab.device != vol.device:
            raise ValueError("lab must be a contiguous int8 tensor on the volume's device")
        if layout != "linear":
            need = _lib.lib().ndrr_packed_label_bytes(*vol.shape, _lib.LAYOUTS[layout])
            if lab.numel() != need:
                raise ValueError(f"the {layout} layout reads its packed labels (pack_labels): {need} bytes, "
                                 f"got {lab.numel()}")
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


def _fused_raymarch_impl(
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
    B, N, D, H, W, strides = _geometry(vol, lab, M, src, tgt, step, n_samples, n_classes, width,
                                       layout if packed is not None else "linear")
    vptr, code = _volume_arg(vol, packed, layout)
    out = torch.empty(B, n_classes, N, device=vol.device, dtype=torch.float32)
    with _on(vol.device):
        rc = _lib.lib().ndrr_raymarch_fwd(
            vptr, _DTYPE_CODE[vol.dtype], code, _ptr(lab) if n_classes > 1 else None,
            M.data_ptr(), src.data_ptr(), tgt.data_ptr(), step.data_ptr(), out.data_ptr(),
            B, N, D, H, W, n_samples, n_classes, width, *strides, _stream(vol.device),
        )
    _lib.check(rc, "ndrr_raymarch_fwd")
    return out


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
    return _fused_raymarch_impl(vol, lab, M, src, tgt, step, n_samples, n_classes, width, packed, layout)


@fused_raymarch.register_fake
def _(vol, lab, M, src, tgt, step, n_samples, n_classes, width, packed=None, layout="linear"):
    return torch.empty(M.shape[0], n_classes, tgt.shape[1], device=vol.device, dtype=torch.float32)


# Volume gradients scatter through 4 x 4 x 4-voxel bricks (ndrr_raymarch_bwd_bricked): a wave's
# atomics at one step fill about half as many cache lines as in the (D, H, W) volume. False takes
# ndrr_raymarch_bwd (the same sums in another order of fp32 additions).
GVOL_BRICKS = True


def _fused_raymarch_bwd_impl(
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
    with _on(dev):
        args = (go.data_ptr(), vptr, _DTYPE_CODE[vol.dtype], code, _ptr(lab) if n_classes > 1 else None,
                M.data_ptr(), src.data_ptr(), tgt.data_ptr(), step.data_ptr(),
                ws.data_ptr(), gM.data_ptr(), _ptr(gsrc), _ptr(gtgt), _ptr(gstep), _ptr(gvol))
        tail = (int(need_gcam), B, N, D, H, W, n_samples, n_classes, width, *strides, _stream(dev))
        if need_gvol and GVOL_BRICKS:
            # the scatter's brick buffer, zero on entry (the call leaves it zero)
            bricks = torch.zeros(lib.ndrr_gvol_brick_floats(D, H, W), device=dev, dtype=torch.float32)
            rc = lib.ndrr_raymarch_bwd_bricked(*args, bricks.data_ptr(), *tail)
            _lib.check(rc, "ndrr_raymarch_bwd_bricked")
        else:
            rc = lib.ndrr_raymarch_bwd(*args, *tail)
            _lib.check(rc, "ndrr_raymarch_bwd")
    return gvol, gM, gsrc, gtgt, gstep


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
    return _fused_raymarch_bwd_impl(go, vol, lab, M, src, tgt, step, n_samples, n_classes, width, need_gvol, need_gcam, packed, layout)


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


def _backward(ctx, go, bwd=None):
    """Autograd formula of fused_raymarch; `bwd` is the adjoint to call (the custom op unless the
    eager path passes its implementation)."""
    vol, lab, M, src, tgt, step, packed = ctx.saved_tensors
    need_gvol = ctx.needs_input_grad[0]
    need_gcam = any(ctx.needs_input_grad[i] for i in (3, 4, 5))
    gvol, gM, gsrc, gtgt, gstep = (bwd or fused_raymarch_bwd)(
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


def _check_pose(w2v: torch.Tensor, rt_inv: torch.Tensor, vol: torch.Tensor) -> None:
    if w2v.dtype != torch.float64 or w2v.shape != (4, 4) or not w2v.is_contiguous() or w2v.device != vol.device:
        raise ValueError(f"world_to_voxel must be a contiguous float64 (4, 4) tensor on {vol.device}")
    if rt_inv.dtype != torch.float32 or rt_inv.dim() != 3 or rt_inv.shape[1:] != (4, 4) \
            or not rt_inv.is_contiguous() or rt_inv.device != vol.device:
        raise ValueError(f"rt_inv must be a contiguous float32 (B, 4, 4) tensor on {vol.device}")


def _jac_forward(vol, lab, M, pose, src, tgt, step, n_samples, n_classes, width, jac_planes, packed, layout,
                 lattice=None):
    """The single-pass forward with M (B, 3, 4) given, or pose = (world_to_voxel, rt_inv) and M
    composed by the kernels; jac_planes 0 (pose form only) skips the sums. lattice (B,) fp32, if
    given, offsets pose b's sample lattice by lattice[b] steps along its rays."""
    if jac_planes not in (0, 6, 7):
        raise ValueError(f"jac_planes must be 0, 6 or 7, got {jac_planes}")
    if pose is not None:
        _check_pose(*pose, vol)
    B, N, D, H, W, strides = _geometry(vol, lab, M, src, tgt, step, n_samples, n_classes, width,
                                       layout if packed is not None else "linear",
                                       B=None if pose is None else pose[1].shape[0])
    if lattice is not None and (lattice.dtype != torch.float32 or lattice.shape != (B,)
                                or not lattice.is_contiguous() or lattice.device != vol.device):
        raise ValueError(f"lattice must be a contiguous float32 ({B},) tensor on {vol.device}")
    vptr, code = _volume_arg(vol, packed, layout)
    out = torch.empty(B, n_classes, N, device=vol.device, dtype=torch.float32)
    jac = (torch.empty(B, jac_planes, n_classes, N, device=vol.device, dtype=torch.float32) if jac_planes
           else out.new_empty(0))
    lib = _lib.lib()
    head = (vptr, _DTYPE_CODE[vol.dtype], code, _ptr(lab) if n_classes > 1 else None)
    tail = (src.data_ptr(), tgt.data_ptr(), step.data_ptr(), _ptr(lattice), out.data_ptr(), _ptr(jac), jac_planes, B,
            N, D, H, W,
            n_samples, n_classes, width, *strides, _stream(vol.device))
    with _on(vol.device):
        if pose is None:
            rc = lib.ndrr_raymarch_fwd_jac(*head, M.data_ptr(), *tail)
        else:
            rc = lib.ndrr_raymarch_fwd_jac_pose(*head, pose[0].data_ptr(), pose[1].data_ptr(), *tail)
    _lib.check(rc, "ndrr_raymarch_fwd_jac" if pose is None else "ndrr_raymarch_fwd_jac_pose")
    return out, jac


def _fused_raymarch_jac_impl(
    vol: torch.Tensor,
    lab: torch.Tensor,
    M: torch.Tensor,
    src: torch.Tensor,
    tgt: torch.Tensor,
    step: torch.Tensor,
    n_samples: int,
    n_classes: int,
    width: int,
    jac_planes: int,
    packed: Optional[torch.Tensor] = None,
    layout: str = "linear",
    lattice: Optional[torch.Tensor] = None,
) -> tuple[torch.Tensor, torch.Tensor]:
    return _jac_forward(vol, lab, M, None, src, tgt, step, n_samples, n_classes, width, jac_planes, packed, layout,
                        lattice)


@torch.library.custom_op("nanodrr_amd::fused_raymarch_jac", mutates_args=())
def fused_raymarch_jac(
    vol: torch.Tensor,
    lab: torch.Tensor,
    M: torch.Tensor,
    src: torch.Tensor,
    tgt: torch.Tensor,
    step: torch.Tensor,
    n_samples: int,
    n_classes: int,
    width: int,
    jac_planes: int,
    packed: Optional[torch.Tensor] = None,
    layout: str = "linear",
    lattice: Optional[torch.Tensor] = None,
) -> tuple[torch.Tensor, torch.Tensor]:
    """`fused_raymarch` plus, from the same march, the per-ray sums of the
    single-pass adjoint: jac (B, jac_planes, C, N) fp32 (ndrr_raymarch_fwd_jac).
    Its autograd backward reads jac instead of marching the volume again, so it
    gives pose and camera gradients but no volume gradient. `lattice` (B,), if
    given, slides pose b's samples by lattice[b] steps along its rays."""
    return _fused_raymarch_jac_impl(vol, lab, M, src, tgt, step, n_samples, n_classes, width, jac_planes, packed,
                                    layout, lattice)


@fused_raymarch_jac.register_fake
def _(vol, lab, M, src, tgt, step, n_samples, n_classes, width, jac_planes, packed=None, layout="linear",
      lattice=None):
    B, N = M.shape[0], tgt.shape[1]
    return (torch.empty(B, n_classes, N, device=vol.device, dtype=torch.float32),
            torch.empty(B, jac_planes, n_classes, N, device=vol.device, dtype=torch.float32))


def _jac_backward(go, jac, M, pose, src, tgt, step, width, need_gcam):
    """The contraction with M given (returns gM (B, 3, 4)) or pose = (world_to_voxel, rt_inv) (returns
    the rt_inv gradient (B, 4, 4)), then gsrc, gtgt, gstep (empty without need_gcam)."""
    B, planes, C, N = jac.shape
    if (M is not None and M.shape != (B, 3, 4)) or (pose is not None and pose[1].shape != (B, 4, 4)) \
            or tgt.shape[1] != N:
        raise ValueError("jac does not match the geometry")
    if need_gcam and planes != 7:
        raise ValueError("camera gradients need the 7-plane jac (the forward ran without them)")
    if go.shape != (B, C, N):
        raise ValueError(f"go must be (B, C, N) = {(B, C, N)}, got {tuple(go.shape)}")
    go = go.to(torch.float32)
    # one (C, N) gradient broadcast over the poses is read in place (batch stride 0)
    shared = B > 1 and go.stride(0) == 0 and go[0].is_contiguous()
    go_bs = 0 if shared else C * N
    if not shared:
        go = go.contiguous()
    dev = jac.device
    lib = _lib.lib()
    strides = (0 if src.shape[0] == 1 else src.shape[1] * 3, 0 if src.shape[1] == 1 else 3,
               0 if tgt.shape[0] == 1 else N * 3, 0 if step.shape[0] == 1 else N)
    ws = torch.empty(lib.ndrr_bwd_workspace_floats(B, N, width), device=dev, dtype=torch.float32)
    g = torch.empty((B, 3, 4) if pose is None else (B, 4, 4), device=dev, dtype=torch.float32)
    gsrc = torch.empty(B, N, 3, device=dev, dtype=torch.float32) if need_gcam else g.new_empty(0)
    gtgt = torch.empty(B, N, 3, device=dev, dtype=torch.float32) if need_gcam else g.new_empty(0)
    gstep = torch.empty(B, N, device=dev, dtype=torch.float32) if need_gcam else g.new_empty(0)
    mid = (src.data_ptr(), tgt.data_ptr(), step.data_ptr(), ws.data_ptr(), g.data_ptr(), _ptr(gsrc), _ptr(gtgt),
           _ptr(gstep), int(need_gcam), B, N, C, width, *strides, _stream(dev))
    with _on(dev):
        if pose is None:
            rc = lib.ndrr_raymarch_bwd_jac(go.data_ptr(), go_bs, jac.data_ptr(), planes, M.data_ptr(), *mid)
        else:
            rc = lib.ndrr_raymarch_bwd_jac_pose(go.data_ptr(), go_bs, jac.data_ptr(), planes, pose[0].data_ptr(),
                                                pose[1].data_ptr(), *mid)
    _lib.check(rc, "ndrr_raymarch_bwd_jac" if pose is None else "ndrr_raymarch_bwd_jac_pose")
    return g, gsrc, gtgt, gstep


def _fused_raymarch_bwd_jac_impl(
    go: torch.Tensor,
    jac: torch.Tensor,
    M: torch.Tensor,
    src: torch.Tensor,
    tgt: torch.Tensor,
    step: torch.Tensor,
    width: int,
    need_gcam: bool,
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    return _jac_backward(go, jac, M, None, src, tgt, step, width, need_gcam)


@torch.library.custom_op("nanodrr_amd::fused_raymarch_bwd_jac", mutates_args=())
def fused_raymarch_bwd_jac(
    go: torch.Tensor,
    jac: torch.Tensor,
    M: torch.Tensor,
    src: torch.Tensor,
    tgt: torch.Tensor,
    step: torch.Tensor,
    width: int,
    need_gcam: bool,
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """Backward of `fused_raymarch_jac` from its jac output (ndrr_raymarch_bwd_jac):
    (gM (B,3,4), gsrc (B,N,3)|empty, gtgt (B,N,3)|empty, gstep (B,N)|empty), fp32."""
    return _fused_raymarch_bwd_jac_impl(go, jac, M, src, tgt, step, width, need_gcam)


@fused_raymarch_bwd_jac.register_fake
def _(go, jac, M, src, tgt, step, width, need_gcam):
    B, _, _, N = jac.shape
    gM = torch.empty(B, 3, 4, device=M.device, dtype=torch.float32)
    gsrc = torch.empty(B, N, 3, device=M.device, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    gtgt = torch.empty(B, N, 3, device=M.device, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    gstep = torch.empty(B, N, device=M.device, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    return gM, gsrc, gtgt, gstep


def _setup_context_jac(ctx, inputs, output):
    vol, lab, M, src, tgt, step, n_samples, n_classes, width, jac_planes, packed, layout = inputs[:12]
    ctx.save_for_backward(M, src, tgt, step, output[1])
    ctx.mark_non_differentiable(output[1])
    ctx.set_materialize_grads(False)  # no zero-filled gradient for the (large) jac planes
    ctx.width = width


def _backward_jac(ctx, go, g_jac, bwd=None):
    """Autograd formula of fused_raymarch_jac; `bwd` as in _backward."""
    if go is None:
        return (None,) * 13
    M, src, tgt, step, jac = ctx.saved_tensors
    if ctx.needs_input_grad[0]:
        raise RuntimeError("fused_raymarch_jac gives no volume gradient; use fused_raymarch for it")
    if jac.numel() == 0:
        raise RuntimeError("fused_raymarch_jac ran with jac_planes=0: it has no gradient")
    need_gcam = any(ctx.needs_input_grad[i] for i in (3, 4, 5))
    gM, gsrc, gtgt, gstep = (bwd or fused_raymarch_bwd_jac)(go, jac, M, src, tgt, step, ctx.width, need_gcam)
    return (
        None,  # vol
        None,  # lab
        gM.to(M.dtype),
        _reduce_to(gsrc, src) if need_gcam else None,
        _reduce_to(gtgt, tgt) if need_gcam else None,
        _reduce_to(gstep, step) if need_gcam else None,
        None, None, None, None, None, None, None,
    )


torch.library.register_autograd("nanodrr_amd::fused_raymarch_jac", _backward_jac, setup_context=_setup_context_jac)


def _fused_raymarch_pose_impl(
    vol: torch.Tensor,
    lab: torch.Tensor,
    world_to_voxel: torch.Tensor,
    rt_inv: torch.Tensor,
    src: torch.Tensor,
    tgt: torch.Tensor,
    step: torch.Tensor,
    n_samples: int,
    n_classes: int,
    width: int,
    jac_planes: int,
    packed: Optional[torch.Tensor] = None,
    layout: str = "linear",
    lattice: Optional[torch.Tensor] = None,
) -> tuple[torch.Tensor, torch.Tensor]:
    return _jac_forward(vol, lab, None, (world_to_voxel, rt_inv), src, tgt, step, n_samples, n_classes, width,
                        jac_planes, packed, layout, lattice)


@torch.library.custom_op("nanodrr_amd::fused_raymarch_pose", mutates_args=())
def fused_raymarch_pose(
    vol: torch.Tensor,
    lab: torch.Tensor,
    world_to_voxel: torch.Tensor,
    rt_inv: torch.Tensor,
    src: torch.Tensor,
    tgt: torch.Tensor,
    step: torch.Tensor,
    n_samples: int,
    n_classes: int,
    width: int,
    jac_planes: int,
    packed: Optional[torch.Tensor] = None,
    layout: str = "linear",
    lattice: Optional[torch.Tensor] = None,
) -> tuple[torch.Tensor, torch.Tensor]:
    """`fused_raymarch_jac` taking the pose itself: M = (world_to_voxel @ rt_inv)[:, :3] is composed
    in FP64 by each workgroup (ndrr_raymarch_fwd_jac_pose), bitwise equal to camera_to_voxel then
    fused_raymarch_jac. jac_planes 0 returns an empty jac (no gradient); the backward gives the
    rt_inv gradient (B, 4, 4) and the camera gradients, no volume gradient."""
    return _fused_raymarch_pose_impl(vol, lab, world_to_voxel, rt_inv, src, tgt, step, n_samples, n_classes, width,
                                     jac_planes, packed, layout, lattice)


@fused_raymarch_pose.register_fake
def _(vol, lab, world_to_voxel, rt_inv, src, tgt, step, n_samples, n_classes, width, jac_planes, packed=None,
      layout="linear", lattice=None):
    B, N = rt_inv.shape[0], tgt.shape[1]
    out = torch.empty(B, n_classes, N, device=vol.device, dtype=torch.float32)
    return out, (torch.empty(B, jac_planes, n_classes, N, device=vol.device, dtype=torch.float32) if jac_planes
                 else out.new_empty(0))


def _fused_raymarch_pose_bwd_impl(
    go: torch.Tensor,
    jac: torch.Tensor,
    world_to_voxel: torch.Tensor,
    rt_inv: torch.Tensor,
    src: torch.Tensor,
    tgt: torch.Tensor,
    step: torch.Tensor,
    width: int,
    need_gcam: bool,
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    return _jac_backward(go, jac, None, (world_to_voxel, rt_inv), src, tgt, step, width, need_gcam)


@torch.library.custom_op("nanodrr_amd::fused_raymarch_pose_bwd", mutates_args=())
def fused_raymarch_pose_bwd(
    go: torch.Tensor,
    jac: torch.Tensor,
    world_to_voxel: torch.Tensor,
    rt_inv: torch.Tensor,
    src: torch.Tensor,
    tgt: torch.Tensor,
    step: torch.Tensor,
    width: int,
    need_gcam: bool,
) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """Backward of `fused_raymarch_pose` from its jac output (ndrr_raymarch_bwd_jac_pose):
    (g_rt_inv (B,4,4), gsrc (B,N,3)|empty, gtgt (B,N,3)|empty, gstep (B,N)|empty), fp32."""
    return _fused_raymarch_pose_bwd_impl(go, jac, world_to_voxel, rt_inv, src, tgt, step, width, need_gcam)


@fused_raymarch_pose_bwd.register_fake
def _(go, jac, world_to_voxel, rt_inv, src, tgt, step, width, need_gcam):
    B, _, _, N = jac.shape
    g = torch.empty(B, 4, 4, device=jac.device, dtype=torch.float32)
    gsrc = torch.empty(B, N, 3, device=jac.device, dtype=torch.float32) if need_gcam else g.new_empty(0)
    gtgt = torch.empty(B, N, 3, device=jac.device, dtype=torch.float32) if need_gcam else g.new_empty(0)
    gstep = torch.empty(B, N, device=jac.device, dtype=torch.float32) if need_gcam else g.new_empty(0)
    return g, gsrc, gtgt, gstep


def _setup_context_pose(ctx, inputs, output):
    vol, lab, w2v, rt_inv, src, tgt, step, n_samples, n_classes, width, jac_planes, packed, layout = inputs[:13]
    ctx.save_for_backward(w2v, rt_inv, src, tgt, step, output[1])
    ctx.mark_non_differentiable(output[1])
    ctx.set_materialize_grads(False)
    ctx.width = width


def _backward_pose(ctx, go, g_jac, bwd=None):
    """Autograd formula of fused_raymarch_pose; `bwd` as in _backward."""
    if go is None:
        return (None,) * 14
    w2v, rt_inv, src, tgt, step, jac = ctx.saved_tensors
    if ctx.needs_input_grad[0] or ctx.needs_input_grad[2]:
        raise RuntimeError("fused_raymarch_pose gives no volume or world_to_voxel gradient")
    if jac.numel() == 0:
        raise RuntimeError("fused_raymarch_pose ran with jac_planes=0: it has no gradient")
    need_gcam = any(ctx.needs_input_grad[i] for i in (4, 5, 6))
    g_rt, gsrc, gtgt, gstep = (bwd or fused_raymarch_pose_bwd)(go, jac, w2v, rt_inv, src, tgt, step, ctx.width,
                                                               need_gcam)
    return (
        None, None, None,  # vol, lab, world_to_voxel
        g_rt if ctx.needs_input_grad[3] else None,
        _reduce_to(gsrc, src) if need_gcam else None,
        _reduce_to(gtgt, tgt) if need_gcam else None,
        _reduce_to(gstep, step) if need_gcam else None,
        None, None, None, None, None, None, None,
    )


torch.library.register_autograd("nanodrr_amd::fused_raymarch_pose", _backward_pose, setup_context=_setup_context_pose)


def _make_rays_impl(k_inv: torch.Tensor, sdd: torch.Tensor, height: int, width: int, orthographic: bool,
                    n_samples: int) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    if k_inv.dtype != torch.float32 or sdd.dtype != torch.float32 or not k_inv.is_cuda or sdd.device != k_inv.device:
        raise ValueError("make_rays needs float32 device tensors k_inv and sdd")
    k = k_inv.contiguous()
    d = sdd.contiguous()
    if k.dim() != 3 or k.shape[1:] != (3, 3) or d.dim() != 1:
        raise ValueError("k_inv must be (B, 3, 3) and sdd (B,)")
    B = max(k.shape[0], d.shape[0])
    if k.shape[0] not in (1, B) or d.shape[0] not in (1, B):
        raise ValueError("k_inv and sdd batches must match or be 1")
    N = height * width
    dev = k.device
    tgt = torch.empty(B, N, 3, device=dev, dtype=torch.float32)
    src = torch.empty(B, N if orthographic else 1, 3, device=dev, dtype=torch.float32)
    step = torch.empty(B, N, device=dev, dtype=torch.float32)
    with _on(dev):
        rc = _lib.lib().ndrr_make_rays(k.data_ptr(), 0 if k.shape[0] == 1 else 9, d.data_ptr(),
                                       0 if d.shape[0] == 1 else 1, B, height, width, int(orthographic), n_samples,
                                       tgt.data_ptr(), src.data_ptr(), step.data_ptr(), _stream(dev))
    _lib.check(rc, "ndrr_make_rays")
    return tgt, src, step


@torch.library.custom_op("nanodrr_amd::make_rays", mutates_args=())
def make_rays(k_inv: torch.Tensor, sdd: torch.Tensor, height: int, width: int, orthographic: bool,
              n_samples: int) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """(tgt (B,N,3), src (B,1,3) or (B,N,3), step (B,N)) fp32 of `render`'s rays in one
    launch (ndrr_make_rays); B = max of the batches of k_inv (B|1,3,3) and sdd (B|1,)."""
    return _make_rays_impl(k_inv, sdd, height, width, orthographic, n_samples)


@make_rays.register_fake
def _(k_inv, sdd, height, width, orthographic, n_samples):
    B, N = max(k_inv.shape[0], sdd.shape[0]), height * width
    return (torch.empty(B, N, 3, device=k_inv.device, dtype=torch.float32),
            torch.empty(B, N if orthographic else 1, 3, device=k_inv.device, dtype=torch.float32),
            torch.empty(B, N, device=k_inv.device, dtype=torch.float32))


def _c2v_check(world_to_voxel, t, name):
    if world_to_voxel.shape != (4, 4) or world_to_voxel.dtype != torch.float64:
        raise ValueError("world_to_voxel must be a (4, 4) float64 tensor")
    if t.dtype != torch.float32 or not t.is_contiguous() or t.device != world_to_voxel.device or not t.is_cuda:
        raise ValueError(f"{name} must be a contiguous float32 device tensor on the device of world_to_voxel")


def _camera_to_voxel_impl(world_to_voxel: torch.Tensor, rt_inv: torch.Tensor) -> torch.Tensor:
    world_to_voxel = world_to_voxel.contiguous()
    _c2v_check(world_to_voxel, rt_inv, "rt_inv")
    if rt_inv.dim() != 3 or rt_inv.shape[1:] != (4, 4):
        raise ValueError("rt_inv must be (B, 4, 4)")
    B = rt_inv.shape[0]
    M = torch.empty(B, 3, 4, device=rt_inv.device, dtype=torch.float32)
    with _on(rt_inv.device):
        rc = _lib.lib().ndrr_camera_to_voxel(world_to_voxel.data_ptr(), rt_inv.data_ptr(), M.data_ptr(), B,
                                             _stream(rt_inv.device))
    _lib.check(rc, "ndrr_camera_to_voxel")
    return M


@torch.library.custom_op("nanodrr_amd::camera_to_voxel", mutates_args=())
def camera_to_voxel(world_to_voxel: torch.Tensor, rt_inv: torch.Tensor) -> torch.Tensor:
    """M = (world_to_voxel @ rt_inv)[:, :3] (B,3,4) fp32, in FP64 with one rounding (ndrr_camera_to_voxel)."""
    return _camera_to_voxel_impl(world_to_voxel, rt_inv)


@camera_to_voxel.register_fake
def _(world_to_voxel, rt_inv):
    return torch.empty(rt_inv.shape[0], 3, 4, device=rt_inv.device, dtype=torch.float32)


def _camera_to_voxel_bwd_impl(world_to_voxel: torch.Tensor, gM: torch.Tensor) -> torch.Tensor:
    gM = gM.to(torch.float32).contiguous()
    world_to_voxel = world_to_voxel.contiguous()
    _c2v_check(world_to_voxel, gM, "gM")
    B = gM.shape[0]
    g = torch.empty(B, 4, 4, device=gM.device, dtype=torch.float32)
    with _on(gM.device):
        rc = _lib.lib().ndrr_camera_to_voxel_bwd(world_to_voxel.data_ptr(), gM.data_ptr(), g.data_ptr(), B,
                                                 _stream(gM.device))
    _lib.check(rc, "ndrr_camera_to_voxel_bwd")
    return g


@torch.library.custom_op("nanodrr_amd::camera_to_voxel_bwd", mutates_args=())
def camera_to_voxel_bwd(world_to_voxel: torch.Tensor, gM: torch.Tensor) -> torch.Tensor:
    """Adjoint of camera_to_voxel: world_to_voxel[:3]^T @ gM (B,4,4) fp32, FP64 sums."""
    return _camera_to_voxel_bwd_impl(world_to_voxel, gM)


@camera_to_voxel_bwd.register_fake
def _(world_to_voxel, gM):
    return torch.empty(gM.shape[0], 4, 4, device=gM.device, dtype=torch.float32)


def _c2v_setup(ctx, inputs, output):
    ctx.save_for_backward(inputs[0])


def _c2v_backward(ctx, gM, bwd=None):
    (w2v,) = ctx.saved_tensors
    return None, (bwd or camera_to_voxel_bwd)(w2v, gM)


torch.library.register_autograd("nanodrr_amd::camera_to_voxel", _c2v_backward, setup_context=_c2v_setup)


@torch.library.custom_op("nanodrr_amd::labels_to_i8", mutates_args=())
def labels_to_i8(label: torch.Tensor, n_classes: int) -> torch.Tensor:
    """int8 label map for the kernels: trunc(label) if in [0, n_classes) else -1."""
    if not label.is_cuda:
        raise ValueError("labels_to_i8 runs on the GPU: pass a device tensor")
    src = label.detach().to(torch.float32).contiguous()
    dst = torch.empty(src.shape, device=src.device, dtype=torch.int8)
    with _on(src.device):
        rc = _lib.lib().ndrr_labels_to_i8(src.data_ptr(), dst.data_ptr(), src.numel(), n_classes,
                                          _stream(src.device))
    _lib.check(rc, "ndrr_labels_to_i8")
    return dst


@labels_to_i8.register_fake
def _(label, n_classes):
    return torch.empty(label.shape, device=label.device, dtype=torch.int8)


# ---------------------------------------------------------------------------
# Eager entry points. A torch.library custom op costs tens of microseconds of
# host dispatch per call (more than most of the kernels it launches), so eager
# callers reach the implementations directly: through an autograd.Function
# that carries the op's own formula when an input needs a gradient, else as a
# plain call. While torch.compile traces, the registered custom ops (with
# their fake implementations) are used instead.
# ---------------------------------------------------------------------------
def _eager(name, op, impl, setup=None, backward=None):
    fn = None
    if setup is not None:
        def forward(ctx, *args):
            out = impl(*args)
            setup(ctx, args, out)
            return out

        fn = type(name, (torch.autograd.Function,), {"forward": staticmethod(forward),
                                                      "backward": staticmethod(backward)})

    def call(*args):
        if torch.compiler.is_compiling():
            return op(*args)
        if fn is not None and torch.is_grad_enabled() and any(
                isinstance(a, torch.Tensor) and a.requires_grad for a in args):
            return fn.apply(*args)
        return impl(*args)

    call.__name__ = call.__qualname__ = name
    call.__doc__ = f"Eager entry point of the custom op {name} (same arguments and result)."
    return call


eager = SimpleNamespace(
    fused_raymarch=_eager("fused_raymarch", fused_raymarch, _fused_raymarch_impl, _setup_context,
                          lambda ctx, go: _backward(ctx, go, _fused_raymarch_bwd_impl)),
    fused_raymarch_jac=_eager("fused_raymarch_jac", fused_raymarch_jac, _fused_raymarch_jac_impl, _setup_context_jac,
                              lambda ctx, go, g_jac: _backward_jac(ctx, go, g_jac, _fused_raymarch_bwd_jac_impl)),
    fused_raymarch_pose=_eager("fused_raymarch_pose", fused_raymarch_pose, _fused_raymarch_pose_impl,
                               _setup_context_pose,
                               lambda ctx, go, g_jac: _backward_pose(ctx, go, g_jac, _fused_raymarch_pose_bwd_impl)),
    make_rays=_eager("make_rays", make_rays, _make_rays_impl),
    camera_to_voxel=_eager("camera_to_voxel", camera_to_voxel, _camera_to_voxel_impl, _c2v_setup,
                           lambda ctx, gM: _c2v_backward(ctx, gM, _camera_to_voxel_bwd_impl)),
)
