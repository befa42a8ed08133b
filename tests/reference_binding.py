"""The binding a nanodrr maintainer would add: nanodrr's own fused ops, under their
names and schemas, on libnanodrr_amd through ctypes.

`nanodrr::fused_raymarch(vol, lab, M, src, tgt, step, n_samples, n_classes,
width) -> Tensor` and `nanodrr::fused_raymarch_bwd(go, vol, lab, M, src, tgt,
step, n_samples, n_classes, width, need_gvol, need_gcam)` as registered in
src/nanodrr/drr/_fused.py:415-460 and :477-531, with the reference's fake
impls and autograd formula (:463-467, :534-542, :545-567): `lab` is the float
label volume (any of the volume's dtypes; the reference passes `vol` itself
when there is one class), the forward returns `vol.dtype`, gradients come back
in their inputs' dtypes. Only the C ABI of include/nanodrr_amd.h is used (no
nanodrr_amd Python code), so this module is what a nanodrr `_hip.py` would
hold: the float labels go through ndrr_labels_to_i8 once per call, the volume
is read in its linear layout, and every launch runs on torch's current stream.
Test infrastructure: imported by the reference-schema tests only.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import torch

_LIB_PATH = Path(__file__).resolve().parents[1] / "nanodrr_amd" / "libnanodrr_amd.so"
P, I, L = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
_DT = {torch.float32: 0, torch.bfloat16: 1, torch.float16: 2}
_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        h = ctypes.CDLL(str(_LIB_PATH))
        h.ndrr_strerror.restype, h.ndrr_strerror.argtypes = ctypes.c_char_p, [I]
        h.ndrr_labels_to_i8.restype, h.ndrr_labels_to_i8.argtypes = I, [P, P, L, I, P]
        h.ndrr_raymarch_fwd.restype = I
        h.ndrr_raymarch_fwd.argtypes = [P, I, I, P, P, P, P, P, P] + [I] * 8 + [L] * 4 + [P]
        h.ndrr_bwd_workspace_floats.restype, h.ndrr_bwd_workspace_floats.argtypes = ctypes.c_size_t, [I, I, I]
        h.ndrr_raymarch_bwd.restype = I
        h.ndrr_raymarch_bwd.argtypes = [P, P, I, I, P, P, P, P, P, P, P, P, P, P, P, I] + [I] * 8 + [L] * 4 + [P]
        _lib = h
    return _lib


def _check(rc: int, what: str) -> None:
    if rc:
        raise RuntimeError(f"{what}: {lib().ndrr_strerror(rc).decode()}")


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _strides(src, tgt, step):
    N = tgt.shape[1]
    return (src.shape[1] * 3 if src.shape[0] > 1 else 0, 0 if src.shape[1] == 1 else 3,
            N * 3 if tgt.shape[0] > 1 else 0, N if step.shape[0] > 1 else 0)


def _classes(lab: torch.Tensor, vol: torch.Tensor, n_classes: int):
    """The float label volume as the kernels' int8 classes (trunc, -1 outside [0, C))."""
    if n_classes == 1:
        return None
    lab8 = torch.empty(vol.shape, device=vol.device, dtype=torch.int8)
    src = lab.detach().to(torch.float32).contiguous()
    _check(lib().ndrr_labels_to_i8(src.data_ptr(), lab8.data_ptr(), src.numel(), n_classes, _stream()),
           "ndrr_labels_to_i8")
    return lab8


@torch.library.custom_op("nanodrr::fused_raymarch", mutates_args=())
def fused_raymarch(vol: torch.Tensor, lab: torch.Tensor, M: torch.Tensor, src: torch.Tensor, tgt: torch.Tensor,
                   step: torch.Tensor, n_samples: int, n_classes: int, width: int) -> torch.Tensor:
    B, N, _ = tgt.shape
    D, H, W = vol.shape
    vol = vol.contiguous()
    M = M.contiguous().float()
    lab8 = _classes(lab, vol, n_classes)
    out = torch.empty(B, n_classes, N, device=vol.device, dtype=torch.float32)
    rc = lib().ndrr_raymarch_fwd(vol.data_ptr(), _DT[vol.dtype], 0, None if lab8 is None else lab8.data_ptr(),
                                 M.data_ptr(), src.data_ptr(), tgt.data_ptr(), step.data_ptr(), out.data_ptr(),
                                 B, N, D, H, W, n_samples, n_classes, width, *_strides(src, tgt, step), _stream())
    _check(rc, "ndrr_raymarch_fwd")
    return out.to(vol.dtype)


@fused_raymarch.register_fake
def _(vol, lab, M, src, tgt, step, n_samples, n_classes, width):
    B, N, _ = tgt.shape
    return torch.empty(B, n_classes, N, device=vol.device, dtype=vol.dtype)


@torch.library.custom_op("nanodrr::fused_raymarch_bwd", mutates_args=())
def fused_raymarch_bwd(go: torch.Tensor, vol: torch.Tensor, lab: torch.Tensor, M: torch.Tensor, src: torch.Tensor,
                       tgt: torch.Tensor, step: torch.Tensor, n_samples: int, n_classes: int, width: int,
                       need_gvol: bool, need_gcam: bool
                       ) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    if need_gvol and torch.are_deterministic_algorithms_enabled():
        raise RuntimeError("Volume gradients use nondeterministic atomics; disable deterministic algorithms")
    B, N, _ = tgt.shape
    D, H, W = vol.shape
    vol = vol.contiguous()
    M = M.contiguous().float()
    go = go.contiguous().float()
    lab8 = _classes(lab, vol, n_classes)
    dev = vol.device
    ws = torch.empty(lib().ndrr_bwd_workspace_floats(B, N, width), device=dev, dtype=torch.float32)
    gM = torch.empty(B, 3, 4, device=dev, dtype=torch.float32)
    gvol = torch.zeros(vol.shape, device=dev, dtype=torch.float32) if need_gvol else gM.new_empty(0)
    gsrc = torch.empty(B, N, 3, device=dev, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    gtgt = torch.empty(B, N, 3, device=dev, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    gstep = torch.empty(B, N, device=dev, dtype=torch.float32) if need_gcam else gM.new_empty(0)

    def ptr(t):
        return t.data_ptr() if t.numel() else None

    rc = lib().ndrr_raymarch_bwd(go.data_ptr(), vol.data_ptr(), _DT[vol.dtype], 0,
                                 None if lab8 is None else lab8.data_ptr(), M.data_ptr(), src.data_ptr(),
                                 tgt.data_ptr(), step.data_ptr(), ws.data_ptr(), gM.data_ptr(), ptr(gsrc), ptr(gtgt),
                                 ptr(gstep), ptr(gvol), int(need_gcam), B, N, D, H, W, n_samples, n_classes, width,
                                 *_strides(src, tgt, step), _stream())
    _check(rc, "ndrr_raymarch_bwd")
    return gvol, gM, gsrc, gtgt, gstep


@fused_raymarch_bwd.register_fake
def _(go, vol, lab, M, src, tgt, step, n_samples, n_classes, width, need_gvol, need_gcam):
    B, N, _ = tgt.shape
    gM = torch.empty(B, 3, 4, device=M.device, dtype=torch.float32)
    gvol = torch.empty(vol.shape, device=M.device, dtype=torch.float32) if need_gvol else gM.new_empty(0)
    gsrc = torch.empty(B, N, 3, device=M.device, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    gtgt = torch.empty(B, N, 3, device=M.device, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    gstep = torch.empty(B, N, device=M.device, dtype=torch.float32) if need_gcam else gM.new_empty(0)
    return gvol, gM, gsrc, gtgt, gstep


def _setup_context(ctx, inputs, output):
    vol, lab, M, src, tgt, step, n_samples, n_classes, width = inputs
    ctx.save_for_backward(vol, lab, M.contiguous(), src, tgt, step)
    ctx.n_samples, ctx.n_classes, ctx.width = n_samples, n_classes, width


def _backward(ctx, go):
    vol, lab, M, src, tgt, step = ctx.saved_tensors
    need_gvol = ctx.needs_input_grad[0]
    need_gcam = any(ctx.needs_input_grad[i] for i in (3, 4, 5))
    gvol, gM, gsrc, gtgt, gstep = fused_raymarch_bwd(go, vol, lab, M, src, tgt, step, ctx.n_samples, ctx.n_classes,
                                                     ctx.width, need_gvol, need_gcam)
    if need_gcam and src.shape[1] == 1:
        gsrc = gsrc.sum(dim=1, keepdim=True)
    return (gvol.to(vol.dtype) if need_gvol else None, None, gM.to(M.dtype),
            gsrc.to(src.dtype) if need_gcam else None, gtgt.to(tgt.dtype) if need_gcam else None,
            gstep.to(step.dtype) if need_gcam else None, None, None, None)


torch.library.register_autograd("nanodrr::fused_raymarch", _backward, setup_context=_setup_context)
