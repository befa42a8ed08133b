"""Image and pose metrics for registration.

`zncc` is the native zero-normalised cross-correlation (ndrr_zncc: chunk
moments over (4096-pixel chunk, image) workgroups merged in a fixed order,
deterministic) with its gradient. `NormalizedCrossCorrelation2d`,
`MultiscaleNormalizedCrossCorrelation2d`, `GradientNormalizedCrossCorrelation2d`
and `DoubleGeodesicSE3` keep the API of src/nanodrr/metrics/ncc.py:8-137 and
src/nanodrr/metrics/geo.py:7-45. Every NCC variant is a mean of ZNCCs on the
native kernels: a multi-channel image is one image per channel, a patch size
p averages the ZNCC of every p x p window (the reference unfolds the windows
and standardises them one by one; `ndrr_patch_zncc` reads them in place), and
the gradient NCC runs on the two Sobel responses. `mtre` is the mean target
registration error on fiducials (no counterpart in the reference).
"""

from __future__ import annotations

import torch
import torch.nn.functional as F

from . import _lib
from ._ops import _eager, _on
from .geometry import matmul3, rotmat_geodesic_distance


def _zncc_impl(x: torch.Tensor, y: torch.Tensor, eps: float) -> tuple[torch.Tensor, torch.Tensor]:
    B = x.shape[0]
    x2 = x.detach().reshape(B, -1).to(torch.float32).contiguous()
    y2 = y.detach().reshape(B, -1).to(torch.float32).contiguous()
    if x2.shape != y2.shape:
        raise ValueError("Input images must be the same size")
    if not x2.is_cuda:
        raise ValueError("zncc runs on the GPU: pass device tensors")
    score = torch.empty(B, device=x.device, dtype=torch.float32)
    grad = torch.empty_like(x2)
    ws = torch.empty(_lib.lib().ndrr_zncc_workspace_floats(B, x2.shape[1]), device=x.device, dtype=torch.float32)
    with _on(x.device):
        rc = _lib.lib().ndrr_zncc(x2.data_ptr(), y2.data_ptr(), B, x2.shape[1], float(eps), ws.data_ptr(),
                                  score.data_ptr(), grad.data_ptr(), torch.cuda.current_stream(x.device).cuda_stream)
    _lib.check(rc, "ndrr_zncc")
    return score, grad.view(x.shape)


@torch.library.custom_op("nanodrr_amd::zncc", mutates_args=())
def _zncc(x: torch.Tensor, y: torch.Tensor, eps: float) -> tuple[torch.Tensor, torch.Tensor]:
    return _zncc_impl(x, y, eps)


@_zncc.register_fake
def _(x, y, eps):
    return x.new_empty(x.shape[0], dtype=torch.float32), torch.empty(x.shape, device=x.device, dtype=torch.float32)


def _zncc_setup(ctx, inputs, output):
    ctx.save_for_backward(output[1])
    ctx.mark_non_differentiable(output[1])
    ctx.set_materialize_grads(False)  # no zero-filled gradient for the unused per-pixel output
    ctx.x_dtype = inputs[0].dtype


def _zncc_backward(ctx, g_score, g_grad):
    if g_score is None:
        return None, None, None
    (grad,) = ctx.saved_tensors
    gx = grad * g_score.reshape(-1, *([1] * (grad.dim() - 1)))
    return gx.to(ctx.x_dtype), None, None


torch.library.register_autograd("nanodrr_amd::zncc", _zncc_backward, setup_context=_zncc_setup)

# eager callers skip the custom-op dispatch (see nanodrr_amd._ops._eager)
_zncc_eager = _eager("zncc", _zncc, _zncc_impl, _zncc_setup, _zncc_backward)


def _patch_zncc_impl(x: torch.Tensor, y: torch.Tensor, patch: int, eps: float) -> tuple[torch.Tensor, torch.Tensor]:
    if x.shape != y.shape or x.dim() != 3:
        raise ValueError("patch_zncc takes two (B, H, W) image batches of the same size")
    if not x.is_cuda:
        raise ValueError("patch_zncc runs on the GPU: pass device tensors")
    B, H, W = x.shape
    if not 1 <= patch <= min(H, W, 32):
        raise ValueError(f"patch size {patch} must lie in [1, min(H, W, 32)] = [1, {min(H, W, 32)}]")
    x2 = x.detach().to(torch.float32).contiguous()
    y2 = y.detach().to(torch.float32).contiguous()
    lib = _lib.lib()
    score = torch.empty(B, device=x.device, dtype=torch.float32)
    grad = torch.empty_like(x2)
    ws = torch.empty(lib.ndrr_patch_zncc_workspace_floats(B, H, W, patch), device=x.device, dtype=torch.float32)
    with _on(x.device):
        rc = lib.ndrr_patch_zncc(x2.data_ptr(), y2.data_ptr(), B, H, W, patch, float(eps), ws.data_ptr(),
                                 score.data_ptr(), grad.data_ptr(), torch.cuda.current_stream(x.device).cuda_stream)
    _lib.check(rc, "ndrr_patch_zncc")
    return score, grad


@torch.library.custom_op("nanodrr_amd::patch_zncc", mutates_args=())
def _patch_zncc(x: torch.Tensor, y: torch.Tensor, patch: int, eps: float) -> tuple[torch.Tensor, torch.Tensor]:
    return _patch_zncc_impl(x, y, patch, eps)


@_patch_zncc.register_fake
def _(x, y, patch, eps):
    return x.new_empty(x.shape[0], dtype=torch.float32), torch.empty(x.shape, device=x.device, dtype=torch.float32)


def _patch_zncc_backward(ctx, g_score, g_grad):
    gx = _zncc_backward(ctx, g_score, g_grad)[0]
    return gx, None, None, None


torch.library.register_autograd("nanodrr_amd::patch_zncc", _patch_zncc_backward, setup_context=_zncc_setup)

_patch_zncc_eager = _eager("patch_zncc", _patch_zncc, _patch_zncc_impl, _zncc_setup, _patch_zncc_backward)


def patch_zncc(x: torch.Tensor, y: torch.Tensor, patch: int, eps: float = 1e-7) -> torch.Tensor:
    """(B,) mean ZNCC over every patch x patch window (stride 1) of (B, H, W) image pairs
    (ndrr_patch_zncc), differentiable in x."""
    return _patch_zncc_eager(x, y, patch, eps)[0]


def zncc(x: torch.Tensor, y: torch.Tensor, eps: float = 1e-7) -> torch.Tensor:
    """(B,) zero-normalised cross-correlation of x and y over all non-batch dims."""
    return _zncc_eager(x, y, eps)[0]


def _channel_zncc(x1: torch.Tensor, x2: torch.Tensor, eps: float, patch: int | None = None) -> torch.Tensor:
    """(B,) mean over the channels of (B, C, H, W) image pairs of each channel's ZNCC, whole
    images (ndrr_zncc) or, with `patch`, the mean over every patch x patch window (ndrr_patch_zncc)."""
    if x1.shape != x2.shape:
        raise ValueError(f"Input images must be the same size, got {tuple(x1.shape)} and {tuple(x2.shape)}")
    B, C = x1.shape[:2]
    if patch is None:
        rows = zncc(x1.reshape(B * C, -1), x2.reshape(B * C, -1), eps)
    else:
        rows = patch_zncc(x1.reshape(B * C, *x1.shape[-2:]), x2.reshape(B * C, *x2.shape[-2:]), patch, eps)
    return rows.view(B, C).mean(dim=1)


class NormalizedCrossCorrelation2d(torch.nn.Module):
    """(B,) NCC of two image batches (B, C, H, W) (ncc.py:8-37): every channel (or, with
    `patch_size` p, every p x p window of it, stride 1: ncc.py:83-88) is standardised with its
    population variance + eps and the score is the mean product over them. Runs on the native
    ZNCC kernels."""

    def __init__(self, patch_size: int | None = None, eps: float = 1e-7):
        super().__init__()
        self.patch_size = patch_size
        self.eps = eps

    def forward(self, x1: torch.Tensor, x2: torch.Tensor) -> torch.Tensor:
        return _channel_zncc(x1, x2, self.eps, self.patch_size)


class MultiscaleNormalizedCrossCorrelation2d(torch.nn.Module):
    """(B,) weighted sum of NCCs at several patch sizes (ncc.py:40-62); None is the whole image."""

    def __init__(self, patch_sizes=(None,), patch_weights=(1.0,), eps: float = 1e-4):
        super().__init__()
        if len(patch_sizes) != len(patch_weights):
            raise ValueError("Each scale must have a weight")
        self.nccs = torch.nn.ModuleList([NormalizedCrossCorrelation2d(p, eps) for p in patch_sizes])
        self.patch_weights = [float(w) for w in patch_weights]

    def forward(self, x1: torch.Tensor, x2: torch.Tensor) -> torch.Tensor:
        total = None
        for weight, ncc in zip(self.patch_weights, self.nccs):
            term = weight * ncc(x1, x2)
            total = term if total is None else total + term
        return total


def gaussian_taps(sigma: float) -> torch.Tensor:
    """Normalised Gaussian taps over int(6 sigma + 1) | 1 points (ncc.py:91-100)."""
    k = int(6 * sigma + 1) | 1
    r = torch.arange(k, dtype=torch.float32) - (k - 1) / 2.0
    w = torch.exp(-0.5 * (r / sigma) ** 2)
    return w / w.sum()


def sobel(img: torch.Tensor, sigma: float = 1.0) -> torch.Tensor:
    """(B, 2, H, W) Sobel responses (x, then y) of (B, 1, H, W) images after a Gaussian blur of
    `sigma` (none when sigma <= 0) with mirrored borders; the Sobel stencils see zeros beyond the
    border (ncc.py:103-137). The 3 x 3 stencils are applied as their separable factors
    [1, 2, 1] (smoothing) and [1, 0, -1] (difference)."""
    x = img.float()
    if sigma > 0:
        taps = gaussian_taps(sigma).to(x.device)
        r = taps.numel() // 2
        x = F.pad(x, (r, r, r, r), mode="reflect")
        x = F.conv2d(F.conv2d(x, taps.view(1, 1, 1, -1)), taps.view(1, 1, -1, 1))
    z = F.pad(x, (1, 1, 1, 1))
    down = z[..., :-2, :] + 2.0 * z[..., 1:-1, :] + z[..., 2:, :]    # smoothing along y, (B, 1, H, W + 2)
    across = z[..., :, :-2] + 2.0 * z[..., :, 1:-1] + z[..., :, 2:]  # smoothing along x, (B, 1, H + 2, W)
    gx = down[..., :, :-2] - down[..., :, 2:]
    gy = across[..., :-2, :] - across[..., 2:, :]
    return torch.cat([gx, gy], dim=1)


class GradientNormalizedCrossCorrelation2d(NormalizedCrossCorrelation2d):
    """(B,) NCC of the Sobel gradients of two image batches (ncc.py:65-80): the mean over
    the x and y responses (and windows) of their ZNCC."""

    def __init__(self, patch_size: int | None = None, sigma: float = 1.0, **kwargs):
        super().__init__(patch_size, **kwargs)
        self.sigma = sigma

    def forward(self, x1: torch.Tensor, x2: torch.Tensor) -> torch.Tensor:
        return super().forward(sobel(x1, self.sigma), sobel(x2, self.sigma))


class DoubleGeodesicSE3(torch.nn.Module):
    """Rotational and translational distances between (B, 4, 4) poses (geo.py:7-45).

    rot = sdd / 2 * angle(R1^T R2) (the arc a rotation sweeps at half the source-to-detector
    distance); xyz = |R1^T t1 - R2^T t2|; double = sqrt(rot^2 + xyz^2); eps keeps each
    square root and the angle's arccos differentiable. Returns (rot, xyz, double)."""

    def __init__(self, sdd: float, eps: float = 1e-7):
        super().__init__()
        self.sdr = sdd / 2
        self.eps = eps

    def forward(self, pose_1: torch.Tensor, pose_2: torch.Tensor):
        rt1 = pose_1[:, :3, :3].transpose(-1, -2)
        rt2 = pose_2[:, :3, :3].transpose(-1, -2)
        d = (matmul3(rt1, pose_1[:, :3, 3:]) - matmul3(rt2, pose_2[:, :3, 3:]))[..., 0]
        xyz = torch.sqrt(torch.clamp((d * d).sum(dim=-1), min=self.eps))
        rot = self.sdr * rotmat_geodesic_distance(pose_1[:, :3, :3], pose_2[:, :3, :3], clamping=1.0 - self.eps)
        return rot, xyz, torch.sqrt(rot * rot + xyz * xyz + self.eps)


def mtre(rt_inv_est: torch.Tensor, rt_inv_true: torch.Tensor, fiducials: torch.Tensor) -> torch.Tensor:
    """(B,) mean target registration error (mm): the mean over fiducials x (world
    coordinates, (K, 3)) of |rt_est^-1 x - rt_true^-1 x|, i.e. the fiducials'
    displacement in the camera frame between the estimated and true poses."""
    est = torch.linalg.inv(rt_inv_est.double())
    true = torch.linalg.inv(rt_inv_true.double())
    xh = torch.cat([fiducials.double(), torch.ones_like(fiducials[:, :1]).double()], dim=-1)  # (K, 4)
    d = torch.einsum("bij,kj->bki", est - true, xh)[..., :3]
    return d.norm(dim=-1).mean(dim=-1)
