"""Image and pose metrics for registration.

`zncc` is the native zero-normalised cross-correlation (ndrr_zncc: chunk
moments over (4096-pixel chunk, image) workgroups merged in a fixed order,
deterministic) with its gradient. `NormalizedCrossCorrelation2d`, `MultiscaleNormalizedCrossCorrelation2d`,
`GradientNormalizedCrossCorrelation2d` and `DoubleGeodesicSE3` keep the API of
src/nanodrr/metrics/ncc.py:8-137 and src/nanodrr/metrics/geo.py:7-45; the
whole-image NCC runs on the native kernel, patch and gradient variants on
torch ops. `mtre` is the mean target registration error on fiducials (no
counterpart in the reference).
"""

from __future__ import annotations

import torch
import torch.nn.functional as F

from . import _lib
from ._ops import _eager
from .geometry import rotmat_geodesic_distance


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
    with torch.cuda.device(x.device):
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


def zncc(x: torch.Tensor, y: torch.Tensor, eps: float = 1e-7) -> torch.Tensor:
    """(B,) zero-normalised cross-correlation of x and y over all non-batch dims."""
    return _zncc_eager(x, y, eps)[0]


class NormalizedCrossCorrelation2d(torch.nn.Module):
    """NCC of two batches of images (ncc.py:8-37); whole images on the native kernel."""

    def __init__(self, patch_size: int | None = None, eps: float = 1e-7):
        super().__init__()
        self.patch_size = patch_size
        self.eps = eps

    def forward(self, x1: torch.Tensor, x2: torch.Tensor) -> torch.Tensor:
        if self.patch_size is not None:
            x1 = _to_patches(x1, self.patch_size)
            x2 = _to_patches(x2, self.patch_size)
        assert x1.shape == x2.shape, "Input images must be the same size"
        if self.patch_size is None and x1.shape[1] == 1:
            return zncc(x1, x2, self.eps)
        _, c, h, w = x1.shape
        x1, x2 = self._norm(x1), self._norm(x2)
        return (x1 * x2).sum(dim=(1, 2, 3)) / (c * h * w)

    def _norm(self, x):
        mu = x.mean(dim=(-1, -2), keepdim=True)
        var = x.var(dim=(-1, -2), keepdim=True, correction=0) + self.eps
        return (x - mu) / var.sqrt()


class MultiscaleNormalizedCrossCorrelation2d(torch.nn.Module):
    """Weighted sum of NCCs at several patch sizes (ncc.py:40-62)."""

    def __init__(self, patch_sizes=(None,), patch_weights=(1.0,), eps: float = 1e-4):
        super().__init__()
        assert len(patch_sizes) == len(patch_weights), "Each scale must have a weight"
        self.nccs = torch.nn.ModuleList([NormalizedCrossCorrelation2d(p, eps) for p in patch_sizes])
        self.patch_weights = list(patch_weights)

    def forward(self, x1, x2):
        return torch.stack([w * ncc(x1, x2) for w, ncc in zip(self.patch_weights, self.nccs)], dim=0).sum(dim=0)


class GradientNormalizedCrossCorrelation2d(NormalizedCrossCorrelation2d):
    """NCC of Sobel image gradients after an optional Gaussian blur (ncc.py:65-137)."""

    def __init__(self, patch_size: int | None = None, sigma: float = 1.0, **kwargs):
        super().__init__(patch_size, **kwargs)
        self.sobel = _Sobel(sigma)

    def forward(self, x1, x2):
        return super().forward(self.sobel(x1), self.sobel(x2))


def _to_patches(x, patch_size: int):
    x = x.unfold(2, patch_size, step=1).unfold(3, patch_size, step=1).contiguous()
    b, c, p1, p2, h, w = x.shape
    return x.reshape(b, c * p1 * p2, h, w)


class _Sobel(torch.nn.Module):
    def __init__(self, sigma: float):
        super().__init__()
        self.sigma = sigma
        gx = torch.tensor([[1, 0, -1], [2, 0, -2], [1, 0, -1]], dtype=torch.float32)
        gy = torch.tensor([[1, 2, 1], [0, 0, 0], [-1, -2, -1]], dtype=torch.float32)
        self.register_buffer("sobel_weight", torch.stack([gx, gy]).unsqueeze(1))
        if sigma > 0:
            k = int(6 * sigma + 1) | 1
            xs = torch.arange(k, dtype=torch.float32) - (k - 1) / 2.0
            g = torch.exp(-0.5 * (xs / sigma) ** 2)
            g = g / g.sum()
            self.register_buffer("gauss_h", g.reshape(1, 1, 1, -1))
            self.register_buffer("gauss_v", g.reshape(1, 1, -1, 1))
            self.gauss_pad = k // 2
        else:
            self.gauss_pad = 0

    def forward(self, img):
        x = img.float()
        if self.sigma > 0:
            x = F.pad(x, [self.gauss_pad] * 4, mode="reflect")
            x = F.conv2d(x, self.gauss_h)
            x = F.conv2d(x, self.gauss_v)
        return F.conv2d(x, self.sobel_weight, padding=1)


class DoubleGeodesicSE3(torch.nn.Module):
    """Angular (scaled by sdd/2) and translational geodesics between poses (geo.py:7-45)."""

    def __init__(self, sdd: float, eps: float = 1e-7):
        super().__init__()
        self.sdr = sdd / 2
        self.eps = eps

    def forward(self, pose_1, pose_2):
        r1, t1 = pose_1[:, :3, :3], pose_1[:, :3, 3]
        r2, t2 = pose_2[:, :3, :3], pose_2[:, :3, 3]
        t1 = torch.einsum("bij,bj->bi", r1.transpose(-1, -2), t1)
        t2 = torch.einsum("bij,bj->bi", r2.transpose(-1, -2), t2)
        rot = self.sdr * rotmat_geodesic_distance(r1, r2, clamping=1.0 - self.eps)
        xyz = (t1 - t2).square().sum(dim=-1).clamp(min=self.eps).sqrt()
        return rot, xyz, (rot.square() + xyz.square() + self.eps).sqrt()


def mtre(rt_inv_est: torch.Tensor, rt_inv_true: torch.Tensor, fiducials: torch.Tensor) -> torch.Tensor:
    """(B,) mean target registration error (mm): the mean over fiducials x (world
    coordinates, (K, 3)) of |rt_est^-1 x - rt_true^-1 x|, i.e. the fiducials'
    displacement in the camera frame between the estimated and true poses."""
    est = torch.linalg.inv(rt_inv_est.double())
    true = torch.linalg.inv(rt_inv_true.double())
    xh = torch.cat([fiducials.double(), torch.ones_like(fiducials[:, :1]).double()], dim=-1)  # (K, 4)
    d = torch.einsum("bij,kj->bki", est - true, xh)[..., :3]
    return d.norm(dim=-1).mean(dim=-1)
