"""Pose-proposal CNN: a learned initialiser for the registration loop.

No counterpart in nanodrr. Six 3x3 convolutions (32 -> 256 channels, stride 2
on every second layer) on a 128 x 128 log-normalised DRR, global average
pooling and a two-layer MLP regress the 6-dof perturbation (rotation vector,
translation in mm) of `Registration`'s parameterisation about a nominal
pose. The convolutions run on the native bf16 matrix-core kernels
(ndrr_conv3x3_*: NHWC bf16 activations and weights, fp32 accumulation,
fp32 master weights); pooling, MLP and losses are small and stay in torch.
Training DRRs are rendered on the fly by the native renderer.
"""

from __future__ import annotations

import math

import torch

from . import _lib

# (Cin, Cout, stride) of the six convolutions
LAYERS = ((1, 32, 1), (32, 64, 2), (64, 64, 1), (64, 128, 2), (128, 128, 1), (128, 256, 2))


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _out(n: int, stride: int) -> int:
    return (n - 1) // stride + 1


def conv3x3(x: torch.Tensor, w: torch.Tensor, bias: torch.Tensor | None, stride: int, relu: bool) -> torch.Tensor:
    """y = relu?(conv3x3(x, w) + bias): x (N, H, W, Cin) bf16, w (Cout, 3, 3, Cin) bf16, y (N, Ho, Wo, Cout) bf16."""
    N, H, W, Cin = x.shape
    Cout = w.shape[0]
    if x.dtype != torch.bfloat16 or w.dtype != torch.bfloat16 or not x.is_cuda:
        raise ValueError("conv3x3 takes bf16 device tensors")
    if w.shape != (Cout, 3, 3, Cin):
        raise ValueError(f"w must be (Cout, 3, 3, Cin), got {tuple(w.shape)}")
    x, w = x.contiguous(), w.contiguous()
    y = torch.empty(N, _out(H, stride), _out(W, stride), Cout, device=x.device, dtype=torch.bfloat16)
    b = None if bias is None else bias.to(torch.float32).contiguous()
    rc = _lib.lib().ndrr_conv3x3_fwd(x.data_ptr(), w.data_ptr(), None if b is None else b.data_ptr(), y.data_ptr(),
                                     N, H, W, Cin, Cout, stride, int(relu), _stream(x))
    _lib.check(rc, "ndrr_conv3x3_fwd")
    return y


def conv3x3_bwd_data(dy: torch.Tensor, w: torch.Tensor, x: torch.Tensor | None, stride: int,
                     in_hw: tuple[int, int]) -> torch.Tensor:
    """dL/dx (N, H, W, Cin) bf16 of conv3x3 for dL/dy `dy`, times [x > 0] when `x` is given."""
    N, Ho, Wo, Cout = dy.shape
    _, _, _, Cin = w.shape
    H, W = in_hw
    wt = torch.empty(3, 3, Cin, Cout, device=w.device, dtype=torch.bfloat16)
    lib = _lib.lib()
    _lib.check(lib.ndrr_conv3x3_transpose_weights(w.contiguous().data_ptr(), wt.data_ptr(), Cin, Cout, _stream(w)),
               "ndrr_conv3x3_transpose_weights")
    dx = torch.empty(N, H, W, Cin, device=dy.device, dtype=torch.bfloat16)
    dy = dy.contiguous()
    rc = lib.ndrr_conv3x3_bwd_data(dy.data_ptr(), wt.data_ptr(), None if x is None else x.contiguous().data_ptr(),
                                   dx.data_ptr(), N, H, W, Cin, Cout, stride, int(x is not None), _stream(dy))
    _lib.check(rc, "ndrr_conv3x3_bwd_data")
    return dx


def conv3x3_bwd_weight(dy: torch.Tensor, x: torch.Tensor, stride: int) -> torch.Tensor:
    """dL/dw (Cout, 3, 3, Cin) fp32 of conv3x3 for dL/dy `dy` (bf16) and input `x` (bf16)."""
    N, H, W, Cin = x.shape
    Cout = dy.shape[-1]
    lib = _lib.lib()
    ws = torch.empty(lib.ndrr_conv3x3_bwd_weight_workspace_floats(N, H, W, Cin, Cout, stride), device=x.device,
                     dtype=torch.float32)
    dw = torch.empty(Cout, 3, 3, Cin, device=x.device, dtype=torch.float32)
    dy, x = dy.contiguous(), x.contiguous()
    rc = lib.ndrr_conv3x3_bwd_weight(dy.data_ptr(), x.data_ptr(), ws.data_ptr(), dw.data_ptr(), N, H, W, Cin, Cout,
                                     stride, _stream(x))
    _lib.check(rc, "ndrr_conv3x3_bwd_weight")
    return dw


class _ConvStack(torch.autograd.Function):
    """The six ReLU convolutions as one autograd node. Backward runs the chain on
    the native kernels: the ReLU mask of each layer's input is fused into the
    input-gradient kernel of the layer above it."""

    @staticmethod
    def forward(ctx, x, *params):
        ws, bs = params[0::2], params[1::2]
        acts = [x]
        wbf = [w.to(torch.bfloat16) for w in ws]
        for (cin, cout, s), w, b in zip(LAYERS, wbf, bs):
            acts.append(conv3x3(acts[-1], w, b, s, relu=True))
        ctx.save_for_backward(*acts, *wbf)
        return acts[-1]

    @staticmethod
    def backward(ctx, gy):
        saved = ctx.saved_tensors
        acts, wbf = saved[: len(LAYERS) + 1], saved[len(LAYERS) + 1:]
        dz = (gy * (acts[-1] > 0)).to(torch.bfloat16)  # through the last ReLU
        grads = [None] * (2 * len(LAYERS))
        for i in reversed(range(len(LAYERS))):
            cin, cout, s = LAYERS[i]
            x = acts[i]
            grads[2 * i] = conv3x3_bwd_weight(dz, x, s)
            grads[2 * i + 1] = dz.float().sum(dim=(0, 1, 2))
            if i > 0:
                dz = conv3x3_bwd_data(dz, wbf[i], x, s, (x.shape[1], x.shape[2]))
        return (None, *grads)


class PoseNet(torch.nn.Module):
    """(N, 1, 128, 128) DRRs -> (N, 6) = (rotation vector [rad], translation [mm] / xyz_scale)."""

    def __init__(self, xyz_scale: float = 50.0, seed: int = 0):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        self.xyz_scale = xyz_scale
        self.weights = torch.nn.ParameterList()
        self.biases = torch.nn.ParameterList()
        for cin, cout, _ in LAYERS:
            std = math.sqrt(2.0 / (9 * cin))  # He initialisation for ReLU
            self.weights.append(torch.nn.Parameter(torch.randn(cout, 3, 3, cin, generator=g) * std))
            self.biases.append(torch.nn.Parameter(torch.zeros(cout)))
        self.fc1 = torch.nn.Linear(256, 256)
        self.fc2 = torch.nn.Linear(256, 6)
        with torch.no_grad():
            self.fc2.weight.mul_(0.1)
            self.fc2.bias.zero_()

    @staticmethod
    def prepare(drr: torch.Tensor) -> torch.Tensor:
        """log(1 + DRR), standardised per image, as (N, 128, 128, 1) bf16."""
        x = torch.log1p(drr.float().clamp(min=0.0)).flatten(1)
        x = (x - x.mean(dim=1, keepdim=True)) / (x.std(dim=1, keepdim=True) + 1e-6)
        return x.view(drr.shape[0], drr.shape[-2], drr.shape[-1], 1).to(torch.bfloat16)

    def forward(self, drr: torch.Tensor) -> torch.Tensor:
        x = self.prepare(drr)
        params = [t for wb in zip(self.weights, self.biases) for t in wb]
        feat = _ConvStack.apply(x, *params)            # (N, 16, 16, 256) bf16
        h = feat.float().mean(dim=(1, 2))              # global average pool
        return self.fc2(torch.relu(self.fc1(h)))

    def split(self, out: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
        """(rotation vector, translation in mm) of the network output."""
        return out[:, :3], out[:, 3:] * self.xyz_scale
