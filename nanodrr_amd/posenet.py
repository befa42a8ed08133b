"""Pose-proposal CNN: a learned initialiser for the registration loop.

No counterpart in nanodrr. Six 3x3 convolutions (32 -> 256 channels, stride 2
on every second layer) on a 128 x 128 log-normalised DRR, 2 x 2 average
pooling and a two-layer MLP regress the 6-dof perturbation (rotation vector,
translation in mm) of `Registration`'s parameterisation about a nominal
pose. The convolutions and the head (pooling and the two-layer MLP) run on
the native bf16 matrix-core kernels (ndrr_conv3x3_*, ndrr_head_*: NHWC bf16
activations and weights, fp32 accumulation, fp32 master weights); the losses
and the optimiser (torch's fused Adam) stay in torch. Training DRRs are
rendered on the fly by the native renderer.
"""

from __future__ import annotations

import math

import torch

from . import _lib
from ._ops import _stream as _ops_stream

# (Cin, Cout, stride) of the six convolutions
LAYERS = ((1, 32, 1), (32, 64, 2), (64, 64, 1), (64, 128, 2), (128, 128, 1), (128, 256, 2))
# Volume layout of every render here (single-channel bf16 phantoms): the quad records stay in the
# Infinity Cache, and single-channel renders run fastest on them.
RENDER_LAYOUT = "quad"


def _stream(t: torch.Tensor) -> int:
    return _ops_stream(t.device)


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


def conv3x3_bwd_weight(dy: torch.Tensor, x: torch.Tensor, stride: int, bias: bool = False):
    """dL/dw (Cout, 3, 3, Cin) fp32 of conv3x3 for dL/dy `dy` (bf16) and input `x` (bf16); with
    bias=True also dL/dbias (Cout,) fp32 from the same pass, returned as (dw, db)."""
    N, H, W, Cin = x.shape
    Cout = dy.shape[-1]
    lib = _lib.lib()
    ws = torch.empty(lib.ndrr_conv3x3_bwd_weight_workspace_floats(N, H, W, Cin, Cout, stride), device=x.device,
                     dtype=torch.float32)
    dw = torch.empty(Cout, 3, 3, Cin, device=x.device, dtype=torch.float32)
    db = torch.empty(Cout, device=x.device, dtype=torch.float32) if bias else None
    dy, x = dy.contiguous(), x.contiguous()
    rc = lib.ndrr_conv3x3_bwd_weight(dy.data_ptr(), x.data_ptr(), ws.data_ptr(), dw.data_ptr(),
                                     db.data_ptr() if bias else None, N, H, W, Cin, Cout, stride, _stream(x))
    _lib.check(rc, "ndrr_conv3x3_bwd_weight")
    return (dw, db) if bias else dw


HIDDEN = 256  # units of the head's fully connected layer
OUTPUTS = 6


def head_forward(feat: torch.Tensor, w1: torch.Tensor, b1: torch.Tensor, w2: torch.Tensor, b2: torch.Tensor):
    """The head on the native kernels: feat (N, FH, FW, C) bf16 -> (out (N, O) fp32, saved) where
    saved = (H, w1t_bf16, Pt) feeds head_backward. w1 (J, K) fp32 in the pooled order
    k = (py (FW/2) + px) C + c; b1 (J,), w2 (O, J), b2 (O,) fp32."""
    if feat.dim() != 4 or feat.dtype != torch.bfloat16 or not feat.is_cuda:
        raise ValueError("head_forward takes an (N, FH, FW, C) bf16 device feature map")
    N, FH, FW, C = feat.shape
    J, K = w1.shape
    O = w2.shape[0]
    if FH % 2 or FW % 2 or C % 64 or K != (FH // 2) * (FW // 2) * C or K % 1024 or J != HIDDEN:
        raise ValueError(f"head shapes: feature map {tuple(feat.shape)} (even sides, C a multiple of 64) pools to "
                         f"{(FH // 2) * (FW // 2) * C} inputs (a multiple of 1024), w1 is {tuple(w1.shape)} "
                         f"(needs ({HIDDEN}, that))")
    if b1.shape != (J,) or w2.shape != (O, J) or b2.shape != (O,) or not 1 <= O <= 8:
        raise ValueError("b1 must be (J,), w2 (O, J) and b2 (O,) with O <= 8")
    dev = feat.device
    lib = _lib.lib()
    st = _stream(feat)
    feat = feat.contiguous()
    w1 = w1.detach().to(torch.float32).contiguous()
    w1b = torch.empty(J, K, device=dev, dtype=torch.bfloat16)
    w1t = torch.empty(K, J, device=dev, dtype=torch.bfloat16)
    _lib.check(lib.ndrr_head_cast_weights(w1.data_ptr(), w1b.data_ptr(), w1t.data_ptr(), J, K, st),
               "ndrr_head_cast_weights")
    Np = (N + 31) // 32 * 32
    P = torch.empty(N, K, device=dev, dtype=torch.bfloat16)
    Pt = torch.empty(K, Np, device=dev, dtype=torch.bfloat16)
    _lib.check(lib.ndrr_head_pool(feat.data_ptr(), P.data_ptr(), Pt.data_ptr(), N, FH, FW, C, st), "ndrr_head_pool")
    ws = torch.empty(lib.ndrr_head_workspace_floats(N, FH, FW, C, J), device=dev, dtype=torch.float32)
    H = torch.empty(N, J, device=dev, dtype=torch.float32)
    out = torch.empty(N, O, device=dev, dtype=torch.float32)
    b1, w2, b2 = (t.detach().to(torch.float32).contiguous() for t in (b1, w2, b2))
    _lib.check(lib.ndrr_head_fwd(P.data_ptr(), w1b.data_ptr(), b1.data_ptr(), w2.data_ptr(), b2.data_ptr(),
                                 ws.data_ptr(), H.data_ptr(), out.data_ptr(), N, K, J, O, st), "ndrr_head_fwd")
    return out, (H, w1t, Pt)


def head_backward(dout: torch.Tensor, feat: torch.Tensor, w2: torch.Tensor, saved):
    """(dfeat (feat's shape, bf16, times the ReLU mask of feat), dw1, db1, dw2, db2 (fp32)) of the
    head for the upstream gradient dout (N, O)."""
    H, w1t, Pt = saved
    N, FH, FW, C = feat.shape
    K, J = w1t.shape
    O = w2.shape[0]
    dev = feat.device
    Np = Pt.shape[1]
    dout = dout.to(torch.float32).contiguous()
    w2 = w2.detach().to(torch.float32).contiguous()
    dh = torch.empty(N, J, device=dev, dtype=torch.float32)
    dhb = torch.empty(N, J, device=dev, dtype=torch.bfloat16)
    dht = torch.empty(J, Np, device=dev, dtype=torch.bfloat16)
    dw1 = torch.empty(J, K, device=dev, dtype=torch.float32)
    db1 = torch.empty(J, device=dev, dtype=torch.float32)
    dw2 = torch.empty(O, J, device=dev, dtype=torch.float32)
    db2 = torch.empty(O, device=dev, dtype=torch.float32)
    dfeat = torch.empty_like(feat)
    rc = _lib.lib().ndrr_head_bwd(dout.data_ptr(), H.data_ptr(), w2.data_ptr(), w1t.data_ptr(), Pt.data_ptr(),
                                  feat.contiguous().data_ptr(), dh.data_ptr(), dhb.data_ptr(), dht.data_ptr(),
                                  dw1.data_ptr(), db1.data_ptr(), dw2.data_ptr(), db2.data_ptr(), dfeat.data_ptr(), N,
                                  FH, FW, C, J, O, _stream(feat))
    _lib.check(rc, "ndrr_head_bwd")
    return dfeat, dw1, db1, dw2, db2


class _PoseNetFn(torch.autograd.Function):
    """The whole network as one autograd node: the six ReLU convolutions, the pooling and the two
    fully connected layers, forward and backward, all on the native kernels. The head's input
    gradient comes back with the last convolution's ReLU mask applied; each convolution's
    input-gradient kernel applies the mask of the layer below."""

    @staticmethod
    def forward(ctx, x, *params):
        conv, (w1, b1, w2, b2) = params[:-4], params[-4:]
        ws, bs = conv[0::2], conv[1::2]
        acts = [x]
        wbf = [w.to(torch.bfloat16) for w in ws]
        for (cin, cout, s), w, b in zip(LAYERS, wbf, bs):
            acts.append(conv3x3(acts[-1], w, b, s, relu=True))
        out, saved = head_forward(acts[-1], w1, b1, w2, b2)
        ctx.save_for_backward(*acts, *wbf, w2, *saved)
        return out

    @staticmethod
    def backward(ctx, dout):
        saved = ctx.saved_tensors
        n = len(LAYERS)
        acts, wbf = saved[: n + 1], saved[n + 1: 2 * n + 1]
        w2, head_saved = saved[2 * n + 1], saved[2 * n + 2:]
        dz, dw1, db1, dw2, db2 = head_backward(dout, acts[-1], w2, head_saved)
        grads = [None] * (2 * n)
        for i in reversed(range(n)):
            cin, cout, s = LAYERS[i]
            x = acts[i]
            grads[2 * i], grads[2 * i + 1] = conv3x3_bwd_weight(dz, x, s, bias=True)
            if i > 0:
                dz = conv3x3_bwd_data(dz, wbf[i], x, s, (x.shape[1], x.shape[2]))
        return (None, *grads, dw1, db1, dw2, db2)


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
            grads[2 * i], grads[2 * i + 1] = conv3x3_bwd_weight(dz, x, s, bias=True)
            if i > 0:
                dz = conv3x3_bwd_data(dz, wbf[i], x, s, (x.shape[1], x.shape[2]))
        return (None, *grads)


class PoseNet(torch.nn.Module):
    """(N, 1, 128, 128) DRRs -> (N, 6) = (rotation vector [rad], translation [mm] / xyz_scale).

    The last feature map (16 x 16 x 256 for 128 x 128 inputs) is average-pooled over 2 x 2
    cells to an 8 x 8 grid and flattened into the MLP, so the head sees where features
    lie: a global average makes the features nearly invariant to in-plane shifts of the
    anatomy, i.e. to the translation the network has to estimate. The MLP's first layer
    (w1: (256, 16384), inputs ordered (cell, channel)) and its output layer are native
    bf16 matrix-core kernels with fp32 master weights, like the convolutions."""

    def __init__(self, xyz_scale: float = 50.0, seed: int = 0, pool: int = 8):
        super().__init__()
        if pool != 8:
            raise ValueError("the native head pools the 16 x 16 feature map over 2 x 2 cells: pool must be 8")
        g = torch.Generator().manual_seed(seed)
        self.xyz_scale = xyz_scale
        self.pool = pool
        self.weights = torch.nn.ParameterList()
        self.biases = torch.nn.ParameterList()
        for cin, cout, _ in LAYERS:
            std = math.sqrt(2.0 / (9 * cin))  # He initialisation for ReLU
            self.weights.append(torch.nn.Parameter(torch.randn(cout, 3, 3, cin, generator=g) * std))
            self.biases.append(torch.nn.Parameter(torch.zeros(cout)))
        k = LAYERS[-1][1] * pool * pool

        def uniform(*shape, fan_in):  # U(-1/sqrt(fan_in), 1/sqrt(fan_in)), torch.nn.Linear's default
            return (torch.rand(*shape, generator=g) * 2 - 1) / math.sqrt(fan_in)

        self.w1 = torch.nn.Parameter(uniform(HIDDEN, k, fan_in=k))
        self.b1 = torch.nn.Parameter(uniform(HIDDEN, fan_in=k))
        self.w2 = torch.nn.Parameter(0.1 * uniform(OUTPUTS, HIDDEN, fan_in=HIDDEN))
        self.b2 = torch.nn.Parameter(torch.zeros(OUTPUTS))

    @staticmethod
    def prepare(drr: torch.Tensor) -> torch.Tensor:
        """log(1 + DRR), standardised per image, as (N, 128, 128, 1) bf16."""
        x = torch.log1p(drr.float().clamp(min=0.0)).flatten(1)
        x = (x - x.mean(dim=1, keepdim=True)) / (x.std(dim=1, keepdim=True) + 1e-6)
        return x.view(drr.shape[0], drr.shape[-2], drr.shape[-1], 1).to(torch.bfloat16)

    def forward(self, drr: torch.Tensor) -> torch.Tensor:
        x = self.prepare(drr)
        params = [t for wb in zip(self.weights, self.biases) for t in wb]
        return _PoseNetFn.apply(x, *params, self.w1, self.b1, self.w2, self.b2)

    def features(self, drr: torch.Tensor) -> torch.Tensor:
        """(N, 16, 16, 256) bf16 feature map of the convolutions (no gradient)."""
        params = [t for wb in zip(self.weights, self.biases) for t in wb]
        with torch.no_grad():
            return _ConvStack.apply(self.prepare(drr), *params)

    def split(self, out: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
        """(rotation vector, translation in mm) of the network output."""
        return out[:, :3], out[:, 3:] * self.xyz_scale


# ---------------------------------------------------------------------------
# Training data, training loop and the registration comparison
# ---------------------------------------------------------------------------


class DrrSource:
    """DRRs of seeded 256^3 bf16 phantoms at poses perturbed about the nominal AP
    pose, in `Registration`'s parameterisation (rotation vector about the
    isocentre, translation in mm), rendered by the native kernels."""

    def __init__(self, seeds, size: int = 256, det: int = 128, det_mm: float = 200.0, n_samples: int = 256,
                 device="cuda:0"):
        from .phantom import detector, make_phantom, nominal_pose, phantom_subject
        from .registration import Registration

        self.regs = []
        self.n_samples = n_samples
        k_inv, sdd = detector(1020.0, det_mm / det, det, det)
        for s in seeds:
            subject = phantom_subject(make_phantom((size,) * 3, (1.0,) * 3, seed=s), 1, torch.bfloat16, device)
            rt = nominal_pose(subject).float().to(device)
            self.regs.append(Registration(subject, rt, k_inv.to(device), sdd.to(device), det, det,
                                          layout=RENDER_LAYOUT))

    @torch.no_grad()
    def render(self, k: int, rot: torch.Tensor, xyz: torch.Tensor, offset: torch.Tensor | None = None,
               n_samples: int | None = None) -> torch.Tensor:
        """(B, 1, det, det) DRRs of phantom k at parameters (rot, xyz) (B, 3)."""
        from .registration import _pose_compose_eager as pose_compose

        reg = self.regs[k]
        B = rot.shape[0]
        S = n_samples or self.n_samples
        M = pose_compose(reg._prefix.expand(B, 4, 4).contiguous(), reg._suffix.expand(B, 4, 4).contiguous(),
                         rot.float().contiguous(), xyz.float().contiguous())
        step = reg._ray_steps(S)
        from .renderer import march

        return march(reg.subject, M, reg._src, reg._tgt, step, S, reg.height, reg.width, offset=offset,
                     layout=RENDER_LAYOUT)


def sample_params(n: int, gen: torch.Generator, rot_deg: float, trans_mm: float, device):
    rot = (torch.rand(n, 3, generator=gen, device=device) * 2 - 1) * (rot_deg * math.pi / 180.0)
    xyz = (torch.rand(n, 3, generator=gen, device=device) * 2 - 1) * trans_mm
    return rot, xyz


def pose_loss(net: PoseNet, out, rot, xyz, sdd: float = 1020.0):
    """MSE on the (rotation vector, translation / xyz_scale) pair plus the mean double
    geodesic distance (angle x sdd/2 and translation, mm) / xyz_scale."""
    from .geometry import convert
    from .metrics import DoubleGeodesicSE3

    rot_hat, xyz_hat = net.split(out)
    mse = (rot_hat - rot).square().mean() + ((xyz_hat - xyz) / net.xyz_scale).square().mean()
    T_hat = convert(rot_hat.double(), xyz_hat.double(), "so3_log")
    T = convert(rot.double(), xyz.double(), "so3_log")
    geo = DoubleGeodesicSE3(sdd)(T_hat, T)[2].mean().float() / net.xyz_scale
    return mse + geo, mse.detach(), geo.detach()


def train(steps: int = 20000, batch: int = 256, seeds=tuple(range(8)), rot_deg: float = 45.0, trans_mm: float = 60.0,
          lr: float = 1e-3, seed: int = 0, device="cuda:0", log=print, log_every: int = 500, pool: int = 8):
    """Train PoseNet on DRRs rendered on the fly (each step: batch / len(seeds) poses per
    phantom, a fresh sub-step lattice offset per DRR). Returns (net, history)."""
    import time

    src = DrrSource(seeds, device=device)
    net = PoseNet(seed=seed, pool=pool).to(device)
    opt = torch.optim.Adam(net.parameters(), lr=lr, fused=True)  # one launch for all parameters
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps, eta_min=lr * 0.02)
    gen = torch.Generator(device=device).manual_seed(seed)
    per = batch // len(seeds)
    hist = []
    t0 = time.perf_counter()
    for it in range(steps):
        rot, xyz = sample_params(per * len(seeds), gen, rot_deg, trans_mm, device)
        off = torch.rand(per * len(seeds), generator=gen, device=device) - 0.5
        drr = torch.cat([src.render(k, rot[k * per:(k + 1) * per], xyz[k * per:(k + 1) * per],
                                    off[k * per:(k + 1) * per]) for k in range(len(seeds))])
        loss, mse, geo = pose_loss(net, net(drr), rot, xyz)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        sched.step()
        if (it + 1) % log_every == 0 or it == 0:
            torch.cuda.synchronize()
            h = dict(step=it + 1, loss=float(loss.detach()), mse=float(mse), geo=float(geo), s=time.perf_counter() - t0)
            hist.append(h)
            log(h)
    torch.cuda.synchronize()
    return net, hist, time.perf_counter() - t0


@torch.no_grad()
def predict(net: PoseNet, drr: torch.Tensor):
    return net.split(net(drr))


def _settled(trace, j):
    """First checked iteration from which problem j's mTRE stays below 1 mm to the end (None if it ends above)."""
    settled = None
    for k, e in trace:
        if e[j] < 1.0:
            if settled is None:
                settled = k
        else:
            settled = None
    return settled


def compare_registration(net: PoseNet, n_test: int = 64, seed0: int = 1000, rot_deg: float = 30.0,
                         trans_mm: float = 40.0, n_iters: int = 300, device="cuda:0", log=print,
                         lr_rot: float = 1e-2, lr_xyz: float = 1.0, short_iters: int | None = None,
                         refine: int = 0):
    """On n_test unseen phantoms, register one target each (200 x 200, targets at
    +-rot_deg / +-trans_mm about the nominal pose, S = 1024 target DRR) from the
    nominal pose and from the CNN's proposal (its input: the target pose rendered at
    128 x 128); both runs share one batch of two poses. Reports mTRE < 1 mm rates,
    median mTRE and final ZNCC per start, and the render iterations each start needs:
    a problem has settled at the first check (every 5 iterations) from which its mTRE
    stays below 1 mm to the end. The CNN start is also registered alone with a whole
    schedule of `short_iters` iterations (default n_iters // 5); every run uses the
    learning rates lr_rot (rad) and lr_xyz (mm).

    refine > 0 corrects the proposal p that many times by the network's own error at p:
    with q the network's estimate from a DRR rendered at p, the next proposal is 2p - q
    (rotation vector and translation), i.e. the target's error is taken to be the error at
    p. Each correction renders one 128 x 128 DRR and is counted as a render iteration of
    the CNN start."""
    short_iters = short_iters or max(n_iters // 5, 1)
    from .metrics import mtre, zncc
    from .phantom import detector, fiducials, make_phantom, nominal_pose, phantom_subject
    from .registration import Registration, register

    res = {"nominal": [], "cnn": [], "proposal_mtre": [], "raw_proposal_mtre": [], "start_mtre": [],
           "zncc": {"nominal": [], "cnn": []},
           "settled": {"nominal": [], "cnn": []}, "cnn_fifth": []}
    check_every = 5
    k200, sdd = detector(1020.0, 1.0, 200, 200)
    k200, sdd = k200.to(device), sdd.to(device)
    for i in range(n_test):
        g = torch.Generator().manual_seed(seed0 + i)
        subject = phantom_subject(make_phantom((256,) * 3, (1.0,) * 3, seed=seed0 + i), 1, torch.bfloat16, device)
        rt_nom = nominal_pose(subject).float().to(device)
        rot_t = ((torch.rand(1, 3, generator=g) * 2 - 1) * (rot_deg * math.pi / 180.0)).to(device)
        xyz_t = ((torch.rand(1, 3, generator=g) * 2 - 1) * trans_mm).to(device)
        fid = fiducials(subject, 32, g).to(device)
        # target and CNN input: the same pose through Registration's own parameterisation
        tgt_reg = Registration(subject, rt_nom, k200, sdd, 200, 200, per_pose=True, layout=RENDER_LAYOUT)
        tgt_reg._rot.data.copy_(rot_t)
        tgt_reg._xyz.data.copy_(xyz_t)
        rt_true = tgt_reg.pose.detach()
        target = tgt_reg(1024).detach()
        src128 = DrrSource.__new__(DrrSource)
        k128, sdd128 = detector(1020.0, 200.0 / 128, 128, 128)
        src128.regs = [Registration(subject, rt_nom, k128.to(device), sdd128.to(device), 128, 128,
                                    layout=RENDER_LAYOUT)]
        src128.n_samples = 512
        rot_p, xyz_p = predict(net, src128.render(0, rot_t, xyz_t))
        raw = Registration(subject, rt_nom, k200, sdd, 200, 200, per_pose=True, layout=RENDER_LAYOUT)
        raw._rot.data.copy_(rot_p)
        raw._xyz.data.copy_(xyz_p)
        res["raw_proposal_mtre"].append(float(mtre(raw.pose.detach(), rt_true, fid)[0]))
        res["start_mtre"].append(float(mtre(rt_nom.double(), rt_true, fid)[0]))
        for _ in range(refine):
            rot_q, xyz_q = predict(net, src128.render(0, rot_p, xyz_p))
            rot_p, xyz_p = 2 * rot_p - rot_q, 2 * xyz_p - xyz_q
        reg = Registration(subject, rt_nom.expand(2, 4, 4).contiguous(), k200, sdd, 200, 200, per_pose=True,
                           layout=RENDER_LAYOUT)
        reg._rot.data.copy_(torch.cat([torch.zeros_like(rot_p), rot_p]))
        reg._xyz.data.copy_(torch.cat([torch.zeros_like(xyz_p), xyz_p]))
        res["proposal_mtre"].append(float(mtre(reg.pose.detach()[1:], rt_true, fid)[0]))
        trace = []  # (iteration, mTRE of both starts) every check_every iterations

        def monitor(k, r):
            trace.append((k, mtre(r.pose.detach(), rt_true.expand(2, 4, 4), fid).tolist()))

        score = register(reg, target.expand(2, -1, -1, -1).contiguous(), n_iters=n_iters, lr_rot=lr_rot,
                         lr_xyz=lr_xyz, n_samples=256, jitter=True, seed=i, monitor=monitor,
                         monitor_every=check_every)
        for j, name in enumerate(("nominal", "cnn")):
            res["settled"][name].append(_settled(trace, j))
        # the CNN start again, with a whole schedule of a fifth of the iterations
        short = Registration(subject, rt_nom, k200, sdd, 200, 200, per_pose=True, layout=RENDER_LAYOUT)
        short._rot.data.copy_(rot_p)
        short._xyz.data.copy_(xyz_p)
        register(short, target, n_iters=short_iters, lr_rot=lr_rot, lr_xyz=lr_xyz, n_samples=256, jitter=True, seed=i)
        res["cnn_fifth"].append(float(mtre(short.pose.detach(), rt_true, fid)[0]))
        e = mtre(reg.pose.detach(), rt_true.expand(2, 4, 4), fid)
        res["nominal"].append(float(e[0]))
        res["cnn"].append(float(e[1]))
        res["zncc"]["nominal"].append(float(score[0]))
        res["zncc"]["cnn"].append(float(score[1]))
        if (i + 1) % 8 == 0:
            log(f"{i + 1}/{n_test}: nominal<1mm {sum(x < 1 for x in res['nominal'])}, "
                f"cnn<1mm {sum(x < 1 for x in res['cnn'])}, "
                f"cnn<1mm in {short_iters} iterations {sum(x < 1 for x in res['cnn_fifth'])}")
    t = torch.tensor
    curves = {name: [sum(1 for x in res["settled"][name] if x is not None and x <= k)
                     for k in range(check_every, n_iters + 1, check_every)] for name in ("nominal", "cnn")}
    plain_final = curves["nominal"][-1]
    # render iterations (the refinement renders included) after which the CNN start has as many
    # settled problems as the plain start has at the end
    cnn_match = next(((i + 1) * check_every + refine for i, c in enumerate(curves["cnn"]) if c >= plain_final),
                     None)

    def med(xs):
        xs = [x for x in xs if x is not None]
        return float(t(xs, dtype=torch.float64).median()) if xs else None

    return {
        "problems": n_test,
        "iterations": n_iters,
        "check_every": check_every,
        "settled_by_iteration": {"nominal": curves["nominal"], "cnn": curves["cnn"]},
        "median_iterations_to_1mm_nominal": med(res["settled"]["nominal"]),
        "median_iterations_to_1mm_cnn": med(res["settled"]["cnn"]),
        "cnn_iterations_to_match_plain_final": cnn_match,
        "cnn_over_plain_iteration_ratio": None if cnn_match is None else cnn_match / n_iters,
        "learning_rates": [lr_rot, lr_xyz],
        "short_iterations": short_iters,
        "refine_renders": refine,
        "render_iterations_of_the_short_cnn_run": short_iters + refine,
        "converged_from_cnn_with_the_short_schedule": int((t(res["cnn_fifth"]) < 1).sum()),
        "median_mtre_mm_cnn_with_the_short_schedule": float(t(res["cnn_fifth"]).median()),
        "converged_from_nominal": int((t(res["nominal"]) < 1).sum()),
        "converged_from_cnn": int((t(res["cnn"]) < 1).sum()),
        "median_mtre_mm_nominal": float(t(res["nominal"]).median()),
        "median_mtre_mm_cnn": float(t(res["cnn"]).median()),
        "median_mtre_mm_cnn_proposal": float(t(res["proposal_mtre"]).median()),
        "median_mtre_mm_cnn_proposal_unrefined": float(t(res["raw_proposal_mtre"]).median()),
        "median_mtre_mm_nominal_start": float(t(res["start_mtre"]).median()),
        "median_zncc_nominal": float(t(res["zncc"]["nominal"]).median()),
        "median_zncc_cnn": float(t(res["zncc"]["cnn"]).median()),
    }


def main(argv=None):
    import argparse
    import json
    import sys

    ap = argparse.ArgumentParser(description="Train the pose-proposal CNN and compare registration starts")
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--test", type=int, default=64)
    ap.add_argument("--out", default=None, help="write the summary JSON here")
    ap.add_argument("--save", default=None, help="write the trained network's state dict here")
    ap.add_argument("--load", default=None, help="skip training: load a state dict written by --save")
    ap.add_argument("--pool", type=int, default=8, help="side of the pooled feature grid fed to the MLP")
    ap.add_argument("--lr", type=float, default=1e-3, help="peak training learning rate (cosine to 2 %%)")
    ap.add_argument("--iters", type=int, default=300, help="registration iterations of the comparison")
    ap.add_argument("--short-iters", type=int, default=None, help="whole schedule of the short CNN-started run")
    ap.add_argument("--lr-scale", type=float, default=1.0, help="multiplies the 1e-2 rad / 1 mm learning rates")
    ap.add_argument("--refine", type=int, default=0,
                    help="corrections of the CNN proposal by the network's error at it (one 128 x 128 render each)")
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if args.load:
        net = PoseNet(pool=args.pool).to("cuda:0")
        net.load_state_dict(torch.load(args.load, map_location="cuda:0", weights_only=True))
        summary = {"train": {"loaded": args.load}}
    else:
        net, hist, seconds = train(args.steps, args.batch, lr=args.lr, log=log, pool=args.pool)
        summary = {"train": {"steps": args.steps, "batch": args.batch, "seconds": seconds,
                             "drr_per_s": args.steps * args.batch / seconds, "final": hist[-1]}}
    if args.save:
        torch.save(net.state_dict(), args.save)
    if args.test > 0:
        summary["registration"] = compare_registration(net, n_test=args.test, log=log, n_iters=args.iters,
                                                       lr_rot=1e-2 * args.lr_scale, lr_xyz=1.0 * args.lr_scale,
                                                       short_iters=args.short_iters, refine=args.refine)
    line = json.dumps(summary)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
