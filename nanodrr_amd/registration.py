"""Differentiable 2D/3D registration on the native kernels.

`Registration` keeps the API of nanodrr's module
(src/nanodrr/registration/registration.py:9-84): pose =
rt_inv @ pivot @ exp(so3_log rot, xyz) @ pivot_inv with the pivot at the
isocentre in camera coordinates. With `per_pose=True` each of the B poses
gets its own (rotation, translation) parameters instead of the reference's
shared (1, 3) pair, so B independent problems run in one launch. The pose
step runs in FP64 on the GPU (ndrr_pose_compose and its adjoint) and hands
the renderer its camera -> voxel affine directly.

`register` is the batched optimisation loop: ZNCC loss (native kernel) and
Adam with per-group learning rates on a cosine schedule, every problem of the
batch advancing in the same launches. It chains the kernels' adjoints itself
(no autograd graph, no custom-op dispatch), which keeps the host's share of a
small batch's iteration low; `adjoint="autograd"` runs the same iteration
through autograd and gives the same parameters bit for bit.
"""

from __future__ import annotations

import math

import torch

from . import _lib
from ._memo import Derived
from .geometry import convert, transform_point
from ._ops import _eager, _fused_raymarch_bwd_jac_impl, _fused_raymarch_jac_impl, _on
from ._ops import _stream as _ops_stream
from .metrics import _zncc_impl, zncc
from .renderer import _make_src, _make_tgt, march, march_inputs

# register()'s loop makes the pose step and the jitter draw in one launch, and both Adam updates
# and the counter's advance in another (ndrr_pose_compose_lattice, ndrr_adam_step_scheduled_pair):
# bit for bit the separate launches, two to three fewer per iteration. False makes the separate ones.
FUSED_LOOP_LAUNCHES = True


def _stream(device):
    return _ops_stream(device)


def _pose_compose_impl(prefix: torch.Tensor, suffix: torch.Tensor, rot: torch.Tensor,
                       xyz: torch.Tensor) -> torch.Tensor:
    B = prefix.shape[0]
    for name, t, dt in (("prefix", prefix, torch.float64), ("suffix", suffix, torch.float64),
                        ("rot", rot, torch.float32), ("xyz", xyz, torch.float32)):
        if t.dtype != dt or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"{name} must be a contiguous {dt} device tensor")
    if prefix.shape != (B, 4, 4) or suffix.shape != (B, 4, 4) or rot.shape != (B, 3) or xyz.shape != (B, 3):
        raise ValueError("pose_compose expects prefix, suffix (B, 4, 4) and rot, xyz (B, 3)")
    M = torch.empty(B, 3, 4, device=prefix.device, dtype=torch.float32)
    with _on(prefix.device):
        rc = _lib.lib().ndrr_pose_compose(prefix.data_ptr(), suffix.data_ptr(), rot.data_ptr(), xyz.data_ptr(), 3,
                                          M.data_ptr(), B, _stream(prefix.device))
    _lib.check(rc, "ndrr_pose_compose")
    return M


@torch.library.custom_op("nanodrr_amd::pose_compose", mutates_args=())
def pose_compose(prefix: torch.Tensor, suffix: torch.Tensor, rot: torch.Tensor, xyz: torch.Tensor) -> torch.Tensor:
    """M (B, 3, 4) fp32 = (prefix @ [exp([rot]x), exp([rot]x) xyz; 0 1] @ suffix)[:3] in FP64."""
    return _pose_compose_impl(prefix, suffix, rot, xyz)


@pose_compose.register_fake
def _(prefix, suffix, rot, xyz):
    return torch.empty(prefix.shape[0], 3, 4, device=prefix.device, dtype=torch.float32)


def _pose_compose_bwd_impl(prefix: torch.Tensor, suffix: torch.Tensor, rot: torch.Tensor, xyz: torch.Tensor,
                           gM: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    B = prefix.shape[0]
    gM = gM.to(torch.float32).contiguous()
    grot = torch.empty(B, 3, device=prefix.device, dtype=torch.float32)
    gxyz = torch.empty(B, 3, device=prefix.device, dtype=torch.float32)
    with _on(prefix.device):
        rc = _lib.lib().ndrr_pose_compose_bwd(prefix.data_ptr(), suffix.data_ptr(), rot.data_ptr(), xyz.data_ptr(),
                                              3, gM.data_ptr(), grot.data_ptr(), gxyz.data_ptr(), B,
                                              _stream(prefix.device))
    _lib.check(rc, "ndrr_pose_compose_bwd")
    return grot, gxyz


@torch.library.custom_op("nanodrr_amd::pose_compose_bwd", mutates_args=())
def pose_compose_bwd(prefix: torch.Tensor, suffix: torch.Tensor, rot: torch.Tensor, xyz: torch.Tensor,
                     gM: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    return _pose_compose_bwd_impl(prefix, suffix, rot, xyz, gM)


@pose_compose_bwd.register_fake
def _(prefix, suffix, rot, xyz, gM):
    B = prefix.shape[0]
    return (torch.empty(B, 3, device=prefix.device, dtype=torch.float32),
            torch.empty(B, 3, device=prefix.device, dtype=torch.float32))


def _pc_setup(ctx, inputs, output):
    prefix, suffix, rot, xyz = inputs
    ctx.save_for_backward(prefix, suffix, rot, xyz)


def _pc_backward(ctx, gM):
    prefix, suffix, rot, xyz = ctx.saved_tensors
    grot, gxyz = pose_compose_bwd(prefix, suffix, rot, xyz, gM)
    return None, None, grot, gxyz


torch.library.register_autograd("nanodrr_amd::pose_compose", _pc_backward, setup_context=_pc_setup)

# eager callers skip the custom-op dispatch (see nanodrr_amd._ops._eager)
_pose_compose_eager = _eager(
    "pose_compose", pose_compose, _pose_compose_impl, _pc_setup,
    lambda ctx, gM: (None, None, *_pose_compose_bwd_impl(*ctx.saved_tensors, gM)))


def _translation(t: torch.Tensor) -> torch.Tensor:
    T = torch.eye(4, device=t.device, dtype=t.dtype).repeat(len(t), 1, 1)
    T[:, :3, 3] = t
    return T


class Registration(torch.nn.Module):
    """Optimise camera poses by aligning rendered DRRs with target images.

    Args match the reference (registration.py:32-57); `per_pose` gives every
    pose its own parameters (B independent problems).
    """

    def __init__(self, subject, rt_inv, k_inv, sdd, height, width, eps: float = 1e-8, per_pose: bool = False,
                 layout: str | None = None):
        super().__init__()
        self.subject = subject
        self.layout = layout  # volume layout of the renders (None: renderer.VOLUME_LAYOUT)
        self.register_buffer("rt_inv", rt_inv)
        self.register_buffer("k_inv", k_inv)
        self.register_buffer("sdd", sdd)
        self.height = height
        self.width = width
        iso = subject.isocenter.to(rt_inv.device, torch.float64).expand(len(rt_inv), 3)
        c = transform_point(torch.linalg.inv(rt_inv.double()), iso)  # isocentre in camera coordinates
        self.register_buffer("pivot", _translation(c))
        self.register_buffer("pivot_inv", _translation(-c))
        n = len(rt_inv) if per_pose else 1
        self._rot = torch.nn.Parameter(eps * torch.randn(n, 3, device=rt_inv.device))
        self._xyz = torch.nn.Parameter(eps * torch.randn(n, 3, device=rt_inv.device))
        self._refresh_prefix()

    def _refresh_prefix(self):
        """FP64 constants of the pose step: world_to_voxel @ rt_inv @ pivot and pivot_inv."""
        w2v = self.subject.world_to_voxel.double()
        self.register_buffer("_prefix", (w2v @ self.rt_inv.double() @ self.pivot).contiguous(), persistent=False)
        self.register_buffer("_suffix", self.pivot_inv.contiguous(), persistent=False)
        tgt = _make_tgt(self.k_inv, self.sdd, self.height, self.width, self.k_inv.device, torch.float32)
        src = _make_src(False, tgt, self.sdd)
        self.register_buffer("_tgt", tgt.contiguous(), persistent=False)
        self.register_buffer("_src", src.contiguous(), persistent=False)

    def _params(self):
        B = len(self.rt_inv)
        rot, xyz = self._rot, self._xyz
        if rot.shape[0] != B:
            rot, xyz = rot.expand(B, 3), xyz.expand(B, 3)
        return rot.float().contiguous(), xyz.float().contiguous()

    def camera_to_voxel(self) -> torch.Tensor:
        """(B, 3, 4) fp32 camera -> voxel affine of the current pose (FP64 pose step)."""
        rot, xyz = self._params()
        return _pose_compose_eager(self._prefix, self._suffix, rot, xyz)

    def forward(self, n_samples: int = 500, backend: str = "auto", offset: torch.Tensor | None = None) -> torch.Tensor:
        """DRRs of the current poses (registration.py:59-62). `offset` (B,), in units of the
        sample step, slides each pose's sample lattice along its rays (sample s at
        t = (s + offset) / (n_samples - 1), the step unchanged; the march kernels take it
        directly); None renders exactly as `render`."""
        if backend not in ("auto", "hip", "triton"):
            raise ValueError(f"Unknown backend {backend!r}; expected 'auto', 'hip' or 'triton'")
        step = self._ray_steps(n_samples)
        return march(self.subject, self.camera_to_voxel(), self._src, self._tgt, step, n_samples, self.height,
                     self.width, offset=offset, layout=self.layout)

    def _ray_steps(self, n_samples: int) -> torch.Tensor:
        """The ray step (1, N) for n_samples, kept until the rays change (the same tensors at the
        same versions: a replaced or edited ray buffer is seen)."""
        def make():
            return (self._tgt - self._src).norm(dim=-1) / float(n_samples - 1)

        if torch.compiler.is_compiling():
            return make()
        if not hasattr(self, "_derived"):
            self._derived = Derived()
        inner = self._derived.get("tgt", self._tgt, None, Derived)
        return inner.get(f"src:{n_samples}", self._src, None, make)

    @property
    def pose(self) -> torch.Tensor:
        """(B, 4, 4) camera-to-world pose (registration.py:64-68), float64."""
        rot, xyz = self._params()
        T = convert(rot.double(), xyz.double(), "so3_log")
        return self.rt_inv.double() @ self.pivot @ T @ self.pivot_inv

    @torch.no_grad()
    def rescale_(self, scale: float) -> None:
        """Change the rendering resolution by rescaling the intrinsics (registration.py:70-76)."""
        self.k_inv[..., 0, 0] /= scale
        self.k_inv[..., 1, 1] /= scale
        self.height = int(round(self.height * scale))
        self.width = int(round(self.width * scale))
        self._refresh_prefix()


class _Adam:
    """Adam with register()'s cosine schedule in torch operations, for CPU tensors:
    the statement of the update that register() makes with ndrr_adam_step."""

    def __init__(self, groups, n_steps: int, betas=(0.9, 0.999), eps=1e-8, final_frac: float = 0.02):
        self.groups = groups  # [(param, lr)]
        self.b1, self.b2, self.eps = betas[0], betas[1], eps
        dev = groups[0][0].device
        k = torch.arange(n_steps, dtype=torch.float64)
        cos = final_frac + (1 - final_frac) * 0.5 * (1 + torch.cos(math.pi * k / max(n_steps - 1, 1)))
        self.sched = cos.to(torch.float32).to(dev)
        self.t = torch.zeros((), device=dev, dtype=torch.float32)
        self.idx = torch.zeros((), device=dev, dtype=torch.int64)
        self.m = [torch.zeros_like(p) for p, _ in groups]
        self.v = [torch.zeros_like(p) for p, _ in groups]

    def state(self) -> list[torch.Tensor]:
        return [self.t, self.idx, *self.m, *self.v]

    @torch.no_grad()
    def step(self):
        self.t.add_(1.0)
        frac = self.sched.index_select(0, self.idx.clamp(max=self.sched.numel() - 1).view(1))
        self.idx.add_(1)
        bc1 = 1.0 - torch.pow(self.b1, self.t)
        bc2 = 1.0 - torch.pow(self.b2, self.t)
        for (p, lr), m, v in zip(self.groups, self.m, self.v):
            g = p.grad
            m.mul_(self.b1).add_(g, alpha=1.0 - self.b1)
            v.mul_(self.b2).addcmul_(g, g, value=1.0 - self.b2)
            p.sub_((lr * frac) * (m / bc1) / ((v / bc2).sqrt() + self.eps))


def _adam_step(p: torch.Tensor, g: torch.Tensor, m: torch.Tensor, v: torch.Tensor, lr: float, b1: float, b2: float,
               eps: float, t: int) -> None:
    """One native Adam step (ndrr_adam_step) on a contiguous fp32 device parameter."""
    g = g.contiguous()
    with _on(p.device):
        rc = _lib.lib().ndrr_adam_step(p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), p.numel(), lr, b1, b2,
                                       eps, 1.0 - b1 ** t, 1.0 - b2 ** t, _stream(p.device))
    _lib.check(rc, "ndrr_adam_step")


class _ExplicitLoop:
    """The explicit adjoint chain of one register() call. Inputs are checked and every buffer
    is allocated once; each iteration then makes the launches that autograd would (pose step,
    march with the single-pass sums, ZNCC, contraction, pose adjoint), in the same order and
    with the same arguments, through the C ABI with cached pointers."""

    def __init__(self, reg: Registration, target: torch.Tensor, n_samples: int, eps: float, jitter: bool):
        from ._ops import _DTYPE_CODE, _geometry, _ptr, _volume_arg

        self.reg, self.eps, self.S = reg, float(eps), n_samples
        dev = target.device
        self.device = dev
        self.lib = _lib.lib()
        B = len(reg.rt_inv)
        self.B = B
        self.shared = reg._rot.shape[0] != B
        self.rot_b = torch.empty(B, 3, device=dev) if self.shared else reg._rot.detach()
        self.xyz_b = torch.empty(B, 3, device=dev) if self.shared else reg._xyz.detach()
        self.M = torch.empty(B, 3, 4, device=dev)
        self.step = reg._ray_steps(n_samples)
        src, tgt = reg._src, reg._tgt
        self.src, self.tgt = src, tgt
        args, packed, layout, _, C = march_inputs(reg.subject, self.M, src, tgt, self.step, n_samples, reg.height,
                                                  reg.width, reg.layout)
        vol, lab = args[0], args[1]
        for a, b in ((args[2], self.M), (args[3], src), (args[4], tgt), (args[5], self.step)):
            if a.data_ptr() != b.data_ptr():
                raise ValueError("register() needs contiguous float32 rays and step")
        _, N, D, H, W, self.strides = _geometry(*args, layout if packed is not None else "linear")
        self.vptr, self.code = _volume_arg(vol, packed, layout)
        self.dt, self.lab = _DTYPE_CODE[vol.dtype], (_ptr(lab) if C > 1 else None)
        self.dims = (B, N, D, H, W, n_samples, C, reg.width)
        self.target = target.reshape(B, -1)
        if self.target.shape[1] != C * N:
            raise ValueError("Input images must be the same size")
        self.img = torch.empty(B, C, N, device=dev)
        self.jac = torch.empty(B, 6, C, N, device=dev)
        self.score = torch.empty(B, device=dev)
        self.gimg = torch.empty(B, C, N, device=dev)
        self.zws = torch.empty(self.lib.ndrr_zncc_workspace_floats(B, C * N), device=dev)
        self.bws = torch.empty(self.lib.ndrr_bwd_workspace_floats(B, N, reg.width), device=dev)
        self.grot = torch.empty(B, 3, device=dev)
        self.gxyz = torch.empty(B, 3, device=dev)

    def step_(self, offset, jitter=None):
        """One iteration on torch's current stream (the capturing stream under HIP-graph capture):
        returns (score (B,), d(-sum score)/d rot, d(-sum score)/d xyz). With jitter = (counter,
        seed), the pose step also draws the lattice offsets into `offset` (one launch for both)."""
        reg, lib, st = self.reg, self.lib, _stream(self.device)
        B, N, D, H, W, S, C, width = self.dims
        if offset is not None and (offset.dtype != torch.float32 or offset.shape != (B,) or not offset.is_contiguous()):
            raise ValueError(f"offset must be a contiguous float32 ({B},) tensor")
        lattice = None if offset is None else offset.data_ptr()
        if self.shared:  # the expanded parameters, as _params() hands them to the pose step
            self.rot_b.copy_(reg._rot.detach().expand(B, 3))
            self.xyz_b.copy_(reg._xyz.detach().expand(B, 3))
        pre, suf = reg._prefix.data_ptr(), reg._suffix.data_ptr()
        rot, xyz, M = self.rot_b.data_ptr(), self.xyz_b.data_ptr(), self.M.data_ptr()
        src, tgt, step = self.src.data_ptr(), self.tgt.data_ptr(), self.step.data_ptr()
        if jitter is not None:
            counter, seed32 = jitter
            _lib.check(lib.ndrr_pose_compose_lattice(pre, suf, rot, xyz, 3, M, B, counter.data_ptr(), seed32,
                                                     offset.data_ptr(), st), "ndrr_pose_compose_lattice")
        else:
            _lib.check(lib.ndrr_pose_compose(pre, suf, rot, xyz, 3, M, B, st), "ndrr_pose_compose")
        _lib.check(lib.ndrr_raymarch_fwd_jac(self.vptr, self.dt, self.code, self.lab, M, src, tgt, step, lattice,
                                             self.img.data_ptr(), self.jac.data_ptr(), 6, B, N, D, H, W, S, C,
                                             width, *self.strides, st), "ndrr_raymarch_fwd_jac")
        _lib.check(lib.ndrr_zncc(self.img.data_ptr(), self.target.data_ptr(), B, C * N, self.eps,
                                 self.zws.data_ptr(), self.score.data_ptr(), self.gimg.data_ptr(), st), "ndrr_zncc")
        # d(-sum score)/d img = -gimg; the contraction and the pose adjoint are linear and
        # sign-symmetric, so scaling the reduced gM by -1 (inside the pose reduction, which also
        # runs the pose step's adjoint) gives the bits that -gimg would
        _lib.check(lib.ndrr_raymarch_bwd_jac_se3(self.gimg.data_ptr(), C * N, self.jac.data_ptr(), 6, M, src, tgt,
                                                 step, self.bws.data_ptr(), pre, suf, rot, xyz, 3, -1.0,
                                                 self.grot.data_ptr(), self.gxyz.data_ptr(), B, N, C, width,
                                                 *self.strides, st), "ndrr_raymarch_bwd_jac_se3")
        if self.shared:  # the expand's adjoint
            return self.score, self.grot.sum_to_size(reg._rot.shape), self.gxyz.sum_to_size(reg._xyz.shape)
        return self.score, self.grot, self.gxyz


def _schedule(n_iters: int, lr_rot: float, lr_xyz: float, b1: float, b2: float, final_frac: float,
              device) -> torch.Tensor:
    """(n_iters, 4) fp32 rows [lr_rot f, lr_xyz f, 1 - b1^t, 1 - b2^t] of register()'s Adam schedule
    (f the cosine factor down to final_frac, t = iteration + 1), computed in float64."""
    k = torch.arange(n_iters, dtype=torch.float64)
    f = final_frac + (1 - final_frac) * 0.5 * (1 + torch.cos(math.pi * k / max(n_iters - 1, 1)))
    t = k + 1
    rows = torch.stack([lr_rot * f, lr_xyz * f, 1.0 - b1 ** t, 1.0 - b2 ** t], dim=1)
    return rows.to(torch.float32).to(device).contiguous()


def register(reg: Registration, target: torch.Tensor, n_iters: int = 300, lr_rot: float = 1e-2, lr_xyz: float = 1.0,
             n_samples: int = 256, jitter: bool = False, seed: int = 0, graph: bool = False,
             iters_per_graph: int = 10, eps: float = 1e-7, adjoint: str = "explicit", monitor=None,
             monitor_every: int = 10) -> torch.Tensor:
    """Optimise every pose of `reg` to maximise the ZNCC of its DRRs with `target` (B, 1, H, W).

    Each iteration: pose step, render, ZNCC, backward, Adam update (learning
    rates lr_rot rad and lr_xyz mm on a cosine schedule down to 2 %, one native
    launch per parameter tensor); every problem of the batch advances in the
    same launches. Returns the ZNCC of the last iteration's DRRs (B,).

    A DRR sampled every `step` mm along rays fixed to the camera repeats itself
    when the camera moves one step along its axis, so the similarity to a target
    has local maxima spaced one step apart in depth. `jitter` slides each
    problem's sample lattice by a fresh uniform offset in [-1/2, 1/2) step every
    iteration, drawn from a hash of (seed, iteration, problem), which averages
    that ripple out.

    `adjoint="explicit"` (default) calls the kernels and their adjoints
    directly; `adjoint="autograd"` builds the same iteration with autograd
    through the custom ops (the path any other loss would take).

    The explicit loop keeps its state on the device (iteration counter, jitter
    offsets, the Adam schedule as a table), so `graph=True` can capture
    `iters_per_graph` iterations in one HIP graph and replay it: the first
    (n_iters - 1) % iters_per_graph + 1 iterations run eagerly (every kernel
    has then run once), the rest as replays. A captured iteration is the same
    launches as an eager one, on the capturing stream, with no allocation,
    memset or host synchronisation, so the results are bitwise the same.
    Graph mode needs per-problem parameters (per_pose=True).

    `monitor(iteration, reg)`, when given, is called after every
    `monitor_every` iterations (in graph mode: after the eager iterations and
    after every replay) and after the last one, with the parameters of that
    iteration applied (e.g. to record the mTRE trajectory).
    """
    if adjoint not in ("explicit", "autograd"):
        raise ValueError(f"adjoint must be 'explicit' or 'autograd', got {adjoint!r}")
    if graph and adjoint != "explicit":
        raise ValueError("graph=True captures the explicit loop: use adjoint='explicit'")
    if graph and iters_per_graph < 1:
        raise ValueError("iters_per_graph must be at least 1")
    params = [(reg._rot, lr_rot), (reg._xyz, lr_xyz)]
    for p, _ in params:
        if not (p.is_cuda and p.dtype == torch.float32 and p.is_contiguous()):
            raise ValueError("register() updates contiguous float32 device parameters")
    if graph and reg._rot.shape[0] != len(reg.rt_inv):
        raise ValueError("graph=True needs per-problem parameters (Registration(..., per_pose=True))")
    dev = target.device
    moments = [(torch.zeros_like(p), torch.zeros_like(p)) for p, _ in params]
    b1, b2, adam_eps, final_frac = 0.9, 0.999, 1e-8, 0.02
    target = target.detach().to(torch.float32).contiguous()
    B = target.shape[0]
    score = torch.zeros(B, device=dev)  # n_iters == 0
    schedule = _schedule(max(n_iters, 1), lr_rot, lr_xyz, b1, b2, final_frac, dev)
    counter = torch.zeros(1, dtype=torch.int32, device=dev)
    offsets = torch.empty(B, device=dev) if jitter else None
    lib = _lib.lib()
    seed32 = int(seed) & 0xFFFFFFFF

    def draw_offsets():
        if jitter:
            _lib.check(lib.ndrr_lattice_offsets(counter.data_ptr(), seed32, offsets.data_ptr(), B, _stream(dev)),
                       "ndrr_lattice_offsets")
        return offsets

    fused = FUSED_LOOP_LAUNCHES and params[0][0].numel() == params[1][0].numel()

    def adam(grads):
        st = _stream(dev)
        if fused:
            (p0, _), (p1, _) = params
            (m0, v0), (m1, v1) = moments
            _lib.check(lib.ndrr_adam_step_scheduled_pair(p0.data_ptr(), grads[0].data_ptr(), m0.data_ptr(),
                                                         v0.data_ptr(), p1.data_ptr(), grads[1].data_ptr(),
                                                         m1.data_ptr(), v1.data_ptr(), p0.numel(),
                                                         schedule.data_ptr(), schedule.shape[0], counter.data_ptr(),
                                                         b1, b2, adam_eps, st), "ndrr_adam_step_scheduled_pair")
            return
        for col, ((p, _), (m, v), g) in enumerate(zip(params, moments, grads)):
            _lib.check(lib.ndrr_adam_step_scheduled(p.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(),
                                                    p.numel(), schedule.data_ptr(), schedule.shape[0], col,
                                                    counter.data_ptr(), b1, b2, adam_eps, st),
                       "ndrr_adam_step_scheduled")
        _lib.check(lib.ndrr_counter_advance(counter.data_ptr(), st), "ndrr_counter_advance")

    def report(it):
        if monitor is not None:
            monitor(it, reg)

    if adjoint == "explicit":
        with torch.no_grad(), torch.cuda.device(dev):
            loop = _ExplicitLoop(reg, target, n_samples, eps, jitter)

            def iteration():
                if fused and jitter:  # the offsets are drawn by the pose step's launch
                    s, grot, gxyz = loop.step_(offsets, jitter=(counter, seed32))
                else:
                    s, grot, gxyz = loop.step_(draw_offsets())
                adam((grot, gxyz))
                return s, grot, gxyz

            head = n_iters if not graph else (n_iters - 1) % iters_per_graph + 1 if n_iters else 0
            for it in range(head):
                score, grot, gxyz = iteration()
                if (it + 1) % monitor_every == 0 or it + 1 == head:
                    report(it + 1)
            if n_iters > head:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    for _ in range(iters_per_graph):
                        score, grot, gxyz = iteration()
                done = head
                while done < n_iters:
                    g.replay()
                    done += iters_per_graph
                    report(done)
            if n_iters:
                reg._rot.grad, reg._xyz.grad = grot.clone(), gxyz.clone()
        return score.clone()
    descent = torch.full((B,), -1.0, device=dev)  # d(-sum of scores) / d score
    for it in range(n_iters):
        for p, _ in params:
            p.grad = None
        offset = draw_offsets().clone() if jitter else None
        s = zncc(reg(n_samples, offset=offset), target, eps)
        s.backward(descent)
        with torch.no_grad():
            adam([p.grad.contiguous() for p, _ in params])
        score = s.detach()
        if (it + 1) % monitor_every == 0 or it + 1 == n_iters:
            report(it + 1)
    return score.clone()
