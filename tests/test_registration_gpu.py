"""GPU parity of the registration row: the FP64 pose step and its adjoint,
ZNCC and its gradient against the oracle, the Registration module against
the reference's pose form (src/nanodrr/registration/registration.py:59-68),
HIP-graph replay against eager iterations, and convergence on a phantom."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import registration as oreg

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _problem(B=64, seed=0, scale=0.5):
    rng = np.random.default_rng(seed)
    prefix = np.tile(np.eye(4), (B, 1, 1))
    suffix = np.tile(np.eye(4), (B, 1, 1))
    prefix[:, :3, :] = rng.normal(size=(B, 3, 4)) * 50
    from scipy.spatial.transform import Rotation

    suffix[:, :3, :3] = Rotation.random(B, random_state=seed).as_matrix()
    suffix[:, :3, 3] = rng.normal(size=(B, 3)) * 20
    rot = (rng.normal(size=(B, 3)) * scale).astype(np.float32)
    rot[0] = 0.0
    rot[1] = [1e-9, -2e-9, 5e-10]
    rot[2] = [3e-4, 0.0, -2e-4]  # series / closed-form boundary region
    xyz = (rng.normal(size=(B, 3)) * 10).astype(np.float32)
    return prefix, suffix, rot, xyz


def _dev(*arrays):
    return [torch.from_numpy(np.ascontiguousarray(a)).to(DEV) for a in arrays]


def test_pose_compose_matches_oracle():
    from nanodrr_amd.registration import pose_compose

    prefix, suffix, rot, xyz = _problem()
    M = pose_compose(*_dev(prefix, suffix, rot, xyz)).cpu().double().numpy()
    ref = oreg.pose_compose(prefix, suffix, rot.astype(np.float64), xyz.astype(np.float64))
    # computed in FP64 and rounded once to fp32: within one fp32 ulp of each entry
    np.testing.assert_allclose(M, ref, rtol=6e-8, atol=1e-12 * np.abs(ref).max())


def test_pose_compose_bwd_matches_oracle():
    from nanodrr_amd.registration import pose_compose_bwd

    prefix, suffix, rot, xyz = _problem(seed=1, scale=1.0)
    gM = np.random.default_rng(5).normal(size=(len(rot), 3, 4)).astype(np.float32)
    grot, gxyz = pose_compose_bwd(*_dev(prefix, suffix, rot, xyz, gM))
    ref_rot, ref_xyz = oreg.pose_compose_bwd(prefix, suffix, rot.astype(np.float64), xyz.astype(np.float64),
                                             gM.astype(np.float64))
    scale_r = np.abs(ref_rot).max()
    scale_x = np.abs(ref_xyz).max()
    np.testing.assert_allclose(grot.cpu().double().numpy(), ref_rot, rtol=1e-6, atol=1e-7 * scale_r)
    np.testing.assert_allclose(gxyz.cpu().double().numpy(), ref_xyz, rtol=1e-6, atol=1e-7 * scale_x)


def test_pose_compose_autograd():
    from nanodrr_amd.registration import pose_compose

    prefix, suffix, rot, xyz = _problem(B=8, seed=2)
    p, s, r, v = _dev(prefix, suffix, rot, xyz)
    r.requires_grad_(True)
    v.requires_grad_(True)
    gM = torch.randn(8, 3, 4, device=DEV, generator=torch.Generator(DEV).manual_seed(0))
    (pose_compose(p, s, r, v) * gM).sum().backward()
    ref_rot, ref_xyz = oreg.pose_compose_bwd(prefix, suffix, rot.astype(np.float64), xyz.astype(np.float64),
                                             gM.cpu().double().numpy())
    np.testing.assert_allclose(r.grad.cpu().double().numpy(), ref_rot, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(v.grad.cpu().double().numpy(), ref_xyz, rtol=1e-6, atol=1e-6)


# one chunk (7 x 13, 300, 64 x 64 = exactly 4096 pixels), ten chunks (200 x 200), and 256 chunks
# (1024 x 1024: more chunks than lanes in the merging wave)
@pytest.mark.parametrize("shape", [(3, 1, 200, 200), (2, 1, 7, 13), (1, 1, 1, 300), (5, 1, 64, 64),
                                   (2, 1, 1024, 1024)])
def test_zncc_matches_oracle(shape):
    from nanodrr_amd.metrics import zncc

    g = torch.Generator().manual_seed(shape[-1])
    x = torch.rand(shape, generator=g) * 3
    y = 0.5 * x + torch.rand(shape, generator=g)
    xd = x.to(DEV).requires_grad_(True)
    s = zncc(xd, y.to(DEV), 1e-7)
    s.sum().backward()
    ref_s, ref_g = oreg.zncc(x.double().numpy(), y.double().numpy(), 1e-7)
    np.testing.assert_allclose(s.detach().cpu().double().numpy(), ref_s, atol=2e-6)
    got_g = xd.grad.cpu().double().numpy().reshape(len(ref_g), -1)
    assert np.abs(got_g - ref_g).max() <= 1e-4 * np.abs(ref_g).max()


def test_zncc_deterministic_and_matches_reference_module():
    from nanodrr_amd.metrics import NormalizedCrossCorrelation2d, zncc

    g = torch.Generator().manual_seed(3)
    x = torch.rand(4, 1, 64, 48, generator=g).to(DEV)
    y = torch.rand(4, 1, 64, 48, generator=g).to(DEV) + x
    a, b = zncc(x, y), zncc(x, y)
    assert torch.equal(a, b)
    # ncc.py:8-37 (population variance, eps in the square root), as the float64 oracle states it
    mod = NormalizedCrossCorrelation2d(patch_size=None, eps=1e-7)
    ref = torch.from_numpy(oreg.zncc(x.cpu().numpy(), y.cpu().numpy(), 1e-7)[0])
    torch.testing.assert_close(mod(x, y).double().cpu(), ref, rtol=0, atol=2e-6)


def _phantom_registration(B=4, size=64, spacing=4.0, h=64, w=64, seed=0, per_pose=True):
    from nanodrr_amd.phantom import make_phantom, phantom_subject, registration_problems
    from nanodrr_amd.registration import Registration

    ph = make_phantom((size, size, size), (spacing, spacing, spacing), seed=seed)
    subject = phantom_subject(ph, 1, torch.float32, DEV)
    prob = registration_problems(subject, B, seed, height=h, width=w, sdd=1020.0)
    # 1 mm detector pixels at 64 x 64 see little; use 4 mm pixels so the phantom fills the view
    prob.k_inv[..., 0, 0] *= 4
    prob.k_inv[..., 1, 1] *= 4
    prob.k_inv[..., :2, 2] *= 4
    torch.manual_seed(seed)
    reg = Registration(subject, prob.rt_init.float().to(DEV), prob.k_inv.float().to(DEV),
                       prob.sdd.float().to(DEV), h, w, per_pose=per_pose)
    return subject, prob, reg


def test_registration_forward_matches_render_of_pose():
    """Registration.forward == render(subject, k_inv, pose) (registration.py:59-68)."""
    from nanodrr_amd import render

    subject, prob, reg = _phantom_registration(B=3)
    with torch.no_grad():
        reg._rot.copy_(torch.tensor([[0.05, -0.02, 0.1], [0.0, 0.0, 0.0], [-0.2, 0.1, 0.03]]))
        reg._xyz.copy_(torch.tensor([[3.0, -2.0, 1.0], [0.0, 0.0, 0.0], [-5.0, 4.0, 2.0]]))
        img = reg(128)
        ref = render(subject, reg.k_inv, reg.pose.float(), reg.sdd, reg.height, reg.width, 128)
    err = (img - ref).abs().max() / ref.abs().max()
    assert err < 1e-5, err


def test_registration_gradients_match_reference_pose_form():
    """d(-zncc)/d(rot, xyz) through the FP64 pose kernel == autograd through the
    reference's pose form (convert(..., "so3_log") in float64, registration.py:64-68)."""
    from nanodrr_amd import render
    from nanodrr_amd.geometry import convert
    from nanodrr_amd.metrics import zncc

    subject, prob, reg = _phantom_registration(B=3, seed=1)
    target = render(subject, reg.k_inv, prob.rt_true.float().to(DEV), reg.sdd, reg.height, reg.width, 128)
    with torch.no_grad():
        reg._rot.copy_(torch.tensor([[0.05, -0.02, 0.1], [0.01, 0.0, 0.0], [-0.2, 0.1, 0.03]]))
        reg._xyz.copy_(torch.tensor([[3.0, -2.0, 1.0], [0.0, 0.5, 0.0], [-5.0, 4.0, 2.0]]))
    (-zncc(reg(128), target).sum()).backward()
    rot = reg._rot.detach().double().clone().requires_grad_(True)
    xyz = reg._xyz.detach().double().clone().requires_grad_(True)
    pose = reg.rt_inv.double() @ reg.pivot @ convert(rot, xyz, "so3_log") @ reg.pivot_inv
    img = render(subject, reg.k_inv, pose.float(), reg.sdd, reg.height, reg.width, 128)
    (-zncc(img, target).sum()).backward()
    for got, ref in ((reg._rot.grad, rot.grad), (reg._xyz.grad, xyz.grad)):
        err = (got.double() - ref).abs().max() / ref.abs().max()
        assert err < 1e-4, err


def test_shared_parameters_sum_over_batch():
    """per_pose=False keeps the reference's shared (1, 3) parameters (registration.py:55-57)."""
    from nanodrr_amd.metrics import zncc

    subject, prob, reg = _phantom_registration(B=3, per_pose=False, seed=2)
    assert reg._rot.shape == (1, 3)
    target = torch.rand(3, 1, reg.height, reg.width, device=DEV)
    (-zncc(reg(64), target).sum()).backward()
    shared = reg._rot.grad.clone()
    _, _, per = _phantom_registration(B=3, per_pose=True, seed=2)
    with torch.no_grad():
        per._rot.copy_(reg._rot.expand(3, 3))
        per._xyz.copy_(reg._xyz.expand(3, 3))
    (-zncc(per(64), target).sum()).backward()
    torch.testing.assert_close(shared[0], per._rot.grad.sum(0), rtol=1e-5, atol=1e-6)


class _RecordingLib:
    """The C ABI with every call recorded: (name, whether torch's current stream was capturing, the
    stream argument, torch's current stream, bytes torch had allocated)."""

    def __init__(self, lib):
        self._lib, self.calls = lib, []

    def __getattr__(self, name):
        fn = getattr(self._lib, name)

        def call(*args):
            capturing = torch.cuda.is_current_stream_capturing()
            self.calls.append((name, capturing, args[-1] if args else None, torch.cuda.current_stream().cuda_stream,
                               torch.cuda.memory_allocated()))
            return fn(*args)

        return call


@pytest.mark.parametrize("fused", [True, False])
def test_graph_replay_matches_eager_iterations(monkeypatch, fused):
    """register(graph=True): 11 iterations as 2 eager ones and three replays of a 3-iteration
    graph, against 11 eager iterations of the same problems. While capturing, every library call is
    a launch on the capturing stream and torch allocates nothing; after each replay (with eager work
    and a host synchronisation in between: the monitor) the parameters equal the eager loop's at the
    same iteration, bit for bit, and so does the final score. With and without the fused loop-state
    launches (pose step + jitter draw, both Adam updates + the counter's advance)."""
    from nanodrr_amd import _lib, render
    from nanodrr_amd import registration as registration_mod
    from nanodrr_amd.registration import register

    monkeypatch.setattr(registration_mod, "FUSED_LOOP_LAUNCHES", fused)
    subject, prob, reg_g = _phantom_registration(B=4, seed=3)
    _, _, reg_e = _phantom_registration(B=4, seed=3)
    target = render(subject, reg_g.k_inv, prob.rt_true.float().to(DEV), reg_g.sdd, reg_g.height, reg_g.width, 64)
    traj = {"eager": {}, "graph": {}}

    def monitor(name):
        def record(it, reg):
            traj[name][it] = (reg._rot.detach().clone(), reg._xyz.detach().clone())
            assert torch.isfinite(reg.pose).all().item()  # eager work and a host sync between replays

        return record

    s_e = register(reg_e, target, n_iters=11, n_samples=64, jitter=True, seed=5, monitor=monitor("eager"),
                   monitor_every=1)
    rec = _RecordingLib(_lib.lib())
    monkeypatch.setattr(registration_mod._lib, "lib", lambda: rec)
    s_g = register(reg_g, target, n_iters=11, n_samples=64, jitter=True, seed=5, graph=True, iters_per_graph=3,
                   monitor=monitor("graph"))
    torch.cuda.synchronize()
    captured = [c for c in rec.calls if c[1]]
    if fused:
        launches = {"ndrr_pose_compose_lattice", "ndrr_raymarch_fwd_jac", "ndrr_zncc", "ndrr_raymarch_bwd_jac_se3",
                    "ndrr_adam_step_scheduled_pair"}
    else:
        launches = {"ndrr_lattice_offsets", "ndrr_pose_compose", "ndrr_raymarch_fwd_jac", "ndrr_zncc",
                    "ndrr_raymarch_bwd_jac_se3", "ndrr_adam_step_scheduled", "ndrr_counter_advance"}
    assert len(captured) == 3 * (5 if fused else 8), [c[0] for c in captured]
    assert {c[0] for c in captured} == launches
    assert all(c[2] == c[3] for c in captured)  # on the capturing stream
    assert len({c[4] for c in captured}) == 1  # no torch allocation inside the capture
    assert sorted(traj["graph"]) == [2, 5, 8, 11]
    for it, (rot, xyz) in traj["graph"].items():
        assert torch.equal(rot, traj["eager"][it][0]) and torch.equal(xyz, traj["eager"][it][1]), it
    assert torch.equal(s_g, s_e)


@pytest.mark.parametrize("jitter", [True, False])
def test_fused_loop_launches_match_the_separate_ones(monkeypatch, jitter):
    """register() with the fused loop-state launches against the separate launches: the same
    parameters after every iteration and the same score, bit for bit."""
    from nanodrr_amd import render
    from nanodrr_amd import registration as registration_mod
    from nanodrr_amd.registration import register

    subject, prob, reg_f = _phantom_registration(B=4, seed=3)
    _, _, reg_s = _phantom_registration(B=4, seed=3)
    target = render(subject, reg_f.k_inv, prob.rt_true.float().to(DEV), reg_f.sdd, reg_f.height, reg_f.width, 64)
    traj = {True: [], False: []}
    scores = {}
    for fused, reg in ((True, reg_f), (False, reg_s)):
        monkeypatch.setattr(registration_mod, "FUSED_LOOP_LAUNCHES", fused)
        scores[fused] = register(reg, target, n_iters=9, n_samples=64, jitter=jitter, seed=7, monitor_every=1,
                                 monitor=lambda it, r, f=fused: traj[f].append(
                                     (r._rot.detach().clone(), r._xyz.detach().clone())))
    assert len(traj[True]) == len(traj[False]) == 9
    for (r1, x1), (r2, x2) in zip(traj[True], traj[False]):
        assert torch.equal(r1, r2) and torch.equal(x1, x2)
    assert torch.equal(scores[True], scores[False])
    assert not torch.equal(traj[True][0][0], traj[True][-1][0])  # the parameters moved


def test_registration_converges_on_phantom():
    """200 iterations on 8 problems (64^3 phantom at 4 mm, 64 x 64 DRRs at 4 mm pixels). A single
    coarse view constrains depth along the viewing axis weakly, so some problems stall a few mm
    deep with ZNCC ~0.999 (a CPU run of the same loop through the oracle renderer stalls at the
    same poses); the check is that the loop improves every problem and solves some to < 1 mm."""
    from nanodrr_amd import render
    from nanodrr_amd.metrics import mtre
    from nanodrr_amd.registration import register

    subject, prob, reg = _phantom_registration(B=8, seed=4)
    target = render(subject, reg.k_inv, prob.rt_true.float().to(DEV), reg.sdd, reg.height, reg.width, 128)
    fid = prob.fiducials.to(DEV)
    before = mtre(reg.pose.detach(), prob.rt_true.to(DEV), fid)
    score = register(reg, target, n_iters=200, n_samples=128, graph=False)
    after = mtre(reg.pose.detach(), prob.rt_true.to(DEV), fid)
    print("mTRE before", before.cpu().numpy(), "after", after.cpu().numpy(), "zncc", score.cpu().numpy())
    assert torch.median(after) < 0.5 * torch.median(before)
    assert (after < 1.0).sum() >= 3
    assert torch.median(score) > 0.995


def test_registration_trajectory_matches_oracle_renderer():
    """20 Adam iterations through the native path == the same loop through the oracle's
    grid_sample renderer, FP64 pose form and torch ZNCC on the CPU."""
    from nanodrr_amd.geometry import convert
    from nanodrr_amd.registration import _Adam, register
    from oracle.torch_ref import render_torch

    from nanodrr_amd.phantom import make_phantom, phantom_subject

    subject, prob, reg = _phantom_registration(B=4, seed=5)
    with torch.no_grad():
        reg._rot.zero_()
        reg._xyz.zero_()
    cpu = phantom_subject(make_phantom((64, 64, 64), (4.0, 4.0, 4.0), seed=5), 1, torch.float32, "cpu")
    img, w2g = cpu.image.float(), cpu.world_to_grid.float()
    k_inv, sdd = reg.k_inv.cpu(), reg.sdd.cpu()

    def rend(rt):
        return render_torch(img, None, 1, w2g, k_inv, rt.float(), sdd, reg.height, reg.width, 96)

    with torch.no_grad():
        target = rend(prob.rt_true)
    register(reg, target.to(DEV), n_iters=20, n_samples=96, graph=False)

    rt0, piv, pinv = reg.rt_inv.double().cpu(), reg.pivot.cpu(), reg.pivot_inv.cpu()
    rot = torch.zeros(4, 3, requires_grad=True)
    xyz = torch.zeros(4, 3, requires_grad=True)
    rot.grad, xyz.grad = torch.zeros_like(rot), torch.zeros_like(xyz)
    opt = _Adam([(rot, 1e-2), (xyz, 1.0)], 20)
    for _ in range(20):
        rot.grad.zero_()
        xyz.grad.zero_()
        x = rend(rt0 @ piv @ convert(rot.double(), xyz.double(), "so3_log") @ pinv).reshape(4, -1)
        y = target.reshape(4, -1)
        x = (x - x.mean(1, keepdim=True)) / (x.var(1, keepdim=True, correction=0) + 1e-7).sqrt()
        y = (y - y.mean(1, keepdim=True)) / (y.var(1, keepdim=True, correction=0) + 1e-7).sqrt()
        (-(x * y).mean(1).sum()).backward()
        opt.step()
    # Adam divides by the gradient's running RMS, so components whose gradient is near zero
    # amplify fp32 differences; the bound is a tenth of one step (1e-2 rad, 1 mm) after 20 steps.
    d_rot = (reg._rot.detach().cpu() - rot.detach()).abs().max().item()
    d_xyz = (reg._xyz.detach().cpu() - xyz.detach()).abs().max().item()
    print("trajectory difference: rot", d_rot, "rad, xyz", d_xyz, "mm")
    assert d_rot < 1e-3 and d_xyz < 0.1


@pytest.mark.parametrize("per_pose", [True, False])
def test_explicit_adjoint_matches_autograd_bitwise(per_pose):
    """register()'s explicit adjoint chain (kernels called directly) and the same loop through
    autograd and the custom ops give the same parameters and scores, bit for bit."""
    from nanodrr_amd.registration import register

    runs = []
    for adjoint in ("explicit", "autograd"):
        subject, prob, reg = _phantom_registration(B=3, seed=8, per_pose=per_pose)
        with torch.no_grad():
            target = reg(96).detach()
            reg._rot.add_(0.05)
            reg._xyz.add_(3.0)
        score = register(reg, target, n_iters=8, n_samples=64, jitter=True, seed=2, adjoint=adjoint)
        runs.append((reg._rot.detach().clone(), reg._xyz.detach().clone(), score, reg._rot.grad.clone()))
    for a, b in zip(*runs):
        assert torch.equal(a, b), (a, b)


def test_native_adam_step_matches_torch_optim_adam():
    """ndrr_adam_step over five steps (learning rate changing per step) == torch.optim.Adam."""
    from nanodrr_amd.registration import _adam_step

    g = torch.Generator().manual_seed(4)
    p0 = torch.randn(37, 3, generator=g)
    grads = [torch.randn(37, 3, generator=g) * 10 ** (k - 2) for k in range(5)]
    p = p0.to(DEV).contiguous()
    m, v = torch.zeros_like(p), torch.zeros_like(p)
    ref = p0.clone().requires_grad_(True)
    opt = torch.optim.Adam([ref], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for k, gk in enumerate(grads):
        lr = 1e-2 * (1.0 - 0.1 * k)
        _adam_step(p, gk.to(DEV), m, v, lr, 0.9, 0.999, 1e-8, k + 1)
        opt.param_groups[0]["lr"] = lr
        ref.grad = gk.clone()
        opt.step()
    # the same update in another order of fp32 operations: a few roundings of each quantity's
    # largest entry per step (|p| < 4, ulp 2.4e-7), not of each entry (some end near zero)
    for mine, theirs in ((p, ref.detach()), (m, opt.state[ref]["exp_avg"]), (v, opt.state[ref]["exp_avg_sq"])):
        err = (mine.cpu() - theirs).abs().max() / theirs.abs().max()
        assert err < 1e-6, err


def test_sample_lattice_offset():
    """offset=0 renders exactly as render(); an offset u slides src and tgt by u steps."""
    from nanodrr_amd import render

    subject, prob, reg = _phantom_registration(B=3, seed=6)
    with torch.no_grad():
        base = reg(96)
        zero = reg(96, offset=torch.zeros(3, device=DEV))
        assert torch.equal(base, zero)
        u = torch.tensor([0.5, -0.25, 1.0], device=DEV)
        got = reg(96, offset=u)
        src, tgt = reg._src, reg._tgt
        d = (tgt - src) * (u.reshape(-1, 1, 1) / 95.0)
        ref = render(subject, reg.k_inv, reg.pose.float(), reg.sdd, reg.height, reg.width, 96, src=src + d,
                     tgt=tgt + d)
    err = (got - ref).abs().max() / ref.abs().max()
    assert err < 1e-5, err
    # a whole-step offset moves each sample onto its neighbour's position: interior sums agree
    assert (got[2] - base[2]).abs().max() / base[2].abs().max() < 0.05


def test_lattice_offset_gradients_agree_across_adjoints():
    """With an offset, the single-pass adjoint (offsets taken by the march) and the two-pass
    adjoint (shifted copies of the rays) give the same parameter gradients to fp32 rounding."""
    import nanodrr_amd.renderer as renderer

    subject, prob, reg = _phantom_registration(B=3, seed=8)
    u = torch.tensor([0.4, -0.3, 0.1], device=DEV)
    w = torch.randn(3, 1, 64, 64, generator=torch.Generator().manual_seed(9)).to(DEV)
    grads = {}
    old = renderer.ADJOINT
    try:
        for adjoint in ("single_pass", "two_pass"):
            renderer.ADJOINT = adjoint
            reg._rot.grad = reg._xyz.grad = None
            (reg(96, offset=u) * w).sum().backward()
            grads[adjoint] = (reg._rot.grad.clone(), reg._xyz.grad.clone())
    finally:
        renderer.ADJOINT = old
    for a, b in zip(grads["single_pass"], grads["two_pass"]):
        assert ((a - b).abs().max() / b.abs().max()).item() < 1e-4
