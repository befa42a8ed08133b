"""Numerics of the pose-proposal CNN's bf16 matrix-core convolutions against
the same operations in plain PyTorch fp32 (conv2d and its autograd) on the
same bf16-rounded inputs."""

from __future__ import annotations

import pytest
import torch
import torch.nn.functional as F

pytestmark = pytest.mark.gpu

DEV = "cuda:0"

SHAPES = [  # N, H, W, Cin, Cout, stride
    (2, 13, 11, 1, 32, 1),
    (2, 16, 16, 32, 64, 2),
    (3, 9, 10, 64, 64, 1),
    (2, 17, 15, 64, 128, 2),
    (2, 8, 8, 128, 128, 1),
    (2, 12, 12, 128, 256, 2),
    (1, 7, 9, 32, 32, 1),
]


def _inputs(N, H, W, Cin, Cout, seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(N, H, W, Cin, generator=g).to(torch.bfloat16)
    w = (torch.randn(Cout, 3, 3, Cin, generator=g) / (3 * Cin ** 0.5)).to(torch.bfloat16)
    b = torch.randn(Cout, generator=g) * 0.1
    return x, w, b


def _torch_conv(x, w, b, stride):
    """fp32 reference on NCHW views of the same values."""
    xn = x.float().permute(0, 3, 1, 2)
    wn = w.float().permute(0, 3, 1, 2)
    return F.conv2d(xn, wn, b, stride=stride, padding=1).permute(0, 2, 3, 1)


@pytest.mark.parametrize("shape", SHAPES)
def test_conv_forward_matches_torch(shape):
    from nanodrr_amd.posenet import conv3x3

    N, H, W, Cin, Cout, s = shape
    x, w, b = _inputs(N, H, W, Cin, Cout, sum(shape))
    for relu in (False, True):
        y = conv3x3(x.to(DEV), w.to(DEV), b.to(DEV), s, relu).cpu()
        ref = _torch_conv(x, w, b, s)
        if relu:
            ref = ref.clamp(min=0)
        assert y.shape == ref.shape
        # fp32 accumulation, one rounding to bf16 at the end (2^-8 relative)
        err = (y.float() - ref).abs() / (ref.abs() + 1e-3 * ref.abs().max())
        assert err.max() < 8e-3, err.max()


@pytest.mark.parametrize("shape", [s for s in SHAPES if s[3] != 1])
def test_conv_input_gradient_matches_torch(shape):
    from nanodrr_amd.posenet import conv3x3_bwd_data

    N, H, W, Cin, Cout, s = shape
    x, w, b = _inputs(N, H, W, Cin, Cout, 7 + sum(shape))
    Ho, Wo = (H - 1) // s + 1, (W - 1) // s + 1
    dy = torch.randn(N, Ho, Wo, Cout, generator=torch.Generator().manual_seed(3)).to(torch.bfloat16)
    xr = x.float().requires_grad_(True)
    _torch_conv(xr, w, b, s).backward(dy.float())
    for masked in (False, True):
        dx = conv3x3_bwd_data(dy.to(DEV), w.to(DEV), x.to(DEV) if masked else None, s, (H, W)).cpu().float()
        ref = xr.grad * (x.float() > 0) if masked else xr.grad
        err = (dx - ref).abs() / (ref.abs() + 1e-3 * ref.abs().max())
        assert err.max() < 8e-3, err.max()


@pytest.mark.parametrize("shape", SHAPES)
def test_conv_weight_gradient_matches_torch(shape):
    from nanodrr_amd.posenet import conv3x3_bwd_weight

    N, H, W, Cin, Cout, s = shape
    if Cin != 1 and Cout % 64:
        pytest.skip("the weight-gradient kernel takes 64-channel output tiles")
    x, w, b = _inputs(N, H, W, Cin, Cout, 11 + sum(shape))
    Ho, Wo = (H - 1) // s + 1, (W - 1) // s + 1
    dy = torch.randn(N, Ho, Wo, Cout, generator=torch.Generator().manual_seed(5)).to(torch.bfloat16)
    wr = w.float().requires_grad_(True)
    _torch_conv(x, wr, b, s).backward(dy.float())
    br = b.clone().requires_grad_(True)
    _torch_conv(x, w, br, s).backward(dy.float())
    dw = conv3x3_bwd_weight(dy.to(DEV), x.to(DEV), s).cpu()
    dw2, db = conv3x3_bwd_weight(dy.to(DEV), x.to(DEV), s, bias=True)
    # products of bf16 values are exact in fp32; only the summation order differs
    assert (dw - wr.grad).abs().max() / wr.grad.abs().max() < 1e-5
    assert torch.equal(dw2.cpu(), dw)  # the bias sum does not change the weight gradient
    assert (db.cpu() - br.grad).abs().max() / br.grad.abs().max() < 1e-5


def test_conv_weight_gradient_many_splits_is_deterministic():
    """More pixels than one split: the split partials are added in a fixed order."""
    from nanodrr_amd.posenet import conv3x3_bwd_weight

    x, w, b = _inputs(4, 64, 64, 32, 64, 99)
    dy = torch.randn(4, 32, 32, 64, generator=torch.Generator().manual_seed(9)).to(torch.bfloat16)
    a, da = conv3x3_bwd_weight(dy.to(DEV), x.to(DEV), 2, bias=True)
    c, dc = conv3x3_bwd_weight(dy.to(DEV), x.to(DEV), 2, bias=True)
    assert torch.equal(a, c) and torch.equal(da, dc)
    wr = w.float().requires_grad_(True)
    br = b.clone().requires_grad_(True)
    _torch_conv(x, wr, br, 2).backward(dy.float())
    assert (a.cpu() - wr.grad).abs().max() / wr.grad.abs().max() < 1e-5
    assert (da.cpu() - br.grad).abs().max() / br.grad.abs().max() < 1e-5


def test_posenet_gradients_match_torch_autograd():
    """The whole conv stack (native forward and backward) against torch autograd of the same
    network in fp32 on the bf16-rounded weights and activations, at a small batch."""
    from nanodrr_amd.posenet import LAYERS, PoseNet

    torch.manual_seed(0)
    net = PoseNet().to(DEV)
    drr = torch.rand(2, 1, 32, 32, device=DEV) * 5
    out = net(drr)
    loss = (out ** 2).sum()
    loss.backward()
    # reference: the same layers in torch, activations rounded to bf16 after each layer as the kernels do
    x = PoseNet.prepare(drr).float()
    ws = [w.detach().to(torch.bfloat16).float().requires_grad_(True) for w in net.weights]
    bs = [b.detach().clone().requires_grad_(True) for b in net.biases]
    h = x
    for (cin, cout, s), w, b in zip(LAYERS, ws, bs):
        h = torch.relu(_torch_conv(h, w, b, s))
        h = h + (h.to(torch.bfloat16).float() - h).detach()  # bf16 activations, straight-through
    feat = torch.nn.functional.adaptive_avg_pool2d(h.permute(0, 3, 1, 2), net.pool).flatten(1)
    ref = net.fc2(torch.relu(net.fc1(feat)))
    torch.testing.assert_close(out, ref, rtol=2e-2, atol=2e-3)
    (ref ** 2).sum().backward()
    for i in range(len(LAYERS)):
        g, r = net.weights[i].grad, ws[i].grad
        assert (g - r).abs().max() / r.abs().max() < 5e-2, i  # bf16 activation gradients


def test_posenet_training_and_registration_pipeline_runs():
    """A few training steps on rendered DRRs lower the loss, and the registration comparison
    runs end to end (one phantom, a short loop)."""
    from nanodrr_amd.posenet import compare_registration, train

    net, hist, _ = train(steps=100, batch=32, seeds=(0, 1), log=lambda m: None, log_every=1)
    first = sum(h["loss"] for h in hist[:10]) / 10
    last = sum(h["loss"] for h in hist[-10:]) / 10
    assert last < first, (first, last)
    r = compare_registration(net, n_test=1, n_iters=10, log=lambda m: None, refine=1)
    assert r["problems"] == 1 and 0 <= r["converged_from_cnn"] <= 1
    assert r["refine_renders"] == 1 and r["render_iterations_of_the_short_cnn_run"] == 3
    assert len(r["settled_by_iteration"]["cnn"]) == 2  # checks at iterations 5 and 10
    assert r["median_zncc_cnn"] > 0.5


def test_posenet_proposals_beat_the_nominal_start():
    """What the learned part gains over the method without it: a network trained briefly on
    DRRs this renderer generates proposes starts much closer to unseen targets (+-30 deg /
    +-40 mm) than the nominal pose that plain registration starts from."""
    from nanodrr_amd.posenet import compare_registration, train

    net, hist, _ = train(steps=1500, batch=64, seeds=(0, 1, 2, 3), log=lambda m: None, log_every=500)
    r = compare_registration(net, n_test=6, n_iters=5, log=lambda m: None, refine=1)
    print(f"nominal start {r['median_mtre_mm_nominal_start']:.1f} mm, proposal "
          f"{r['median_mtre_mm_cnn_proposal_unrefined']:.1f} mm, refined {r['median_mtre_mm_cnn_proposal']:.1f} mm")
    assert r["median_mtre_mm_cnn_proposal_unrefined"] < 0.7 * r["median_mtre_mm_nominal_start"]
    assert r["median_mtre_mm_cnn_proposal"] < 0.7 * r["median_mtre_mm_nominal_start"]
