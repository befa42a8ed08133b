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
    (3, 10, 64, 1, 32, 1),  # one-channel layer in whole 32-pixel row groups, a wave's groups crossing images
    (2, 9, 32, 1, 64, 1),
    (2, 16, 16, 32, 64, 2),
    (3, 9, 10, 64, 64, 1),
    (2, 17, 15, 64, 128, 2),
    (2, 8, 8, 128, 128, 1),
    (2, 12, 12, 128, 256, 2),
    (1, 7, 9, 32, 32, 1),
    (2, 11, 13, 64, 32, 1),  # stride-1 input gradient streaming 32 dy channels
    (2, 9, 7, 128, 64, 1),  # stride-1 input gradient with 128 columns per workgroup
    # stride 1 with whole 16-, 32- and 64-pixel rows per 64-pixel tile (side columns from lane shifts)
    (2, 6, 64, 64, 64, 1),
    (3, 5, 32, 128, 128, 1),
    (2, 7, 16, 32, 64, 1),
    (2, 4, 32, 64, 128, 1),
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
    if Cin == 1 and Cout != 32:
        pytest.skip("the one-channel weight-gradient kernel takes 32 output channels (the first layer's)")
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


def _torch_head(feat, w1, b1, w2, b2):
    """The head in torch fp32: 2 x 2 mean pool of the NHWC map, inputs ordered (cell, channel) and
    rounded to bf16 as the kernel stores them; the first layer's weights rounded to bf16."""
    N, FH, FW, C = feat.shape
    P = feat.float().reshape(N, FH // 2, 2, FW // 2, 2, C).mean(dim=(2, 4)).reshape(N, -1)
    P = P + (P.to(torch.bfloat16).float() - P).detach()  # bf16 pooled inputs, straight-through
    H = torch.relu(P @ w1.to(torch.bfloat16).float().T + b1)
    return H @ w2.T + b2, H, P


def test_head_forward_and_backward_against_torch():
    """The native head at full size (batch 256, 16 x 16 x 256 feature map, 16384 -> 256 -> 6)
    against torch fp32 on the same bf16 roundings: outputs and hidden units to 1e-5; bias and
    output-layer gradients to 1e-5; the first layer's weight gradient to 1e-4 (bf16 products are
    exact in fp32, but an element of dH whose fp32 sum lands within an ulp of a bf16 rounding
    boundary may round the other way in the two sums); the input gradient to 4e-3 of its
    largest value (bf16 output)."""
    from nanodrr_amd.posenet import head_backward, head_forward

    g = torch.Generator().manual_seed(3)
    feat = torch.relu(torch.randn(256, 16, 16, 256, generator=g)).to(torch.bfloat16)
    w1 = torch.randn(256, 16384, generator=g) / 128
    b1 = torch.randn(256, generator=g) * 0.1
    w2 = torch.randn(6, 256, generator=g) / 16
    b2 = torch.randn(6, generator=g) * 0.1
    dout = torch.randn(256, 6, generator=g)
    out, saved = head_forward(feat.to(DEV), w1.to(DEV), b1.to(DEV), w2.to(DEV), b2.to(DEV))
    ref, H, P = _torch_head(feat, w1, b1, w2, b2)
    assert (out.cpu() - ref).abs().max() / ref.abs().max() < 1e-5
    assert (saved[0].cpu() - H).abs().max() / H.abs().max() < 1e-5
    assert torch.equal(out, head_forward(feat.to(DEV), w1.to(DEV), b1.to(DEV), w2.to(DEV), b2.to(DEV))[0])
    dfeat, dw1, db1, dw2, db2 = head_backward(dout.to(DEV), feat.to(DEV), w2.to(DEV), saved)
    dH = (dout @ w2) * (H > 0)
    dHb = dH.to(torch.bfloat16).float()
    rel = lambda a, b: ((a.cpu().double() - b.double()).abs().max() / b.double().abs().max()).item()  # noqa: E731
    assert rel(dw1, dHb.double().T @ P.detach().double()) < 1e-4
    assert rel(db1, dH.sum(0)) < 1e-5
    assert rel(dw2, dout.double().T @ H.double()) < 1e-5
    assert rel(db2, dout.sum(0)) < 1e-5
    dP = dHb.double() @ w1.to(torch.bfloat16).double()
    ref_dfeat = (dP.reshape(256, 8, 1, 8, 1, 256) * 0.25).expand(-1, -1, 2, -1, 2, -1).reshape(feat.shape)
    ref_dfeat = ref_dfeat * (feat.double() > 0)
    assert rel(dfeat, ref_dfeat) < 4e-3
    assert torch.equal(dfeat.cpu() != 0, ref_dfeat != 0) or ((dfeat.cpu() != 0) != (ref_dfeat != 0)).float().mean() < 1e-4
    # the fp32 master weights are what the gradients refer to: torch autograd through the reference
    w1r, b1r, w2r, b2r = (t.clone().requires_grad_(True) for t in (w1, b1, w2, b2))
    r, _, _ = _torch_head(feat, w1r, b1r, w2r, b2r)
    (r * dout).sum().backward()
    assert rel(db1, b1r.grad) < 1e-5 and rel(dw2, w2r.grad) < 1e-5 and rel(db2, b2r.grad) < 1e-5
    assert rel(dw1, w1r.grad) < 4e-3  # the reference keeps dH in fp32; the kernels round it to bf16


def test_head_rejects_mismatched_shapes():
    from nanodrr_amd.posenet import head_forward

    feat = torch.zeros(2, 4, 4, 256, dtype=torch.bfloat16, device=DEV)
    with pytest.raises(ValueError, match="pools to"):
        head_forward(feat, torch.zeros(256, 16384, device=DEV), torch.zeros(256, device=DEV),
                     torch.zeros(6, 256, device=DEV), torch.zeros(6, device=DEV))


def test_posenet_gradients_match_torch_autograd():
    """The whole network (native convolutions and head, forward and backward) against torch
    autograd of the same network in fp32 on the bf16-rounded weights and activations."""
    from nanodrr_amd.posenet import LAYERS, PoseNet

    torch.manual_seed(0)
    net = PoseNet().to(DEV)
    drr = torch.rand(2, 1, 128, 128, device=DEV) * 5
    out = net(drr)
    loss = (out ** 2).sum()
    loss.backward()
    # reference: the same layers in torch, activations rounded to bf16 after each layer as the kernels do
    x = PoseNet.prepare(drr).float()
    ws = [w.detach().to(torch.bfloat16).float().requires_grad_(True) for w in net.weights]
    bs = [b.detach().clone().requires_grad_(True) for b in net.biases]
    hw = [t.detach().clone().requires_grad_(True) for t in (net.w1, net.b1, net.w2, net.b2)]
    h = x
    for (cin, cout, s), w, b in zip(LAYERS, ws, bs):
        h = torch.relu(_torch_conv(h, w, b, s))
        h = h + (h.to(torch.bfloat16).float() - h).detach()  # bf16 activations, straight-through
    ref, _, _ = _torch_head(h, *hw)
    torch.testing.assert_close(out, ref, rtol=2e-2, atol=2e-3)
    (ref ** 2).sum().backward()
    for i in range(len(LAYERS)):
        g, r = net.weights[i].grad, ws[i].grad
        assert (g - r).abs().max() / r.abs().max() < 5e-2, i  # bf16 activation gradients
    for p, r in zip((net.w1, net.b1, net.w2, net.b2), hw):
        assert (p.grad - r.grad).abs().max() / r.grad.abs().max() < 5e-2


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


def _training_shapes():
    """(N, H, W, Cin, Cout, stride) of every convolution at the training shape: batch 256 of
    128 x 128 DRRs."""
    from nanodrr_amd.posenet import LAYERS

    out, hw = [], 128
    for cin, cout, s in LAYERS:
        out.append((256, hw, hw, cin, cout, s))
        hw = (hw - 1) // s + 1
    return out


@pytest.mark.parametrize("layer", range(6))
def test_every_layer_at_the_training_shape(layer):
    """Each of the six convolutions at its real size in training (batch 256, 128 x 128 input)
    against torch fp32 on the CPU on the same bf16 values: forward with ReLU and input
    gradient with the ReLU mask (8e-3, one bf16 rounding of an fp32 sum), weight and bias
    gradients (sums of exact bf16 products over 256 H W positions: 1e-4 of the largest
    entry for the different summation order)."""
    from nanodrr_amd.posenet import conv3x3, conv3x3_bwd_data, conv3x3_bwd_weight

    N, H, W, Cin, Cout, s = _training_shapes()[layer]
    x, w, b = _inputs(N, H, W, Cin, Cout, 100 + layer)
    x = torch.relu(x)  # activations after a ReLU, as every layer but the first sees them
    Ho, Wo = (H - 1) // s + 1, (W - 1) // s + 1
    dy = torch.randn(N, Ho, Wo, Cout, generator=torch.Generator().manual_seed(layer)).to(torch.bfloat16)
    xd, wd, dyd = x.to(DEV), w.to(DEV), dy.to(DEV)
    y = conv3x3(xd, wd, b.to(DEV), s, True).cpu()
    dw, db = conv3x3_bwd_weight(dyd, xd, s, bias=True)
    dx = conv3x3_bwd_data(dyd, wd, xd, s, (H, W)).cpu() if Cin > 1 else None
    xr = x.float().requires_grad_(True)
    wr = w.float().requires_grad_(True)
    br = b.clone().requires_grad_(True)
    ref = _torch_conv(xr, wr, br, s)
    ref.backward(dy.float())
    ref = ref.detach().clamp(min=0)
    err = (y.float() - ref).abs() / (ref.abs() + 1e-3 * ref.abs().max())
    assert err.max() < 8e-3, ("forward", err.max())
    assert (dw.cpu() - wr.grad).abs().max() / wr.grad.abs().max() < 1e-4
    assert (db.cpu() - br.grad).abs().max() / br.grad.abs().max() < 1e-4
    if dx is not None:
        gref = xr.grad * (x.float() > 0)
        err = (dx.float() - gref).abs() / (gref.abs() + 1e-3 * gref.abs().max())
        assert err.max() < 8e-3, ("input gradient", err.max())


def test_one_training_step_at_full_batch():
    """Training steps at the training configuration: 256 DRRs rendered on the fly from 8
    phantoms, the native network forward and backward and the fused Adam update. The loss is
    finite and every parameter moves; the first step's gradients are finite and nonzero."""
    from nanodrr_amd.posenet import train

    net, hist, _ = train(steps=2, batch=256, seeds=tuple(range(8)), log=lambda m: None, log_every=1)
    assert len(hist) == 2 and all(torch.isfinite(torch.tensor(h["loss"])) for h in hist)
    for p in net.parameters():
        assert torch.isfinite(p).all()
    fresh = type(net)(seed=0, pool=8).to(DEV)
    for (name, p), q in zip(net.named_parameters(), fresh.parameters()):
        assert not torch.equal(p, q), name
