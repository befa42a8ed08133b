"""Parity of the HIP ray march with the float64 oracle (GPU).

Mirrors nanodrr's backend-parity tests (tests/drr/test_fused.py:68-226) with
the oracle in place of the grid_sample backend and tighter tolerances:
forward <= 1e-5 and gradients <= 1e-4 relative to the largest entry
(the reference's floors are 1e-4 and 1e-3).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

from nanodrr_amd import render
from nanodrr_amd._ops import fused_raymarch
from nanodrr_amd.camera import make_k_inv, make_rt_inv
from oracle import raymarch as orm
from oracle import torch_ref
from tests.helpers import oracle_inputs, random_subject, random_volume, ref_camera_np, rel_to_max

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda")
GOLDEN = Path(__file__).resolve().parent / "golden" / "ref_camera_32.npz"
FWD_TOL = 1e-5
GRAD_TOL = 1e-4


@pytest.fixture(params=["quad", "stencil", "coef", "linear"], autouse=True)
def volume_layout(request, monkeypatch):
    """Every test runs on every volume layout the kernels read."""
    import nanodrr_amd.renderer as renderer

    monkeypatch.setattr(renderer, "VOLUME_LAYOUT", request.param)
    return request.param


@pytest.fixture(params=["single_pass", "two_pass"], autouse=True)
def adjoint(request, monkeypatch):
    """... and with both adjoints for pose and camera gradients."""
    import nanodrr_amd.renderer as renderer

    monkeypatch.setattr(renderer, "ADJOINT", request.param)
    return request.param


def camera(device=DEV, height=16, width=16):
    sdd = 100.0
    k_inv = make_k_inv(sdd, 1.0, 1.0, 0.0, 0.0, height, width, device=device)
    rt_inv = make_rt_inv(torch.tensor([[5.0, -3.0, 8.0]]), torch.tensor([[1.0, 80.0, -2.0]]), "AP",
                         torch.zeros(3)).to(device)
    return k_inv, rt_inv, torch.tensor([sdd], device=device), height, width


def _oracle_fwd(n_classes, orthographic=False, seed=0, S=200):
    image, label = random_volume(32, n_classes, seed)
    k_inv, rt_inv, sdd, h, w = ref_camera_np()
    return orm.forward(**oracle_inputs(image, label, n_classes, k_inv, rt_inv, sdd, h, w, S, orthographic))


def _assert_parity(ref, out, n_classes, tol=FWD_TOL):
    """Label routing is discontinuous, so ~1e-6 coordinate differences may flip
    single samples between channels (test_fused.py:56-65)."""
    ref = np.asarray(ref, np.float64)
    out = np.asarray(out, np.float64)
    scale = np.abs(ref).max()
    if n_classes == 1:
        assert np.abs(ref - out).max() / scale < tol
    else:
        assert np.abs(ref.sum(1) - out.sum(1)).max() / scale < tol
        bad = (np.abs(ref - out).max(axis=1) / scale) > tol
        assert bad.mean() < 1e-3


@pytest.mark.parametrize("orthographic", [False, True])
@pytest.mark.parametrize("n_classes", [1, 3])
def test_forward_matches_oracle(orthographic, n_classes):
    subject = random_subject(n_classes=n_classes, device=DEV)
    k_inv, rt_inv, sdd, h, w = camera()
    out = render(subject, k_inv, rt_inv, sdd, h, w, n_samples=200, orthographic=orthographic)
    assert out.shape == (1, n_classes, h, w) and out.dtype == torch.float32
    ref = _oracle_fwd(n_classes, orthographic)
    _assert_parity(ref.reshape(out.shape), out.cpu().numpy(), n_classes)


@pytest.mark.parametrize("name,ortho,C", [("persp_c1", False, 1), ("ortho_c1", True, 1),
                                          ("persp_c3", False, 3), ("ortho_c3", True, 3)])
def test_against_golden_vectors(name, ortho, C, volume_layout):
    """Forward and pose/endpoint gradients vs the committed float64 fixtures."""
    from nanodrr_amd._ops import pack_labels, pack_volume

    g = np.load(GOLDEN)
    vol = torch.from_numpy(g["volume"]).to(DEV)
    packed = None if volume_layout == "linear" else pack_volume(vol, volume_layout)
    lab = torch.from_numpy(g["label3"]).to(DEV) if C > 1 else torch.empty(0, dtype=torch.int8, device=DEV)
    if C > 1:
        lab = pack_labels(lab, volume_layout)
    D, H, W = vol.shape
    w2v = np.eye(4)
    w2v[:3, 3] = [(W - 1) / 2, (H - 1) / 2, (D - 1) / 2]
    tgt = orm.make_tgt(g["k_inv"], g["sdd"], 16, 16)
    src = orm.make_src(ortho, tgt, g["sdd"])
    step = orm.step_size(src, tgt, 200)
    M = orm.camera_to_voxel(w2v, g["rt_inv"])
    t = {k: torch.tensor(v, dtype=torch.float32, device=DEV).requires_grad_(True)
         for k, v in (("M", M), ("src", src), ("tgt", tgt), ("step", step))}
    out = fused_raymarch(vol, lab, t["M"], t["src"], t["tgt"], t["step"], 200, C, 16, packed, volume_layout)
    _assert_parity(g[f"{name}_fwd"], out.detach().cpu().numpy(), C)
    (out * torch.tensor(g[f"{name}_go"], dtype=torch.float32, device=DEV)).sum().backward()
    tol = GRAD_TOL if C == 1 else 2e-3  # label flips move single samples between channels
    assert rel_to_max(g[f"{name}_gM"], t["M"].grad.cpu()) < tol
    gsrc = g[f"{name}_gsrc"] if ortho else g[f"{name}_gsrc"].sum(1, keepdims=True)
    assert rel_to_max(gsrc, t["src"].grad.cpu()) < tol
    assert rel_to_max(g[f"{name}_gtgt"], t["tgt"].grad.cpu()) < tol
    assert rel_to_max(g[f"{name}_gstep"], t["step"].grad.cpu()) < tol


def _torch_ref_grads(n_classes, orthographic, wrt, seed_w):
    """float64 autograd of the grid_sample restatement on the CPU (the reference's torch backend)."""
    image, label = random_volume(32, n_classes, seed=0)
    k_inv, rt_inv, sdd, h, w = camera(torch.device("cpu"))
    k_inv, rt_inv, sdd = k_inv.double(), rt_inv.double(), sdd.double()
    torch.manual_seed(seed_w)
    wts = torch.randn(1, n_classes, h, w)
    leaves = {"rt_inv": rt_inv, "k_inv": k_inv, "sdd": sdd, "image": image.double()}
    for k in wrt:
        leaves[k] = leaves[k].clone().requires_grad_(True)
    w2g = torch_ref.voxel_to_grid((32, 32, 32)) @ torch.linalg.inv(_v2w(32))
    out = torch_ref.render_torch(leaves["image"], label.double(), n_classes, w2g, leaves["k_inv"], leaves["rt_inv"],
                                 leaves["sdd"], h, w, 200, orthographic)
    (out * wts.double()).sum().backward()
    return {k: leaves[k].grad for k in wrt}, wts


def _v2w(size):
    v2w = torch.eye(4, dtype=torch.float64)
    v2w[:3, 3] = -(size - 1) / 2
    return v2w


@pytest.mark.parametrize("orthographic", [False, True])
def test_pose_gradients_match_reference_backend(orthographic):
    ref, wts = _torch_ref_grads(1, orthographic, ["rt_inv"], 1)
    subject = random_subject(device=DEV)
    k_inv, rt_inv, sdd, h, w = camera()
    rt = rt_inv.clone().requires_grad_(True)
    out = render(subject, k_inv, rt, sdd, h, w, n_samples=200, orthographic=orthographic)
    (out * wts.to(DEV)).sum().backward()
    # SE(3) rows only: the homogeneous row's phantom gradient differs by design (test_fused.py:97)
    assert rel_to_max(ref["rt_inv"][0, :3], rt.grad[0, :3].cpu()) < GRAD_TOL


@pytest.mark.parametrize("n_classes", [1, 3])
def test_volume_gradients_match_reference_backend(n_classes):
    """A volume that requires grad: 'auto' would pick the linear layout; 'stencil'
    re-packs it and must give the same gradient (the atomics target the linear volume)."""
    ref, wts = _torch_ref_grads(n_classes, False, ["image"], 2)
    subject = random_subject(n_classes=n_classes, device=DEV)
    subject._image_hu = subject._image_hu.detach().requires_grad_(True)
    k_inv, rt_inv, sdd, h, w = camera()
    out = render(subject, k_inv, rt_inv, sdd, h, w, n_samples=200)
    (out * wts.to(DEV)).sum().backward()
    tol = GRAD_TOL if n_classes == 1 else 1e-3
    assert rel_to_max(ref["image"], subject._image_hu.grad.cpu()) < tol


def test_intrinsics_gradients_match_reference_backend():
    ref, wts = _torch_ref_grads(1, False, ["k_inv", "sdd"], 3)
    subject = random_subject(device=DEV)
    k_inv, rt_inv, sdd, h, w = camera()
    k = k_inv.clone().requires_grad_(True)
    s = sdd.clone().requires_grad_(True)
    out = render(subject, k, rt_inv, s, h, w, n_samples=200)
    (out * wts.to(DEV)).sum().backward()
    assert rel_to_max(ref["k_inv"], k.grad.cpu()) < GRAD_TOL
    assert rel_to_max(ref["sdd"], s.grad.cpu()) < GRAD_TOL


def test_pose_gradients_deterministic():
    """Per-tile partials reduced in a fixed order: bitwise-stable pose gradients."""
    subject = random_subject(device=DEV)
    k_inv, rt_inv, sdd, h, w = camera()
    torch.manual_seed(5)
    wts = torch.randn(1, 1, h, w, device=DEV)
    grads = []
    for _ in range(3):
        rt = rt_inv.clone().requires_grad_(True)
        (render(subject, k_inv, rt, sdd, h, w, n_samples=200) * wts).sum().backward()
        grads.append(rt.grad.clone())
    torch.testing.assert_close(grads[0], grads[1], rtol=0, atol=0)
    torch.testing.assert_close(grads[0], grads[2], rtol=0, atol=0)


@pytest.mark.parametrize("n_classes", [1, 3])
def test_pose_batch_broadcasts_against_shared_detector(n_classes):
    subject = random_subject(n_classes=n_classes, device=DEV)
    k_inv, _, sdd, h, w = camera()
    rots = [[5.0, -3.0, 8.0], [25.0, 4.0, -6.0]]
    trans = [[1.0, 80.0, -2.0], [-3.0, 78.0, 5.0]]
    rt_inv = make_rt_inv(torch.tensor(rots), torch.tensor(trans), "AP", torch.zeros(3)).to(DEV)
    out = render(subject, k_inv, rt_inv, sdd, h, w, n_samples=200)
    assert out.shape == (2, n_classes, h, w)
    image, label = random_volume(32, n_classes, 0)
    from oracle import geometry as og

    kn, _, sn, _, _ = ref_camera_np()
    rtn = og.make_rt_inv(rots, trans, "AP", np.zeros(3))
    ref = orm.forward(**oracle_inputs(image, label, n_classes, kn, rtn, sn, h, w, 200))
    _assert_parity(ref.reshape(out.shape), out.cpu().numpy(), n_classes)


def test_out_of_range_labels_are_dropped():
    """Labels >= n_classes are dropped in forward and masked in backward (test_fused.py:212-226)."""
    subject = random_subject(n_classes=5, device=DEV)
    subject.n_classes = 3
    k_inv, rt_inv, sdd, h, w = camera()
    rt = rt_inv.clone().requires_grad_(True)
    out = render(subject, k_inv, rt, sdd, h, w, n_samples=200)
    out.sum().backward()
    assert out.shape[1] == 3
    assert torch.isfinite(out).all() and torch.isfinite(rt.grad).all()
    image, label = random_volume(32, 5, 0)
    kn, rtn, sn, _, _ = ref_camera_np()
    inp = oracle_inputs(image, label, 3, kn, rtn, sn, h, w, 200)
    ref = orm.forward(**inp)
    _assert_parity(ref.reshape(out.shape), out.detach().cpu().numpy(), 3)


def test_unit_impulse_integral_equals_one():
    """tests/drr/test_renderer.py:52-91 on the GPU."""
    from nanodrr_amd import Subject

    image = torch.zeros(1, 1, 3, 3, 3)
    image[0, 0, 1, 1, 1] = 1.0
    v2w = torch.eye(4)
    v2w[:3, 3] = -1.0
    subject = Subject(image, torch.zeros_like(image), v2w, torch.linalg.inv(v2w),
                      Subject._make_voxel_to_grid(image.shape), torch.zeros(3), max_label=0,
                      convert_to_mu=False).to(DEV)
    out = render(subject, torch.eye(3, device=DEV)[None], torch.eye(4, device=DEV)[None],
                 torch.tensor([1.0], device=DEV), 1, 1, n_samples=1001,
                 src=torch.tensor([[[-1.5, 0.0, 0.0]]], device=DEV), tgt=torch.tensor([[[1.5, 0.0, 0.0]]], device=DEV))
    assert float(out.squeeze()) == pytest.approx(1.0, rel=1e-3, abs=1e-3)


def test_volume_gradient_refused_under_determinism():
    subject = random_subject(device=DEV)
    subject._image_hu = subject._image_hu.detach().requires_grad_(True)
    k_inv, rt_inv, sdd, h, w = camera()
    out = render(subject, k_inv, rt_inv, sdd, h, w, n_samples=64)
    torch.use_deterministic_algorithms(True)
    try:
        with pytest.raises(RuntimeError, match="nondeterministic"):
            out.sum().backward()
    finally:
        torch.use_deterministic_algorithms(False)


@pytest.mark.parametrize("n_classes", [1, 3])
def test_compiles_as_opaque_custom_op(n_classes):
    """torch.compile(fullgraph=True) traces both launches as opaque ops (test_fused.py:147-169);
    aot_eager keeps the surrounding ops in eager ATen (no generated kernels)."""
    subject = random_subject(n_classes=n_classes, device=DEV)
    k_inv, rt_inv, sdd, h, w = camera()
    torch.manual_seed(4)
    wts = torch.randn(1, n_classes, h, w, device=DEV)

    def f(rt):
        return render(subject, k_inv, rt, sdd, h, w, n_samples=200)

    rt = rt_inv.clone().requires_grad_(True)
    (f(rt) * wts).sum().backward()
    torch._dynamo.reset()
    rt_c = rt_inv.clone().requires_grad_(True)
    (torch.compile(f, fullgraph=True, backend="aot_eager")(rt_c) * wts).sum().backward()
    torch.testing.assert_close(rt.grad, rt_c.grad, rtol=0, atol=0)


def test_hip_graph_replays_match_eager():
    """Forward + backward captured in a HIP graph; every replay equals the eager result."""
    subject = random_subject(device=DEV)
    k_inv, rt_inv, sdd, h, w = camera()
    torch.manual_seed(6)
    wts = torch.randn(1, 1, h, w, device=DEV)
    rt_static = rt_inv.clone().requires_grad_(True)

    def step():
        rt_static.grad = None
        out = render(subject, k_inv, rt_static, sdd, h, w, n_samples=200)
        (out * wts).sum().backward()
        return out

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):
            step()
    torch.cuda.current_stream().wait_stream(s)
    eager_out = step().detach().clone()
    eager_grad = rt_static.grad.clone()

    graph = torch.cuda.CUDAGraph()
    rt_static.grad = None
    with torch.cuda.graph(graph):
        out_static = render(subject, k_inv, rt_static, sdd, h, w, n_samples=200)
        (out_static * wts).sum().backward()
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        torch.testing.assert_close(out_static, eager_out, rtol=0, atol=0)
        torch.testing.assert_close(rt_static.grad, eager_grad, rtol=0, atol=0)
        _ = (eager_out * 2).sum().item()  # eager work between replays


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16, torch.float16])
def test_packed_layouts_hold_the_corners(dtype, volume_layout):
    """ndrr_pack_volume, bit for bit: stencil record (x0, y0, z0) holds the
    zero-padded corners vol[z0+a, y0+b, x0+c] in (a, b, c) order; quad record
    (x, y0, z0) holds vol[z0+a, y0+b, x] in (a, b) order; coefficient record
    (x0, y0, z0) holds the cell's interpolant coefficients (oracle.raymarch.coef_records)."""
    from nanodrr_amd._ops import pack_volume

    torch.manual_seed(0)
    D, H, W = 5, 6, 7
    vol = torch.randn(D, H, W, device=DEV).to(dtype)
    rec = pack_volume(vol, volume_layout)
    pad = torch.nn.functional.pad(vol.float(), (1, 1, 1, 1, 1, 1)).to(dtype)  # (D+2, H+2, W+2)
    if volume_layout == "stencil":
        corners = [pad[a:a + D + 1, b:b + H + 1, c:c + W + 1] for a in (0, 1) for b in (0, 1) for c in (0, 1)]
        expect = torch.stack(corners, dim=-1)  # (D+1, H+1, W+1, 8)
    elif volume_layout == "quad":
        corners = [pad[a:a + D + 1, b:b + H + 1, :] for a in (0, 1) for b in (0, 1)]
        expect = torch.stack(corners, dim=-1)  # (D+1, H+1, W+2, 4)
    elif volume_layout == "coef":  # coefficients in fp32 in the documented order, scaled, rounded once
        from tests.helpers import coef_expected

        slots, _ = coef_expected(vol.float().cpu().numpy(), dtype)
        body = rec[256:].cpu().view(slots.dtype).reshape(slots.shape)
        assert torch.equal(body, slots)
        return
    else:
        expect = vol
    got = rec.view(dtype).reshape(expect.shape)
    assert torch.equal(got.view(torch.uint8), expect.contiguous().view(torch.uint8))


def test_label_quads_hold_the_classes(volume_layout):
    """ndrr_pack_labels: quad record (x, y0, z0) holds lab[z0+a, y0+b, x] in byte 2a + b
    (zero outside); the stencil and coefficient layouts read the map padded by two zero voxels
    on every side; the linear layout reads the (D, H, W) map itself."""
    from nanodrr_amd._ops import pack_labels

    torch.manual_seed(1)
    D, H, W = 4, 5, 6
    lab = torch.randint(-1, 5, (D, H, W), dtype=torch.int8, device=DEV)
    got = pack_labels(lab, volume_layout)
    if volume_layout == "linear":
        assert torch.equal(got, lab)
        return
    if volume_layout in ("stencil", "coef"):
        expect = torch.nn.functional.pad(lab.to(torch.int32), (2, 2, 2, 2, 2, 2)).to(torch.int8)
        assert torch.equal(got.view(D + 4, H + 4, W + 4), expect)
        return
    pad = torch.nn.functional.pad(lab.to(torch.int32), (1, 1, 1, 1, 1, 1)).to(torch.int8)
    expect = torch.stack([pad[a:a + D + 1, b:b + H + 1, :] for a in (0, 1) for b in (0, 1)], dim=-1)
    assert torch.equal(got.view(D + 1, H + 1, W + 2, 4), expect)


@pytest.mark.parametrize("adjoint", ["single_pass", "two_pass"])
def test_rays_that_miss_the_volume(adjoint):
    """A pose whose rays all miss the volume renders exact zeros with exact-zero, finite pose
    gradients, and leaves the pose that hits it in the same batch as it renders alone."""
    import nanodrr_amd.renderer as renderer

    subject = random_subject(n_classes=2, device=DEV)
    k_inv, _, sdd, h, w = camera()
    rots = torch.tensor([[5.0, -3.0, 8.0], [5.0, -3.0, 8.0]])
    trans = torch.tensor([[1.0, 80.0, -2.0], [900.0, 80.0, -2.0]])  # the second is 900 mm to the side
    old = renderer.ADJOINT
    try:
        renderer.ADJOINT = adjoint
        rt = make_rt_inv(rots, trans, "AP", torch.zeros(3)).to(DEV).requires_grad_(True)
        out = render(subject, k_inv, rt, sdd, h, w, n_samples=200)
        out.sum().backward()
        rt1 = make_rt_inv(rots[:1], trans[:1], "AP", torch.zeros(3)).to(DEV).requires_grad_(True)
        one = render(subject, k_inv, rt1, sdd, h, w, n_samples=200)
        one.sum().backward()
    finally:
        renderer.ADJOINT = old
    assert out[0].abs().max() > 0 and torch.count_nonzero(out[1]) == 0
    assert torch.isfinite(rt.grad).all() and torch.count_nonzero(rt.grad[1]) == 0
    assert torch.equal(out[0], one[0]) and torch.equal(rt.grad[0], rt1.grad[0])
