"""nanodrr's own op schema on the native library (GPU).

`torch.ops.nanodrr.fused_raymarch` as a nanodrr maintainer would bind it
(tests/reference_binding.py: ctypes over the C ABI, the reference's 9-argument
schema with a float label volume and a `vol.dtype` output, its fake impl and
autograd formula) on the reference test geometry (tests/drr/test_fused.py:
37-47: 32^3 volume, 16 x 16 detector, sdd 100, ZXY (5, -3, 8) deg at 80 mm):
forward and every gradient against the committed float64 fixtures and the
oracle, bf16 and fp16 volumes returning their own dtype, and the op traced
opaquely by torch.compile.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

import tests.reference_binding  # noqa: F401  (registers torch.ops.nanodrr.*)
from oracle import raymarch as orm
from tests.helpers import rel_to_max

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
GOLDEN = Path(__file__).resolve().parent / "golden" / "ref_camera_32.npz"


def _geometry(g, ortho):
    D, H, W = g["volume"].shape
    w2v = np.eye(4)
    w2v[:3, 3] = [(W - 1) / 2, (H - 1) / 2, (D - 1) / 2]
    tgt = orm.make_tgt(g["k_inv"], g["sdd"], 16, 16)
    src = orm.make_src(ortho, tgt, g["sdd"])
    step = orm.step_size(src, tgt, 200)
    M = orm.camera_to_voxel(w2v, g["rt_inv"])
    return {k: torch.tensor(v, dtype=torch.float32, device=DEV) for k, v in
            (("M", M), ("src", src), ("tgt", tgt), ("step", step))}


def _labels(g, C, vol):
    """The reference's `lab`: the float label volume, or `vol` itself with one class (backends.py:89-90)."""
    return torch.from_numpy(g["label3"]).to(torch.float32).to(DEV) if C > 1 else vol


@pytest.mark.parametrize("name,ortho,C", [("persp_c1", False, 1), ("ortho_c1", True, 1),
                                          ("persp_c3", False, 3), ("ortho_c3", True, 3)])
def test_reference_schema_against_golden_vectors(name, ortho, C):
    g = np.load(GOLDEN)
    vol = torch.from_numpy(g["volume"]).to(DEV)
    lab = _labels(g, C, vol)
    t = {k: v.requires_grad_(True) for k, v in _geometry(g, ortho).items()}
    out = torch.ops.nanodrr.fused_raymarch(vol, lab, t["M"], t["src"], t["tgt"], t["step"], 200, C, 16)
    assert out.dtype == torch.float32 and out.shape == (1, C, 256)
    ref = g[f"{name}_fwd"]
    scale = np.abs(ref).max()
    got = out.detach().double().cpu().numpy()
    if C == 1:
        assert np.abs(ref - got).max() / scale < 1e-5
    else:  # label routing: a sample within ~1e-6 voxel of a rounding boundary may change class
        assert np.abs(ref.sum(1) - got.sum(1)).max() / scale < 1e-5
        assert np.mean(np.abs(ref - got).max(axis=1) / scale > 1e-5) < 1e-3
    (out * torch.tensor(g[f"{name}_go"], dtype=torch.float32, device=DEV)).sum().backward()
    tol = 1e-4 if C == 1 else 2e-3
    assert rel_to_max(g[f"{name}_gM"], t["M"].grad.cpu()) < tol
    gsrc = g[f"{name}_gsrc"] if ortho else g[f"{name}_gsrc"].sum(1, keepdims=True)
    assert rel_to_max(gsrc, t["src"].grad.cpu()) < tol
    assert rel_to_max(g[f"{name}_gtgt"], t["tgt"].grad.cpu()) < tol
    assert rel_to_max(g[f"{name}_gstep"], t["step"].grad.cpu()) < tol


@pytest.mark.parametrize("C", [1, 3])
def test_reference_schema_volume_gradient(C):
    """need_gvol through the reference's autograd formula against the oracle's volume gradient
    (on the kernel's fp32 inputs, so no sample moves)."""
    g = np.load(GOLDEN)
    vol = torch.from_numpy(g["volume"]).to(DEV).requires_grad_(True)
    lab = _labels(g, C, vol.detach())
    t = _geometry(g, False)
    go = torch.randn(1, C, 256, generator=torch.Generator().manual_seed(C)).to(DEV)
    out = torch.ops.nanodrr.fused_raymarch(vol, lab, t["M"], t["src"], t["tgt"], t["step"], 200, C, 16)
    (out * go).sum().backward()
    lab_np = orm.label_classes(g["label3"].astype(np.float32)) if C > 1 else None
    ref = orm.backward(go.double().cpu().numpy(), g["volume"], lab_np, C, *(t[k].double().cpu().numpy()
                       for k in ("M", "src", "tgt", "step")), 200, need_gvol=True)["gvol"]
    assert vol.grad.dtype == torch.float32
    assert rel_to_max(ref, vol.grad.cpu()) < 1e-5


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
def test_reference_schema_half_volumes_return_their_dtype(dtype):
    """The forward returns vol.dtype (_fused.py:460) and the volume gradient comes back in it."""
    g = np.load(GOLDEN)
    vol = torch.from_numpy(g["volume"]).to(DEV).to(dtype).requires_grad_(True)
    t = _geometry(g, False)
    out = torch.ops.nanodrr.fused_raymarch(vol, vol.detach(), t["M"], t["src"], t["tgt"], t["step"], 200, 1, 16)
    assert out.dtype == dtype
    ref = orm.forward(vol.detach().double().cpu().numpy(), None, 1, *(t[k].double().cpu().numpy()
                      for k in ("M", "src", "tgt", "step")), 200)
    assert rel_to_max(ref, out.detach().double().cpu().numpy()) < 8e-3  # one rounding of the fp32 sum
    out.float().sum().backward()
    assert vol.grad.dtype == dtype and torch.isfinite(vol.grad.float()).all()


def test_reference_schema_compiles_opaquely():
    """torch.compile(fullgraph=True) traces the op through its fake impl and calls it as is."""
    g = np.load(GOLDEN)
    vol = torch.from_numpy(g["volume"]).to(DEV)
    lab = _labels(g, 3, vol)
    t = _geometry(g, False)

    def f(M):
        return torch.ops.nanodrr.fused_raymarch(vol, lab, M, t["src"], t["tgt"], t["step"], 200, 3, 16).sum(1)

    eager = f(t["M"])
    compiled = torch.compile(f, fullgraph=True)(t["M"])
    assert torch.equal(eager, compiled)
