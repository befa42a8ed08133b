"""Volume gradients where many rays share each cell (GPU).

The two-pass backward merges the atomics of lanes whose samples fall in the
same cell (sums over runs of adjacent lanes, then one atomic per run). These
cases make the runs long: detector pixels far smaller than a voxel put whole
8 x 8 ray patches inside one cell, on a ragged detector (edge lanes of every
wave are not rays) and in the linear ray order. The float64 oracle's
volume gradient (oracle/raymarch.py, restating nanodrr's _fused.py:284-298)
is the reference; a value dropped or added twice by the merge would show as
an error of the order of one sample's contribution, far above the tolerance.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from nanodrr_amd._ops import fused_raymarch_bwd
from oracle import geometry as og
from oracle import raymarch as orm
from tests.helpers import random_volume

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _case(pixel_mm, det, n_classes, seed=0, S=160):
    image, label = random_volume(24, n_classes, seed)
    sdd = 100.0
    k_inv = og.make_k_inv(sdd, pixel_mm, pixel_mm, 0.0, 0.0, det, det)
    rt_inv = og.make_rt_inv([[4.0, -7.0, 3.0]], [[0.4, 70.0, -0.3]], "AP", np.zeros(3))
    vol = image[0, 0].double().numpy()
    D, H, W = vol.shape
    w2v = np.eye(4)
    w2v[:3, 3] = [(W - 1) / 2, (H - 1) / 2, (D - 1) / 2]
    tgt = orm.make_tgt(k_inv, np.array([sdd]), det, det)
    src = orm.make_src(False, tgt, np.array([sdd]))
    step = orm.step_size(src, tgt, S)
    M = orm.camera_to_voxel(w2v, rt_inv)
    lab = orm.label_classes(label[0, 0].numpy()) if n_classes > 1 else None
    go = np.random.default_rng(seed + 1).standard_normal((1, n_classes, det * det))
    # the oracle sees the fp32 values the kernel gets, so input rounding moves no sample
    M, src, tgt, step, go = (a.astype(np.float32).astype(np.float64) for a in (M, src, tgt, step, go))
    ref = orm.backward(go, vol, lab, n_classes, M, src, tgt, step, S, need_gvol=True)["gvol"]
    t = {
        "vol": torch.from_numpy(vol).float().to(DEV),
        "lab": (torch.from_numpy(lab).to(torch.int8).to(DEV) if lab is not None
                else torch.empty(0, dtype=torch.int8, device=DEV)),
        "M": torch.from_numpy(M).float().to(DEV),
        "src": torch.from_numpy(src).float().to(DEV),
        "tgt": torch.from_numpy(tgt).float().to(DEV),
        "step": torch.from_numpy(step).float().to(DEV),
        "go": torch.from_numpy(go).float().to(DEV),
    }
    return t, ref, S


@pytest.mark.parametrize("pixel_mm,det", [(0.02, 64), (0.05, 60), (0.3, 50)])
@pytest.mark.parametrize("row_order", [True, False], ids=["tiles", "linear"])
def test_volume_gradient_with_shared_cells_matches_oracle(pixel_mm, det, row_order):
    t, ref, S = _case(pixel_mm, det, 1)
    width = det if row_order else 0
    gvol = fused_raymarch_bwd(t["go"], t["vol"], t["lab"], t["M"], t["src"], t["tgt"], t["step"], S, 1, width,
                              True, False)[0]
    out = gvol.double().cpu().numpy()
    scale = np.abs(ref).max()
    assert scale > 0
    err = np.abs(out - ref).max() / scale
    print(f"pixel {pixel_mm} mm, {det}x{det}, width {width}: gvol error {err:.2e} of the max")
    assert err < 1e-5
    assert out.sum() == pytest.approx(ref.sum(), rel=1e-5, abs=1e-6 * scale)


def test_volume_gradient_with_shared_cells_and_labels():
    """Three label channels: each sample's upstream gradient comes from its class (label
    routing, _fused.py:237), so lanes merged in one cell carry different channels' gradients.
    With the oracle on the kernel's own fp32 inputs no sample changes class (measured 1.7e-6
    of the maximum; from float64 inputs the oracle itself moves by 5e-3 of the maximum here,
    dense rays repeating each label flip)."""
    t, ref, S = _case(0.05, 48, 3)
    gvol = fused_raymarch_bwd(t["go"], t["vol"], t["lab"], t["M"], t["src"], t["tgt"], t["step"], S, 3, 48,
                              True, False)[0]
    out = gvol.double().cpu().numpy()
    scale = np.abs(ref).max()
    err = np.abs(out - ref).max() / scale
    print(f"labels: gvol error {err:.2e} of the max")
    assert err < 1e-5
