"""Volume gradients where many rays share each cell (GPU).

The two-pass backward gathers each wave's corner contributions at one step in
an LDS box around the middle ray's cell (contributions outside it go to memory
directly), then adds each touched voxel once; through the op the adds go to a
4 x 4 x 4-voxel brick buffer that a second launch adds into gvol and clears
(ndrr_raymarch_bwd_bricked), else straight into gvol (ndrr_raymarch_bwd). These
cases make whole 8 x 8 ray patches share cells (detector pixels far smaller
than a voxel), or spread them past the box (pixels larger than a voxel), on a
ragged detector (edge lanes of every wave are not rays), in the linear ray
order and on a volume whose sides are not multiples of 4. The float64
oracle's volume gradient (oracle/raymarch.py, restating nanodrr's
_fused.py:284-298) is the reference; a value dropped or added twice would
show as an error of the order of one sample's contribution, far above the
tolerance.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from nanodrr_amd import _lib
from nanodrr_amd import _ops
from nanodrr_amd._ops import fused_raymarch_bwd
from oracle import geometry as og
from oracle import raymarch as orm
from tests.helpers import random_volume

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _case(pixel_mm, det, n_classes, seed=0, S=160, size=24):
    image, label = random_volume(size, n_classes, seed)
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


@pytest.fixture(params=[True, False], ids=["bricks", "direct"])
def bricks(request, monkeypatch):
    monkeypatch.setattr(_ops, "GVOL_BRICKS", request.param)
    return request.param


@pytest.mark.parametrize("pixel_mm,det,size", [(0.02, 64, 24), (0.05, 60, 24), (0.3, 50, 24), (1.5, 40, 21)])
@pytest.mark.parametrize("row_order", [True, False], ids=["tiles", "linear"])
def test_volume_gradient_with_shared_cells_matches_oracle(pixel_mm, det, size, row_order, bricks):
    t, ref, S = _case(pixel_mm, det, 1, size=size)
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


def test_volume_gradient_with_shared_cells_and_labels(bricks):
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


def test_brick_buffer_is_left_zero_and_reused():
    """ndrr_raymarch_bwd_bricked through the C ABI: the buffer is zero after each call, two calls
    with it give the same gvol to atomic-order rounding, and ndrr_raymarch_bwd's gvol matches."""
    t, ref, S = _case(0.3, 50, 1, size=21)
    lib = _lib.lib()
    D, H, W = t["vol"].shape
    N = t["tgt"].shape[1]
    assert lib.ndrr_gvol_brick_floats(D, H, W) == 6 * 6 * 6 * 64
    bricks = torch.zeros(lib.ndrr_gvol_brick_floats(D, H, W), device=DEV)
    ws = torch.empty(lib.ndrr_bwd_workspace_floats(1, N, 50), device=DEV)
    gM = torch.empty(1, 3, 4, device=DEV)
    outs = []
    for bricked in (True, True, False):
        gvol = torch.zeros(D, H, W, device=DEV)
        args = (t["go"].data_ptr(), t["vol"].data_ptr(), _lib.NDRR_F32, _lib.LAYOUT_LINEAR, None, t["M"].data_ptr(),
                t["src"].data_ptr(), t["tgt"].data_ptr(), t["step"].data_ptr(), ws.data_ptr(), gM.data_ptr(), None,
                None, None, gvol.data_ptr())
        tail = (0, 1, N, D, H, W, S, 1, 50, 0, 0, 0, 0, None)
        rc = lib.ndrr_raymarch_bwd_bricked(*args, bricks.data_ptr(), *tail) if bricked else \
            lib.ndrr_raymarch_bwd(*args, *tail)
        assert rc == 0
        torch.cuda.synchronize()
        assert int(torch.count_nonzero(bricks)) == 0
        outs.append(gvol.double().cpu().numpy())
    scale = np.abs(ref).max()
    for out in outs:
        assert np.abs(out - ref).max() / scale < 1e-5
    assert lib.ndrr_raymarch_bwd_bricked(*args, None, *tail) == -1  # no buffer
