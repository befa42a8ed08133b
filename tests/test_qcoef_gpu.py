"""Quad coefficient records ("qcoef", NDRR_LAYOUT_QCOEF) on the GPU.

* ndrr_pack_volume's quad coefficient records bit for bit against tests.helpers.qcoef_expected
  (fp32 coefficients in the documented order, times the coefficient layout's scale, rounded once to
  fp16), header included, from fp32, bf16 and fp16 volumes.
* The march (forward, single-pass and two-pass adjoints) against the float64 oracle evaluating the
  records it reads (decoded and converted to the cell coefficients they define): the kernels'
  arithmetic, to the usual 1e-5 / 1e-4, with labels read from the quad grid's label quads.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from nanodrr_amd import _lib
from nanodrr_amd._ops import (fused_raymarch, fused_raymarch_bwd, fused_raymarch_bwd_jac, fused_raymarch_jac,
                              pack_labels, pack_volume)
from oracle import raymarch as orm
from tests.helpers import (coef_header, oracle_inputs, qcoef_decode, qcoef_expected, qcoef_to_coef, random_volume,
                           ref_camera_np, rel_to_max)

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


@pytest.mark.parametrize("src_dtype", [torch.float32, torch.bfloat16, torch.float16])
@pytest.mark.parametrize("magnitude", [1e-4, 3.0, 2e4])
def test_qcoef_records_bit_for_bit(src_dtype, magnitude):
    torch.manual_seed(1)
    vol = (torch.randn(5, 6, 7) * magnitude).to(src_dtype)
    vol[1, 2] = 0.0
    packed = pack_volume(vol.to(DEV), "qcoef")
    expect, s = qcoef_expected(vol.float().numpy())
    assert coef_header(packed) == (s, 1.0 / s, _lib.NDRR_F16)
    got = packed[256:].cpu().view(torch.int16).reshape(expect.shape)
    assert torch.equal(got, expect)
    assert packed.numel() == 256 + 6 * 7 * 9 * 8


def test_qcoef_labels_are_the_label_quads():
    lab = torch.randint(-1, 3, (5, 6, 7), dtype=torch.int8, device=DEV)
    assert torch.equal(pack_labels(lab, "qcoef"), pack_labels(lab, "quad"))


def test_qcoef_needs_a_16bit_volume():
    image, label = random_volume(8, 1, 0)
    vol = image[0, 0].to(DEV)
    packed = pack_volume(vol, "qcoef")
    M = torch.eye(4, device=DEV)[None, :3, :].contiguous()
    tgt = torch.ones(1, 4, 3, device=DEV)
    src = torch.zeros(1, 1, 3, device=DEV)
    step = torch.ones(1, 4, device=DEV)
    with pytest.raises(RuntimeError, match="element type"):
        fused_raymarch(vol, vol.new_empty(0, dtype=torch.int8), M, src, tgt, step, 8, 1, 0, packed, "qcoef")


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("n_classes", [1, 3])
def test_qcoef_march_matches_oracle_of_its_records(dtype, n_classes):
    image, label = random_volume(32, n_classes, 0)
    k_inv, rt_inv, sdd, h, w = ref_camera_np()
    o = oracle_inputs(image, label, n_classes, k_inv, rt_inv, sdd, h, w, 200)
    t = {k: torch.as_tensor(o[k], dtype=torch.float32, device=DEV).contiguous() for k in ("M", "src", "tgt", "step")}
    vol = image[0, 0].to(DEV)
    packed = pack_volume(vol, "qcoef")  # from the fp32 values
    vol_t = vol.to(dtype)
    lab = label[0, 0].to(torch.int8).to(DEV)
    lab_t = pack_labels(lab, "qcoef") if n_classes > 1 else lab.new_empty(0)
    rec = qcoef_to_coef(qcoef_decode(packed, vol.shape))
    out = fused_raymarch(vol_t, lab_t, t["M"], t["src"], t["tgt"], t["step"], 200, n_classes, 16, packed, "qcoef")
    ref = orm.forward(**o, coef=rec)
    out = out.double().cpu().numpy()
    assert rel_to_max(ref.sum(1), out.sum(1)) < 1e-5
    assert ((np.abs(ref - out).max(axis=1) / np.abs(ref).max()) > 1e-5).mean() < 1e-3
    go = torch.randn(1, n_classes, 256, generator=torch.Generator().manual_seed(3)).to(DEV)
    g = orm.backward(go.double().cpu().numpy(), **o, need_gvol=False, coef=rec)
    _, jac = fused_raymarch_jac(vol_t, lab_t, t["M"], t["src"], t["tgt"], t["step"], 200, n_classes, 16, 7, packed,
                                "qcoef")
    gM1, gsrc1, gtgt1, gstep1 = fused_raymarch_bwd_jac(go, jac, t["M"], t["src"], t["tgt"], t["step"], 16, True)
    _, gM2, _, gtgt2, gstep2 = fused_raymarch_bwd(go, vol_t, lab_t, t["M"], t["src"], t["tgt"], t["step"], 200,
                                                  n_classes, 16, False, True, packed, "qcoef")
    for got in (gM1, gM2):
        assert rel_to_max(g["gM"], got.double().cpu().numpy()) < 1e-4
    for got in (gtgt1, gtgt2):
        assert rel_to_max(g["gtgt"], got.double().cpu().numpy()) < 1e-4
    for got in (gstep1, gstep2):
        assert rel_to_max(g["gstep"], got.double().cpu().numpy()) < 1e-4
