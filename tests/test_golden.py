"""The oracle reproduces the committed golden vectors (tests/golden/ref_camera_32.npz)."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import raymarch as orm
from tests.helpers import oracle_inputs, random_volume

GOLDEN = Path(__file__).resolve().parent / "golden" / "ref_camera_32.npz"
CASES = [("persp_c1", False, 1), ("ortho_c1", True, 1), ("persp_c3", False, 3), ("ortho_c3", True, 3)]


@pytest.mark.parametrize("name,ortho,C", CASES)
def test_oracle_reproduces_golden(name, ortho, C):
    g = np.load(GOLDEN)
    image, label = random_volume(32, C, seed=0)
    np.testing.assert_array_equal(image.numpy()[0, 0], g["volume"])
    inp = oracle_inputs(image, label, C, g["k_inv"], g["rt_inv"], g["sdd"], 16, 16, int(g["n_samples"]), ortho)
    np.testing.assert_allclose(orm.forward(**inp), g[f"{name}_fwd"], rtol=1e-12, atol=0)
    back = orm.backward(g[f"{name}_go"], **inp, need_gvol=False)
    for k in ("gM", "gsrc", "gtgt", "gstep"):
        np.testing.assert_allclose(back[k], g[f"{name}_{k}"], rtol=1e-10, atol=1e-12 * np.abs(g[f"{name}_{k}"]).max())


def test_golden_labels_are_torch_seeded():
    g = np.load(GOLDEN)
    _, label = random_volume(32, 3, seed=0)
    np.testing.assert_array_equal(label.numpy()[0, 0].astype(np.int8), g["label3"])
    assert set(np.unique(g["label3"])) == {0, 1, 2}
    assert torch.rand(1).numel() == 1
