"""The labelled coefficient records ("lcoef", NDRR_LAYOUT_LCOEF) as the tests emulate them
(tests.helpers.lcoef_expected / lcoef_decode), pinned on the CPU against what the format must
give: the class a record names for a sample is the nearest-label class of the oracle
(oracle.raymarch._labels, round half to even, class 0 outside), the decoded coefficients are the
exact ones to the slots' rounding, and the scale is the largest power of two that keeps every slot
below 2^-7 (so the two class bits of each slot are free)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import raymarch as orm
from tests.helpers import LCOEF_TOP, lcoef_codes, lcoef_decode, lcoef_expected, lcoef_scale


def _packed(slots: torch.Tensor, s: float) -> torch.Tensor:
    hdr = np.zeros(64, np.uint32)
    hdr[1] = np.array([s], np.float32).view(np.uint32)[0]
    hdr[2] = np.array([1.0 / s], np.float32).view(np.uint32)[0]
    return torch.cat([torch.from_numpy(hdr.view(np.uint8)), slots.reshape(-1).view(torch.uint8)])


def _case(seed, shape=(6, 7, 9), C=3, magnitude=1.0):
    rng = np.random.default_rng(seed)
    vol = (np.abs(rng.standard_normal(shape)) * magnitude).astype(np.float32)
    vol[1] = 0.0
    lab = rng.integers(-1 if C < 4 else 0, C, size=shape)
    return vol, lab


@pytest.mark.parametrize("C", [2, 3, 4])
def test_record_codes_give_the_nearest_label(C):
    """At random positions (and on cell faces and half-integer ties) the code in the sample's
    record, at the corner rint(p) - floor(p), is the oracle's nearest label (lab & 3; -1 -> code 3,
    which no class below 3 takes)."""
    vol, lab = _case(C, C=C)
    codes = lcoef_codes(lab)
    D, H, W = lab.shape
    rng = np.random.default_rng(10 + C)
    lo, hi = np.full(3, -1.5), np.array([W, H, D]) + 0.5
    p = lo + rng.random((4000, 3)) * (hi - lo)
    p[:500] = np.round(p[:500] * 2) / 2  # ties and faces
    fl = np.floor(p).astype(np.int64)
    ok = (fl >= -1).all(axis=1) & (fl[:, 0] < W) & (fl[:, 1] < H) & (fl[:, 2] < D)
    p, fl = p[ok], fl[ok]
    o = np.rint(p).astype(np.int64) - fl
    assert ((o == 0) | (o == 1)).all()
    j = o[:, 0] + 2 * o[:, 1] + 4 * o[:, 2]
    got = codes[fl[:, 2] + 1, fl[:, 1] + 1, fl[:, 0] + 1, j]
    want = orm._labels(lab, p) & 3
    assert len(p) > 2000 and np.array_equal(got, want)
    for c in range(C):  # routing: code == class for every class, never for a dropped voxel
        assert np.array_equal(got == c, orm._labels(lab, p) == c)


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("magnitude", [1e-4, 0.05, 3.0, 2e4])
def test_records_round_trip_to_the_exact_coefficients(dtype, magnitude):
    """Decoded records give back the class codes and the float64 coefficients of the volume to
    the slots' rounding: fp16 slots to half an ulp (2^-11 relative) or, as subnormals, 2^-25 / s;
    a bf16 c0 to 2^-8 relative; an fp16 c0 like the other fp16 slots."""
    vol, lab = _case(1, C=3, magnitude=magnitude)
    slots, s, m = lcoef_expected(vol, lab, dtype)
    coef, codes = lcoef_decode(_packed(slots, s), dtype, vol.shape)
    assert np.array_equal(codes, lcoef_codes(lab))
    exact = orm.coef_records(vol.astype(np.float64), np.float64)
    err = np.abs(coef - exact)
    bound = np.maximum(np.abs(exact) * 2.0 ** -11, 2.0 ** -25 / s) * 1.001 + np.abs(exact) * 2.0 ** -22
    if dtype == torch.bfloat16:
        bound[..., 0] = np.abs(exact[..., 0]) * 2.0 ** -8 * 1.001
    assert (err <= bound).all()


@pytest.mark.parametrize("magnitude", [1e-30, 1e-4, 3.0, 2e4, 1e30])
def test_scale_is_the_largest_that_frees_the_class_bits(magnitude):
    """Every scaled coefficient stays at or below 2^-7 - 2^-18 (fp16 exponent field <= 7, so its two
    top exponent bits are free), and twice the scale would break that."""
    vol, lab = _case(2, magnitude=magnitude)
    coef = orm.coef_records(vol, np.float32)
    s, m = lcoef_scale(coef)
    assert m == np.abs(coef).max() and s == 2.0 ** np.round(np.log2(s))
    assert m * s <= LCOEF_TOP < m * 2 * s
    scaled = torch.from_numpy(coef * np.float32(s)).to(torch.float16).view(torch.int16).to(torch.int32) & 0x7FFF
    assert (scaled >> 10).max() <= 7  # exponent fields
