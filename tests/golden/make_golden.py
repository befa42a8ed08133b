"""Write the golden vectors of the ray march from the pinned float64 oracle.

Run from the repository root:  python -m tests.golden.make_golden
Inputs follow nanodrr's own test set-up (tests/drr/test_fused.py:11-47): a
seeded 32^3 torch.rand volume with 1 mm voxels centred at the origin, the
16 x 16 detector at sdd 100 mm, pose make_rt_inv([[5,-3,8]], [[1,80,-2]], "AP"),
200 samples per ray. Outputs: the forward image and the analytic adjoint
(dL/dM, per-ray endpoint and step gradients) for a seeded upstream gradient.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from oracle import raymarch as orm
from tests.helpers import oracle_inputs, random_volume, ref_camera_np

OUT = Path(__file__).resolve().parent / "ref_camera_32.npz"
CASES = [("persp_c1", False, 1), ("ortho_c1", True, 1), ("persp_c3", False, 3), ("ortho_c3", True, 3)]


def main() -> None:
    arrays = {}
    k_inv, rt_inv, sdd, h, w = ref_camera_np()
    arrays.update(k_inv=k_inv, rt_inv=rt_inv, sdd=sdd, n_samples=np.array(200))
    for name, ortho, C in CASES:
        image, label = random_volume(32, C, seed=0)
        inp = oracle_inputs(image, label, C, k_inv, rt_inv, sdd, h, w, 200, ortho)
        arrays["volume"] = image.numpy()[0, 0]  # the same draw for every case (seed 0)
        if C > 1:
            arrays["label3"] = label.numpy()[0, 0].astype(np.int8)
        out = orm.forward(**inp)
        go = np.random.default_rng(11).standard_normal(out.shape)
        g = orm.backward(go, **inp, need_gvol=False)
        arrays[f"{name}_fwd"] = out
        arrays[f"{name}_go"] = go
        arrays[f"{name}_gM"] = g["gM"]
        arrays[f"{name}_gsrc"] = g["gsrc"]
        arrays[f"{name}_gtgt"] = g["gtgt"]
        arrays[f"{name}_gstep"] = g["gstep"]
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
