"""CPU oracle of nanodrr's fused DRR ray march — TEST INFRASTRUCTURE ONLY.

Nothing in the product (`nanodrr_amd/`) imports this package. Only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may use
it, and only as the checker (or, for the baseline, as the timed pure-PyTorch
CPU backend), never as the code path being measured or shipped.

Modules
  raymarch   float64 numpy restatement of the forward march and its analytic
             adjoint (src/nanodrr/drr/_fused.py:133-300, backends.py:9-63,
             renderer.py:71-132), plus the exact ray-box sample count used
             for roofline bytes.
  torch_ref  pure-PyTorch restatement of the reference `render_torch`
             (grid_sample path, backends.py:9-63), dtype generic: float64 as
             an independent cross-check of `raymarch`, float32 as the CPU
             baseline.
  geometry   float64 numpy camera/pose helpers (intrinsics.py:53-68,
             extrinsics.py:7-57, se3.py:179-221).
  registration float64 numpy pose step of the registration loop
             (registration.py:64-68) with its adjoint, ZNCC with its
             gradient (metrics/ncc.py:8-37) and the window (patch) ZNCC
             built from it (metrics/ncc.py:83-88).

Parity pinning (tests/test_oracle.py, tests/test_geometry.py, tests/test_registration.py):
  raymarch.forward     unit-impulse known answer (tests/drr/test_renderer.py:52-91),
                       constant-volume chord length, torch grid_sample float64,
                       channel-sum invariant
  raymarch.backward    torch autograd of grid_sample (float64), central finite
                       differences, adjoint dot-product identity for the
                       volume gradient
  raymarch.inside_count brute force over every sample
  geometry             scipy Rotation.from_euler("ZXY"), matrix exponential,
                       round trips (tests/camera/test_extrinsics.py:11-54)
  registration         scipy Rotation.from_rotvec, torch float64 autograd of
                       matrix_exp, central finite differences (pose step);
                       scipy.stats.pearsonr, affine invariance, finite
                       differences (ZNCC); one full-size window equals the
                       whole-image ZNCC, a brute-force loop over windows and
                       finite differences (window ZNCC)
No function here is "parity unpinned".
"""
