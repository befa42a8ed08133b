"""Float64 numpy restatement of nanodrr's ray march and its analytic adjoint.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Coordinates: `M` (B, 3, 4) maps camera space to voxel-index space (voxel
centres at integers, x -> W, y -> H, z -> D), as `render_fused` builds it
(src/nanodrr/drr/backends.py:93-96). Sample s of ray (b, n) is at
M[b] @ [lerp(src, tgt, s / (S - 1)), 1] (_fused.py:133-142, 190-194).
The value is the zero-padded trilinear interpolation (_fused.py:11-16,
205-220); the label is the nearest voxel's, rounded half to even and zero
padded (_fused.py:28-32, 222-228); the pixel is step * sum over samples
(_fused.py:300, 326-331). Labels outside [0, C) are dropped, the fused
path's behaviour (_fused.py:237; tests/drr/test_fused.py:212-226).

Every function takes `rays`, an optional index array selecting a subset of
the N rays, so full-size workloads can be checked on a sample of rays.
"""

from __future__ import annotations

import numpy as np

_CHUNK = 1 << 20  # samples per vectorised chunk


def make_tgt(k_inv: np.ndarray, sdd: np.ndarray, height: int, width: int) -> np.ndarray:
    """(Bk, N, 3) detector pixel centres in camera space (renderer.py:101-117)."""
    v, u = np.meshgrid(np.arange(height) + 0.5, np.arange(width) + 0.5, indexing="ij")
    uv1 = np.stack([u, v, np.ones_like(u)], axis=-1).reshape(-1, 3)
    return np.asarray(sdd, np.float64)[:, None, None] * np.einsum("bij,nj->bni", np.asarray(k_inv, np.float64), uv1)


def make_src(orthographic: bool, tgt: np.ndarray, sdd: np.ndarray) -> np.ndarray:
    """(B, 1, 3) zeros, or per-ray sources sdd behind the targets (renderer.py:120-132)."""
    if orthographic:
        src = tgt.copy()
        src[..., 2] -= np.asarray(sdd, np.float64)[:, None]
        return src
    return np.zeros((tgt.shape[0], 1, 3))


def step_size(src: np.ndarray, tgt: np.ndarray, n_samples: int) -> np.ndarray:
    """Per-ray Riemann weight |tgt - src| / (S - 1) (renderer.py:82)."""
    return np.linalg.norm(tgt - src, axis=-1) / float(n_samples - 1)


def camera_to_voxel(world_to_voxel: np.ndarray, rt_inv: np.ndarray) -> np.ndarray:
    """(B, 3, 4) fused camera -> voxel affine (backends.py:96)."""
    return (np.asarray(world_to_voxel, np.float64) @ np.asarray(rt_inv, np.float64))[:, :3, :]


def _ray_endpoints(src, tgt, b, rays):
    sb = src[b if src.shape[0] > 1 else 0]
    tb = tgt[b if tgt.shape[0] > 1 else 0]
    tb = tb[rays]
    sb = np.broadcast_to(sb[rays] if sb.shape[0] > 1 else sb, tb.shape)
    return sb, tb


def _step_rows(step, b, rays):
    return step[b if step.shape[0] > 1 else 0][rays]


def _corners(vol: np.ndarray, p: np.ndarray):
    """Zero-padded 2x2x2 stencil at voxel positions p (..., 3)."""
    D, H, W = vol.shape
    fl = np.floor(p)
    f = p - fl
    i = fl.astype(np.int64)
    vals = {}
    for cz in (0, 1):
        for cy in (0, 1):
            for cx in (0, 1):
                jx, jy, jz = i[..., 0] + cx, i[..., 1] + cy, i[..., 2] + cz
                inb = (jx >= 0) & (jx < W) & (jy >= 0) & (jy < H) & (jz >= 0) & (jz < D)
                v = vol[np.clip(jz, 0, D - 1), np.clip(jy, 0, H - 1), np.clip(jx, 0, W - 1)]
                vals[(cz, cy, cx)] = np.where(inb, v, 0.0)
    return vals, f, i


def coef_records(vol: np.ndarray, dtype=np.float32) -> np.ndarray:
    """(D+1, H+1, W+1, 8) coefficient records of the zero-padded volume, one per cell (z0, y0, x0)
    in [-1, D-1] x [-1, H-1] x [-1, W-1]: [c0, cx, cy, cxy, cz, cxz, cyz, cxyz] of the cell's trilinear
    interpolant, formed with the arithmetic of `dtype` in the order ndrr_pack_coef documents
    (float32 gives the packer's fp32 values before their final rounding; float64 the exact ones)."""
    vol = np.asarray(vol, dtype)
    D, H, W = vol.shape
    P = np.zeros((D + 2, H + 2, W + 2), dtype)
    P[1:-1, 1:-1, 1:-1] = vol
    v = [P[a:a + D + 1, b:b + H + 1, c:c + W + 1] for a in (0, 1) for b in (0, 1) for c in (0, 1)]  # v[4a+2b+c]
    dx00, dx01, dx10, dx11 = v[1] - v[0], v[3] - v[2], v[5] - v[4], v[7] - v[6]
    cy = v[2] - v[0]
    out = [v[0], dx00, cy, dx01 - dx00, v[4] - v[0], dx10 - dx00, (v[6] - v[4]) - cy, (dx11 - dx10) - (dx01 - dx00)]
    return np.stack(out, axis=-1)


def _coef_eval(coef: np.ndarray, p: np.ndarray):
    """Value and gradient (voxel units) at positions p (..., 3) from coefficient records (float64)."""
    D1, H1, W1 = coef.shape[:3]
    fl = np.floor(p)
    f = p - fl
    i = fl.astype(np.int64)
    ok = (i[..., 0] >= -1) & (i[..., 0] < W1 - 1) & (i[..., 1] >= -1) & (i[..., 1] < H1 - 1) & \
        (i[..., 2] >= -1) & (i[..., 2] < D1 - 1)
    rec = coef[np.clip(i[..., 2] + 1, 0, D1 - 1), np.clip(i[..., 1] + 1, 0, H1 - 1), np.clip(i[..., 0] + 1, 0, W1 - 1)]
    rec = np.where(ok[..., None], np.asarray(rec, np.float64), 0.0)
    c0, cx, cy, cxy, cz, cxz, cyz, cxyz = (rec[..., k] for k in range(8))
    fx, fy, fz = f[..., 0], f[..., 1], f[..., 2]
    val = c0 + cx * fx + cy * fy + cxy * fx * fy + cz * fz + cxz * fx * fz + cyz * fy * fz + cxyz * fx * fy * fz
    gx = cx + cxy * fy + cxz * fz + cxyz * fy * fz
    gy = cy + cxy * fx + cyz * fz + cxyz * fx * fz
    gz = cz + cxz * fx + cyz * fy + cxyz * fx * fy
    return val, gx, gy, gz


def _lerp(vals, f):
    fx, fy, fz = f[..., 0], f[..., 1], f[..., 2]
    c00 = vals[0, 0, 0] * (1 - fx) + vals[0, 0, 1] * fx
    c01 = vals[0, 1, 0] * (1 - fx) + vals[0, 1, 1] * fx
    c10 = vals[1, 0, 0] * (1 - fx) + vals[1, 0, 1] * fx
    c11 = vals[1, 1, 0] * (1 - fx) + vals[1, 1, 1] * fx
    return (c00 * (1 - fy) + c01 * fy) * (1 - fz) + (c10 * (1 - fy) + c11 * fy) * fz


def _labels(lab: np.ndarray, p: np.ndarray) -> np.ndarray:
    """Nearest label, round half to even, class 0 outside the volume; lab holds trunc'd classes."""
    D, H, W = lab.shape
    r = np.rint(p).astype(np.int64)
    jx, jy, jz = r[..., 0], r[..., 1], r[..., 2]
    inb = (jx >= 0) & (jx < W) & (jy >= 0) & (jy < H) & (jz >= 0) & (jz < D)
    v = lab[np.clip(jz, 0, D - 1), np.clip(jy, 0, H - 1), np.clip(jx, 0, W - 1)]
    return np.where(inb, v, 0)


def label_classes(label: np.ndarray) -> np.ndarray:
    """Class index of a float label volume as the fused kernel reads it: trunc toward zero."""
    return np.trunc(np.asarray(label, np.float64)).astype(np.int64)


def _iter_chunks(n_rays: int, S: int):
    per = max(1, _CHUNK // max(S, 1))
    for lo in range(0, n_rays, per):
        yield lo, min(n_rays, lo + per)


def _sample_points(M_b, s_rows, t_rows, S):
    t = np.arange(S, dtype=np.float64) / (S - 1)
    e = t_rows - s_rows
    q = s_rows[:, None, :] + t[None, :, None] * e[:, None, :]
    p = q @ M_b[:, :3].T + M_b[:, 3]
    return t, q, p


def forward(vol, lab, n_classes, M, src, tgt, step, n_samples, rays=None, coef=None) -> np.ndarray:
    """(B, C, R) line integrals in float64 (R = len(rays) or N). With `coef` (coefficient records,
    `coef_records`), the interpolant of each cell is the records' polynomial (vol only fixes the shape)."""
    vol = np.asarray(vol, np.float64)
    M = np.asarray(M, np.float64)
    src, tgt, step = (np.asarray(a, np.float64) for a in (src, tgt, step))
    B, N = M.shape[0], tgt.shape[1]
    rays = np.arange(N) if rays is None else np.asarray(rays)
    C = n_classes
    out = np.zeros((B, C, len(rays)))
    for b in range(B):
        s_all, t_all = _ray_endpoints(src, tgt, b, rays)
        st = _step_rows(step, b, rays)
        for lo, hi in _iter_chunks(len(rays), n_samples):
            _, _, p = _sample_points(M[b], s_all[lo:hi], t_all[lo:hi], n_samples)
            if coef is None:
                vals, f, _ = _corners(vol, p)
                val = _lerp(vals, f)
            else:
                val = _coef_eval(coef, p)[0]
            if C == 1:
                out[b, 0, lo:hi] = val.sum(axis=1) * st[lo:hi]
            else:
                cls = _labels(lab, p)
                for c in range(C):
                    out[b, c, lo:hi] = np.where(cls == c, val, 0.0).sum(axis=1) * st[lo:hi]
    return out


def backward(go, vol, lab, n_classes, M, src, tgt, step, n_samples, rays=None, need_gvol=True, coef=None):
    """Analytic adjoint of `forward` (restates _fused.py:230-298 in float64).

    go: (B, C, R). Returns dict with gM (B, 3, 4), gsrc/gtgt (B, R, 3) per ray,
    gstep (B, R) and gvol (D, H, W) (None unless need_gvol). With `coef`, values and
    gradients come from the coefficient records (no volume gradient then).
    """
    if coef is not None and need_gvol:
        raise ValueError("the volume gradient is defined for the corner form only")
    vol = np.asarray(vol, np.float64)
    M = np.asarray(M, np.float64)
    go = np.asarray(go, np.float64)
    src, tgt, step = (np.asarray(a, np.float64) for a in (src, tgt, step))
    B, N = M.shape[0], tgt.shape[1]
    rays = np.arange(N) if rays is None else np.asarray(rays)
    R, C = len(rays), n_classes
    gM = np.zeros((B, 3, 4))
    gsrc = np.zeros((B, R, 3))
    gtgt = np.zeros((B, R, 3))
    gstep = np.zeros((B, R))
    gvol = np.zeros(vol.shape) if need_gvol else None
    D, H, W = vol.shape
    for b in range(B):
        s_all, t_all = _ray_endpoints(src, tgt, b, rays)
        st = _step_rows(step, b, rays)
        for lo, hi in _iter_chunks(R, n_samples):
            t, q, p = _sample_points(M[b], s_all[lo:hi], t_all[lo:hi], n_samples)
            vals, f, i = _corners(vol, p)
            val = _lerp(vals, f)
            if C == 1:
                gg = np.broadcast_to(go[b, 0, lo:hi][:, None], val.shape)
            else:
                cls = _labels(lab, p)
                ok = (cls >= 0) & (cls < C)
                gg = np.where(ok, go[b, np.clip(cls, 0, C - 1), np.arange(lo, hi)[:, None]], 0.0)
            g = gg * st[lo:hi][:, None]
            fx, fy, fz = f[..., 0], f[..., 1], f[..., 2]
            if coef is not None:
                val, dvx, dvy, dvz = _coef_eval(coef, p)
            else:
                v = vals
                dvx = ((v[0, 0, 1] - v[0, 0, 0]) * (1 - fy) + (v[0, 1, 1] - v[0, 1, 0]) * fy) * (1 - fz) + (
                    (v[1, 0, 1] - v[1, 0, 0]) * (1 - fy) + (v[1, 1, 1] - v[1, 1, 0]) * fy) * fz
                dvy = ((v[0, 1, 0] - v[0, 0, 0]) * (1 - fx) + (v[0, 1, 1] - v[0, 0, 1]) * fx) * (1 - fz) + (
                    (v[1, 1, 0] - v[1, 0, 0]) * (1 - fx) + (v[1, 1, 1] - v[1, 0, 1]) * fx) * fz
                dvz = ((v[1, 0, 0] - v[0, 0, 0]) * (1 - fx) + (v[1, 0, 1] - v[0, 0, 1]) * fx) * (1 - fy) + (
                    (v[1, 1, 0] - v[0, 1, 0]) * (1 - fx) + (v[1, 1, 1] - v[0, 1, 1]) * fx) * fy
            gp = np.stack([g * dvx, g * dvy, g * dvz], axis=-1)  # (R, S, 3) dL/dp
            qh = np.concatenate([q, np.ones(q.shape[:-1] + (1,))], axis=-1)
            gM[b] += np.einsum("rsi,rsj->ij", gp, qh)
            gq = gp @ M[b][:, :3]  # M33^T gp
            gsrc[b, lo:hi] = ((1 - t)[None, :, None] * gq).sum(axis=1)
            gtgt[b, lo:hi] = (t[None, :, None] * gq).sum(axis=1)
            gstep[b, lo:hi] = (gg * val).sum(axis=1)
            if need_gvol:
                w = {0: (1 - fx, 1 - fy, 1 - fz), 1: (fx, fy, fz)}
                for cz in (0, 1):
                    for cy in (0, 1):
                        for cx in (0, 1):
                            jx, jy, jz = i[..., 0] + cx, i[..., 1] + cy, i[..., 2] + cz
                            inb = (jx >= 0) & (jx < W) & (jy >= 0) & (jy < H) & (jz >= 0) & (jz < D)
                            wt = w[cx][0] * w[cy][1] * w[cz][2]
                            flat = (jz * H + jy) * W + jx
                            np.add.at(gvol.reshape(-1), flat[inb], (g * wt)[inb])
    return {"gM": gM, "gsrc": gsrc, "gtgt": gtgt, "gstep": gstep, "gvol": gvol}


def inside_count(M, src, tgt, n_samples, shape, rays=None) -> np.ndarray:
    """(B, R) number of samples strictly inside (-1, W) x (-1, H) x (-1, D),
    i.e. with at least one trilinear corner in the volume (brute force)."""
    D, H, W = shape
    M = np.asarray(M, np.float64)
    src, tgt = np.asarray(src, np.float64), np.asarray(tgt, np.float64)
    B, N = M.shape[0], tgt.shape[1]
    rays = np.arange(N) if rays is None else np.asarray(rays)
    out = np.zeros((B, len(rays)), np.int64)
    hi = np.array([W, H, D], np.float64)
    for b in range(B):
        s_all, t_all = _ray_endpoints(src, tgt, b, rays)
        for lo, up in _iter_chunks(len(rays), n_samples):
            _, _, p = _sample_points(M[b], s_all[lo:up], t_all[lo:up], n_samples)
            ins = np.all((p > -1.0) & (p < hi), axis=-1)
            out[b, lo:up] = ins.sum(axis=1)
    return out


def discontinuity_distance(M, src, tgt, n_samples, shape, rays=None, labels=True) -> np.ndarray:
    """(B, R) smallest distance (voxels) over a ray's samples that touch the volume from a position
    where the march is discontinuous: a cell face (integer coordinate), where the trilinear
    derivative jumps, and, with labels, a nearest-label tie (half-integer coordinate), where the
    class jumps. An fp32 march places samples within ~1 ulp of the float64 positions, so rays
    whose distance exceeds a few ulps see the same cells and classes in both."""
    D, H, W = shape
    M = np.asarray(M, np.float64)
    src, tgt = np.asarray(src, np.float64), np.asarray(tgt, np.float64)
    B, N = M.shape[0], tgt.shape[1]
    rays = np.arange(N) if rays is None else np.asarray(rays)
    out = np.full((B, len(rays)), np.inf)
    hi = np.array([W, H, D], np.float64)
    for b in range(B):
        s_all, t_all = _ray_endpoints(src, tgt, b, rays)
        for lo, up in _iter_chunks(len(rays), n_samples):
            _, _, p = _sample_points(M[b], s_all[lo:up], t_all[lo:up], n_samples)
            ins = np.all((p > -1.0) & (p < hi), axis=-1)
            step = 0.5 if labels else 1.0
            q = p / step
            dist = np.abs(q - np.rint(q)).min(axis=-1) * step
            out[b, lo:up] = np.where(ins, dist, np.inf).min(axis=1)
    return out
