/*
 * nanodrr_amd — C ABI of the MI355X (gfx950) fused differentiable DRR ray march.
 *
 * Every entry point takes raw device pointers, element counts and a HIP stream
 * handle (passed as `void*`; NULL means the legacy default stream). Nothing in
 * these signatures depends on PyTorch. Launches are asynchronous on `stream`;
 * no entry point synchronises the host or allocates device memory. Callers own
 * every buffer.
 *
 * Return value: 0 on success, a negative NDRR_E* code for an argument error
 * detected before any launch, or a positive hipError_t from the HIP runtime.
 * `ndrr_strerror` turns either into text.
 *
 * Geometry conventions (all row-major, fp32):
 *   M    [B][3][4]   camera space -> voxel-index space (voxel centres are the
 *                    integers; x indexes W, y indexes H, z indexes D)
 *   src  ray sources in camera space, element (b, n) at
 *        src[b * src_batch_stride + n * src_ray_stride + {0,1,2}]
 *        (perspective rays share one source per pose: src_ray_stride = 0)
 *   tgt  ray targets, element (b, n) at tgt[b * tgt_batch_stride + 3 n + {0,1,2}]
 *        (tgt_batch_stride = 0 shares one detector across the pose batch)
 *   step per-ray Riemann weight, element (b, n) at step[b * step_batch_stride + n]
 *   out  [B][C][N] fp32 line integrals
 *   vol  the volume in NDRR_F32 / NDRR_BF16 / NDRR_F16, laid out as
 *        NDRR_LAYOUT_LINEAR:  [D][H][W], W fastest, or a packed copy made by
 *        ndrr_pack_volume / ndrr_pack_coef / ndrr_pack_coef_labels:
 *        NDRR_LAYOUT_QUAD, NDRR_LAYOUT_STENCIL, NDRR_LAYOUT_COEF,
 *        NDRR_LAYOUT_LCOEF or NDRR_LAYOUT_QCOEF records (vol_dtype is then
 *        the record type; for NDRR_LAYOUT_QCOEF, whose records are fp16,
 *        the 16-bit type of the volume they were made for)
 *   lab  int8 class per voxel (values outside [0, C) are dropped) in the form
 *        vol_layout reads: the [D][H][W] map for NDRR_LAYOUT_LINEAR, its padded
 *        copy for NDRR_LAYOUT_STENCIL and NDRR_LAYOUT_COEF and its label quads
 *        for NDRR_LAYOUT_QUAD and NDRR_LAYOUT_QCOEF (both by ndrr_pack_labels); ignored (may be NULL)
 *        for NDRR_LAYOUT_LCOEF, whose records carry the classes, and when
 *        n_classes == 1
 * Sample s of ray (b, n) sits at M[b] @ [lerp(src, tgt, s / (S - 1)), 1].
 * Rays are mapped to detector tiles of a `width`-wide image when width > 0
 * (N must be a multiple of width), else to runs of consecutive rays; see
 * ndrr_launch_tile. Packed volume and label copies must be smaller than 2 GiB.
 */
#ifndef NANODRR_AMD_H
#define NANODRR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Volume element types. */
enum {
  NDRR_F32 = 0,
  NDRR_BF16 = 1,
  NDRR_F16 = 2
};

/* Volume layouts. */
enum {
  NDRR_LAYOUT_LINEAR = 0,  /* [D][H][W] elements */
  NDRR_LAYOUT_STENCIL = 1, /* one record of 8 trilinear corners per cell */
  NDRR_LAYOUT_QUAD = 2,    /* one record of a (y, z) corner quad per x */
  NDRR_LAYOUT_COEF = 3,    /* one record of 8 multilinear coefficients per cell */
  NDRR_LAYOUT_LCOEF = 4,   /* the same with its 8 corners' classes (<= 4 classes) */
  NDRR_LAYOUT_QCOEF = 5    /* one record of a (y, z) quad's bilinear coefficients per x */
};

/* Argument errors (negative return codes). */
enum {
  NDRR_OK = 0,
  NDRR_EINVAL = -1,  /* null pointer, non-positive size, S < 2, bad width */
  NDRR_EDTYPE = -2,  /* unknown volume element type */
  NDRR_ECLASSES = -3,/* n_classes outside [1, NDRR_MAX_CLASSES] */
  NDRR_ERANGE = -4,  /* a size exceeds the kernel's 32-bit index range */
  NDRR_ESCRATCH = -5 /* the work needs more than the fixed scratch array holds */
};

#define NDRR_MAX_CLASSES 8

/* Library version as major * 10000 + minor * 100 + patch. */
int ndrr_version(void);

/* Text for a return code of any entry point below. Never NULL. */
const char* ndrr_strerror(int code);

/*
 * Forward ray march: out[b][c][n] = step(b,n) * sum_s trilinear(vol, p_s) * [lab(p_s) == c]
 * with zero padding outside the volume and nearest (round-half-even) label
 * lookup. Rays with n_samples >= 512 accumulate with Neumaier-compensated
 * fp32 sums. Output is fp32 whatever the volume type.
 * Replaces the forward launch of nanodrr's fused op:
 *   src/nanodrr/drr/_fused.py:415-460 (fused_raymarch) and the
 *   BACKWARD=0 specialisation of _fused_kernel, _fused.py:36-331.
 */
int ndrr_raymarch_fwd(const void* vol, int vol_dtype, int vol_layout, const int8_t* lab,
                      const float* M, const float* src, const float* tgt,
                      const float* step, float* out,
                      int B, int N, int D, int H, int W,
                      int n_samples, int n_classes, int width,
                      int64_t src_batch_stride, int64_t src_ray_stride,
                      int64_t tgt_batch_stride, int64_t step_batch_stride,
                      void* stream);

/*
 * The launch tile: each workgroup marches a tile_width x tile_height patch of
 * the detector (width > 0) or tile_rays consecutive rays (width == 0).
 * Pure host query; any pointer may be NULL.
 */
void ndrr_launch_tile(int* tile_width, int* tile_height, int* tile_rays);

/*
 * Number of fp32 elements of the pose-gradient workspace that
 * ndrr_raymarch_bwd needs for B poses of N rays at this width
 * (12 per pose per detector tile, or per wave of every tile, as built).
 * Pure host arithmetic.
 * The workspace must be 16-byte aligned (hipMalloc memory always is).
 */
size_t ndrr_bwd_workspace_floats(int B, int N, int width);

/*
 * Analytic adjoint of ndrr_raymarch_fwd for an upstream gradient
 * go[B][C][N] (fp32):
 *   gM[b][3][4]  = sum over the pose's rays and samples of
 *                  g * grad_p trilinear (x) [q; 1], g = go[b][cls][n] * step,
 *                  q = lerp(src, tgt, t) — deterministic: per-tile partials in
 *                  `workspace` (ndrr_bwd_workspace_floats) reduced in fixed order;
 *   gsrc, gtgt   [B][N][3] per-ray gradients of the ray endpoints, and
 *   gstep        [B][N], written when need_gcam != 0 (else may be NULL);
 *   gvol         [D][H][W] fp32, accumulated with atomics when gvol != NULL
 *                (the caller zeroes it; the result is not bitwise reproducible).
 * Replaces the backward launch of nanodrr's fused op:
 *   src/nanodrr/drr/_fused.py:477-531 (_fused_raymarch_bwd) and the
 *   BACKWARD=1 specialisation of _fused_kernel, _fused.py:161-186,230-325.
 */
int ndrr_raymarch_bwd(const float* go, const void* vol, int vol_dtype,
                      int vol_layout, const int8_t* lab, const float* M, const float* src,
                      const float* tgt, const float* step,
                      float* workspace, float* gM,
                      float* gsrc, float* gtgt, float* gstep, float* gvol,
                      int need_gcam,
                      int B, int N, int D, int H, int W,
                      int n_samples, int n_classes, int width,
                      int64_t src_batch_stride, int64_t src_ray_stride,
                      int64_t tgt_batch_stride, int64_t step_batch_stride,
                      void* stream);

/*
 * ndrr_raymarch_bwd with the volume gradient's scatter going through a brick
 * buffer: gvol_bricks holds ndrr_gvol_brick_floats(D, H, W) fp32 values
 * (16-byte aligned), zero on entry: 4 x 4 x 4-voxel bricks, brick
 * (bx, by, bz) at ((bz * ceil(H/4) + by) * ceil(W/4) + bx) * 64, voxel
 * (x, y, z) at (z % 4) * 16 + (y % 4) * 4 + x % 4 within its brick. The march's
 * atomics add there (a 128-byte line then holds a 4 x 4 x 2 block of voxels,
 * so the voxels one wave touches at one step fill about half as many lines as
 * the [D][H][W] rows do), then a second launch adds the bricks into gvol
 * (which must not be NULL) and clears them: the buffer is zero again on
 * return, ready for the next call on the same stream. Results equal
 * ndrr_raymarch_bwd's up to the order of the fp32 atomic additions.
 * No counterpart in the reference (an internal layout of its
 * _fused.py:284-298 volume-gradient scatter).
 */
size_t ndrr_gvol_brick_floats(int D, int H, int W);
int ndrr_raymarch_bwd_bricked(const float* go, const void* vol, int vol_dtype,
                              int vol_layout, const int8_t* lab, const float* M,
                              const float* src, const float* tgt, const float* step,
                              float* workspace, float* gM,
                              float* gsrc, float* gtgt, float* gstep, float* gvol,
                              float* gvol_bricks, int need_gcam,
                              int B, int N, int D, int H, int W,
                              int n_samples, int n_classes, int width,
                              int64_t src_batch_stride, int64_t src_ray_stride,
                              int64_t tgt_batch_stride, int64_t step_batch_stride,
                              void* stream);

/*
 * Single-pass adjoint: the forward and backward above as one volume march
 * plus a per-ray contraction, for callers that need pose (and camera)
 * gradients but not the volume gradient. Together they replace the pair
 * fused_raymarch + fused_raymarch_bwd (src/nanodrr/drr/_fused.py:349-460,
 * 477-531) when need_gvol is false; the reference has no split like this.
 *
 * ndrr_raymarch_fwd_jac writes `out` exactly as ndrr_raymarch_fwd and, in the
 * same march, jac[B][jac_planes][C][N] (fp32): per ray and class c the sums
 * over the samples routed to c of grad v (planes 0-2, voxel units), of
 * t * grad v (planes 3-5) and, with jac_planes == 7, of v (plane 6).
 * jac == NULL makes it ndrr_raymarch_fwd. lattice_offset, if not NULL, is
 * [B] fp32: pose b's samples sit at t = (s + lattice_offset[b]) / (S - 1)
 * along src -> tgt instead of s / (S - 1) (the per-problem lattice jitter of
 * registration, without shifted copies of the rays); the step is unchanged
 * and the t sums use the shifted t, so ndrr_raymarch_bwd_jac needs no offset.
 *
 * ndrr_raymarch_bwd_jac then gives gM (and gsrc, gtgt, gstep when
 * need_gcam != 0, which needs jac_planes == 7) for an upstream gradient
 * go[B][C][N] (go_batch_stride C * N, or 0 for one [C][N] gradient shared by
 * every pose) from jac alone, without reading the volume; the geometry
 * arguments must be those of the forward. The same workspace as
 * ndrr_raymarch_bwd; deterministic.
 */
int ndrr_raymarch_fwd_jac(const void* vol, int vol_dtype, int vol_layout,
                          const int8_t* lab, const float* M, const float* src,
                          const float* tgt, const float* step,
                          const float* lattice_offset, float* out,
                          float* jac, int jac_planes,
                          int B, int N, int D, int H, int W,
                          int n_samples, int n_classes, int width,
                          int64_t src_batch_stride, int64_t src_ray_stride,
                          int64_t tgt_batch_stride, int64_t step_batch_stride,
                          void* stream);
int ndrr_raymarch_bwd_jac(const float* go, int64_t go_batch_stride,
                          const float* jac, int jac_planes,
                          const float* M, const float* src, const float* tgt,
                          const float* step, float* workspace, float* gM,
                          float* gsrc, float* gtgt, float* gstep, int need_gcam,
                          int B, int N, int n_classes, int width,
                          int64_t src_batch_stride, int64_t src_ray_stride,
                          int64_t tgt_batch_stride, int64_t step_batch_stride,
                          void* stream);

/*
 * The same pair with the camera -> voxel composition folded in: in place of
 * M they take world_to_voxel (4 x 4 float64, row-major) and rt_inv
 * ([B][4][4] fp32), and every workgroup composes its pose's
 * M = (world_to_voxel @ rt_inv[b])[:3] in FP64, rounded once to fp32, with
 * the sums of ndrr_camera_to_voxel; the backward's pose reduction returns
 * g_rt_inv ([B][4][4] fp32) = world_to_voxel[:3]^T gM in FP64 from the
 * rounded gM, with the sums of ndrr_camera_to_voxel_bwd. Results are those
 * of ndrr_camera_to_voxel + ndrr_raymarch_fwd_jac and ndrr_raymarch_bwd_jac +
 * ndrr_camera_to_voxel_bwd, bit for bit, in two launches fewer. Replaces the
 * composition at src/nanodrr/drr/backends.py:96 together with the ops of
 * _fused.py:415-460,477-531, and its autograd.
 */
int ndrr_raymarch_fwd_jac_pose(const void* vol, int vol_dtype, int vol_layout,
                               const int8_t* lab, const double* world_to_voxel,
                               const float* rt_inv, const float* src,
                               const float* tgt, const float* step,
                               const float* lattice_offset, float* out,
                               float* jac, int jac_planes, int B, int N, int D,
                               int H, int W, int n_samples, int n_classes, int width,
                               int64_t src_batch_stride, int64_t src_ray_stride,
                               int64_t tgt_batch_stride, int64_t step_batch_stride,
                               void* stream);
int ndrr_raymarch_bwd_jac_pose(const float* go, int64_t go_batch_stride,
                               const float* jac, int jac_planes,
                               const double* world_to_voxel, const float* rt_inv,
                               const float* src, const float* tgt, const float* step,
                               float* workspace, float* g_rt_inv, float* gsrc,
                               float* gtgt, float* gstep, int need_gcam, int B, int N,
                               int n_classes, int width, int64_t src_batch_stride,
                               int64_t src_ray_stride, int64_t tgt_batch_stride,
                               int64_t step_batch_stride, void* stream);

/*
 * ndrr_raymarch_bwd_jac with registration's pose step folded into its pose
 * reduction: in place of gM it returns the parameter gradients grot, gxyz
 * ([B] x 3 at param_stride, as ndrr_pose_compose_bwd) of grad_scale * gM
 * (grad_scale +1 or -1: the descent direction of a score, exactly), with the
 * sums of ndrr_pose_compose_bwd, bit for bit, in two launches fewer (the
 * scaling and the adjoint). No camera gradients. Replaces the autograd of
 * Registration.pose (src/nanodrr/registration/registration.py:64-68) after
 * the render's backward (_fused.py:477-531).
 */
int ndrr_raymarch_bwd_jac_se3(const float* go, int64_t go_batch_stride,
                              const float* jac, int jac_planes, const float* M,
                              const float* src, const float* tgt, const float* step,
                              float* workspace, const double* prefix,
                              const double* suffix, const float* rot, const float* xyz,
                              int64_t param_stride, float grad_scale, float* grot,
                              float* gxyz, int B, int N, int n_classes, int width,
                              int64_t src_batch_stride, int64_t src_ray_stride,
                              int64_t tgt_batch_stride, int64_t step_batch_stride,
                              void* stream);

/*
 * ndrr_raymarch_bwd_jac_pose without the workspace argument: it allocates
 * nothing. No counterpart in the reference: it is for callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). The pose partials go to a fixed
 * device array of NDRR_NOALLOC_SCRATCH_FLOATS fp32 that the library holds
 * on every GPU it runs on until the process ends; its address is looked up
 * once per device (hipGetSymbolAddress) on the first call there, with the
 * stream's device current. A call needs B * ceil(N / NDRR_NOALLOC_RAYS_PER_PART)
 * * 12 floats of it and returns NDRR_ESCRATCH, before any launch, when that
 * exceeds the array. Every partial is written before it is read, so nothing
 * is cleared between calls. All calls on one GPU share its array: calls that
 * overlap in time on one GPU (concurrent streams) race; serialise them.
 * Results are bitwise those of ndrr_raymarch_bwd_jac_pose.
 */
#define NDRR_NOALLOC_SCRATCH_FLOATS (1u << 20)
#define NDRR_NOALLOC_RAYS_PER_PART 768
int ndrr_raymarch_bwd_jac_pose_noalloc(const float* go, int64_t go_batch_stride,
                                       const float* jac, int jac_planes,
                                       const double* world_to_voxel, const float* rt_inv,
                                       const float* src, const float* tgt,
                                       const float* step, float* g_rt_inv, float* gsrc,
                                       float* gtgt, float* gstep, int need_gcam, int B,
                                       int N, int n_classes, int width,
                                       int64_t src_batch_stride, int64_t src_ray_stride,
                                       int64_t tgt_batch_stride, int64_t step_batch_stride,
                                       void* stream);

/*
 * Packed copies of a [D][H][W] volume that the march reads with one load per
 * sample (two 16-byte loads for fp32):
 *   NDRR_LAYOUT_QUAD: one record per (x, y0, z0), x in [-1, W], y0 in [-1, H-1],
 *     z0 in [-1, D-1] (x fastest), holding v[z0+a][y0+b][x] for
 *     (a, b) = 00, 01, 10, 11, zero outside the volume; a sample's cell is the
 *     two adjacent records x0 and x0 + 1. 4 copies of each voxel.
 *   NDRR_LAYOUT_STENCIL: one record per cell (x0, y0, z0), x0 in [-1, W-1] etc.,
 *     holding its eight corners in the order (z, y, x) = 000, 001, 010, 011,
 *     100, 101, 110, 111. 8 copies of each voxel.
 *   NDRR_LAYOUT_COEF: a 256-byte header, then one record per cell, on the
 *     stencil's grid, holding the coefficients of the cell's trilinear interpolant
 *       v(f) = c0 + cx fx + cy fy + cxy fx fy + cz fz + cxz fx fz + cyz fy fz + cxyz fx fy fz
 *     in that order (zero-padded corners vzyx: c0 = v000, cx = v001 - v000,
 *     cy = v010 - v000, cz = v100 - v000, cxy = (v011 - v010) - cx,
 *     cxz = (v101 - v100) - cx, cyz = (v110 - v100) - cy,
 *     cxyz = ((v111 - v110) - (v101 - v100)) - ((v011 - v010) - cx), each in fp32
 *     in that order). fp32 storage keeps them as they are (32-byte records).
 *     16-bit storage keeps s times each, rounded once (16-byte records), with s
 *     the largest power of two such that 8 max|v| s <= 65504 (so no
 *     coefficient overflows fp16 and the differences stay clear of its
 *     subnormals): NDRR_BF16 holds s c0 in bf16 and the seven differences in
 *     fp16, NDRR_F16 all eight in fp16. The header holds, as 32-bit words, the
 *     fp32 bits of max|v|, s, 1/s and the record type. The edge differences
 *     behind the gradient are so rounded relative to their own size, not to the
 *     voxel values.
 *   NDRR_LAYOUT_QCOEF (16-bit volumes): a 256-byte header like NDRR_LAYOUT_COEF's
 *     (max|v|, s, 1/s, record type NDRR_F16), then one 8-byte record per quad
 *     record (x, y0, z0) of NDRR_LAYOUT_QUAD's grid holding the bilinear
 *     coefficients of that quad, a = v00, by = v01 - a, bz = v10 - a,
 *     byz = (v11 - v10) - by (vzy zero-padded at x, fp32 in that order), times
 *     s (NDRR_LAYOUT_COEF's rule: the largest power of two with
 *     8 max|v| s <= 65504), rounded once to fp16. A cell is the two adjacent
 *     records x0 and x0 + 1, v = lerp_x(Q(x0), Q(x0 + 1)): the quad layout's
 *     footprint (half the coefficient layout's) with the y and z differences
 *     rounded relative to their own size. ndrr_pack_volume makes it from a
 *     volume of any vol_dtype (e.g. the fp32 values a bf16 volume was rounded
 *     from); the march then takes the volume's 16-bit type as vol_dtype.
 * ndrr_packed_bytes gives the size of `out` (0 for invalid arguments; for
 * NDRR_LAYOUT_LINEAR the volume's own size); ndrr_pack_volume fills it from the
 * linear volume on `stream`, in records of the volume's type.
 * ndrr_pack_coef fills NDRR_LAYOUT_COEF records of record_dtype from a volume
 * of vol_dtype (e.g. bf16 records from the fp32 volume, rounded once).
 * Replaces the per-sample gathers of _fused.py:205-212 (_sample, :11-16) with
 * a one-off re-layout of the subject's volume.
 */
size_t ndrr_packed_bytes(int D, int H, int W, int vol_dtype, int vol_layout);
int ndrr_pack_volume(const void* vol, int vol_dtype, int vol_layout, int D, int H, int W,
                     void* out, void* stream);
int ndrr_pack_coef(const void* vol, int vol_dtype, int record_dtype, int D, int H, int W,
                   void* out, void* stream);

/*
 * NDRR_LAYOUT_LCOEF: coefficient records (16-bit only, NDRR_LAYOUT_COEF's
 * slots and 256-byte header) that also carry the classes of their cell's
 * eight corners, so the march reads no label map: one 16-byte load per sample
 * instead of a record load and a label load. The scale s is the largest power
 * of two with max|coefficient| s <= 2^-7 - 2^-18, so every slot is below 2^-7
 * and the top two exponent bits of each fp16 slot are zero (a bf16 c0, not
 * negative, leaves its sign and top exponent bit zero); those two bits of
 * slot j = x + 2y + 4z hold the class code of the corner at offsets (x, y, z):
 * bits 13-14 of the slot, bits 14-15 for a bf16 c0. The code is lab & 3 (0
 * outside the volume): classes 0..3, with 3 also meaning dropped (-1) when
 * n_classes <= 3. Differences below 2^-14 / s are fp16 subnormals, rounded to
 * 2^-25 / s absolutely. Header words: [0] max|coefficient| (fp32 bits), [1] s,
 * [2] 1/s, [3] record type, [4] NDRR_LCOEF_* bits of what the volume and
 * labels break (0: the records are valid), [5] n_classes.
 * ndrr_pack_coef_labels fills such records of record_dtype (NDRR_BF16 or
 * NDRR_F16) from a volume of vol_dtype (not negative) and its [D][H][W] int8
 * class map (values in [-1, n_classes), no -1 when n_classes == 4), n_classes
 * in [1, 4]. The records are then marched with vol_dtype = record_dtype,
 * vol_layout = NDRR_LAYOUT_LCOEF and the same n_classes (at most 4). The
 * launch does not synchronise: read header word 4 before the first march.
 * Replaces the nearest-label gather of _fused.py:222-228 as well as the
 * trilinear gathers of _fused.py:205-212 with one record per sample.
 */
enum {
  NDRR_LCOEF_NEGATIVE = 1, /* a voxel below zero */
  NDRR_LCOEF_CLASS = 2,    /* a class outside [-1, n_classes), or -1 with 4 classes */
  NDRR_LCOEF_RANGE = 4     /* max|coefficient| not finite or above 2^119 */
};
int ndrr_pack_coef_labels(const void* vol, int vol_dtype, const int8_t* lab, int n_classes,
                          int record_dtype, int D, int H, int W, void* out, void* stream);

/*
 * Labels in the form a layout's march reads. For NDRR_LAYOUT_QUAD: one uint32
 * per quad record (x, y0, z0) holding the int8 classes lab[z0+a][y0+b][x] in
 * byte 2a + b (zero outside the volume), so the nearest-label lookup shares
 * the volume's record index; for NDRR_LAYOUT_STENCIL and NDRR_LAYOUT_COEF the
 * padded map int8 [D + 4][H + 4][W + 4] with lab[z][y][x] at (z + 2, y + 2,
 * x + 2) and 0 elsewhere, which the kernels read at the rounded position
 * clamped to the border, without bounds tests; for NDRR_LAYOUT_LINEAR the
 * [D][H][W] map itself (copied). The packed layouts also need
 * (H + 4)(W + 4) < 2^23. ndrr_packed_label_bytes gives the size of `out`.
 * Replaces the nearest-label gather of _fused.py:222-228 (rounding,
 * zero padding, `.to(tl.int32)`), which the kernels then resolve from the
 * sample's cell fractions.
 */
size_t ndrr_packed_label_bytes(int D, int H, int W, int vol_layout);
int ndrr_pack_labels(const int8_t* lab, int vol_layout, int D, int H, int W, void* out, void* stream);

/*
 * Label map for the kernels: dst[i] = trunc(src[i]) when that lies in
 * [0, n_classes), else -1 (a class the march drops). `src` holds the
 * reference's float label volume.
 * Replaces the float label read of _fused.py:222-228 (`.to(tl.int32)` of
 * the nearest float label) with a one-off int8 conversion.
 */
int ndrr_labels_to_i8(const float* src, int8_t* dst, int64_t numel,
                      int n_classes, void* stream);

/* ---------------------------------------------------------------------------
 * Ray and camera set-up
 * ------------------------------------------------------------------------- */

/*
 * Detector rays of B cameras (fp32): tgt[b][n] = sdd_b * K^-1_b [u + 1/2,
 * v + 1/2, 1] with n = v * width + u; src = [B][1][3] zeros (perspective) or
 * [B][N][3] = tgt - (0, 0, sdd_b) (orthographic); step[b][n] =
 * |tgt - src| / (n_samples - 1). k_inv is [B or 1][3][3], sdd [B or 1];
 * a batch stride of 0 shares one across the poses. Replaces _make_tgt,
 * _make_src and the step of src/nanodrr/drr/renderer.py:82,101-132.
 */
int ndrr_make_rays(const float* k_inv, int64_t k_batch_stride, const float* sdd,
                   int64_t sdd_batch_stride, int B, int height, int width,
                   int orthographic, int n_samples, float* tgt, float* src, float* step,
                   void* stream);

/*
 * M[b] = (world_to_voxel @ rt_inv[b])[:3, :] computed in FP64 and rounded
 * once to fp32 ([B][3][4]); world_to_voxel is [4][4] float64, rt_inv
 * [B][4][4] fp32. ndrr_camera_to_voxel_bwd writes the adjoint
 * g_rt_inv[b] = world_to_voxel[:3, :]^T @ gM[b] ([B][4][4] fp32, FP64 sums).
 * Replaces the composition in src/nanodrr/drr/backends.py:96 and its
 * autograd.
 */
int ndrr_camera_to_voxel(const double* world_to_voxel, const float* rt_inv, float* M,
                         int B, void* stream);
int ndrr_camera_to_voxel_bwd(const double* world_to_voxel, const float* gM,
                             float* g_rt_inv, int B, void* stream);

/* ---------------------------------------------------------------------------
 * Registration support
 * ------------------------------------------------------------------------- */

/*
 * The SE(3) pose step in FP64 (the only FP64 on the GPU). For each of B
 * problems with parameters w = rot[b*param_stride + 0..2] (rotation vector)
 * and v = xyz[b*param_stride + 0..2]:
 *   T = [exp([w]x), exp([w]x) v; 0 0 0 1],
 *   M[b] = (prefix[b] @ T @ suffix[b])[:3, :]   (fp32, [B][3][4]),
 * prefix/suffix are [B][4][4] float64 (for registration: prefix =
 * world_to_voxel @ rt_inv @ pivot, suffix = pivot_inv). ndrr_pose_compose_bwd
 * maps gM [B][3][4] to the parameter gradients grot, gxyz (same strides).
 * Replaces Registration.pose and convert(..., "so3_log") of
 * src/nanodrr/registration/registration.py:64-68 and
 * src/nanodrr/geometry/se3.py:62-69,109-110, with the render's
 * world_to_voxel @ rt_inv (src/nanodrr/drr/backends.py:96) folded in.
 */
int ndrr_pose_compose(const double* prefix, const double* suffix, const float* rot,
                      const float* xyz, int64_t param_stride, float* M, int B,
                      void* stream);
int ndrr_pose_compose_bwd(const double* prefix, const double* suffix, const float* rot,
                          const float* xyz, int64_t param_stride, const float* gM,
                          float* grot, float* gxyz, int B, void* stream);

/*
 * One Adam step in place on n fp32 parameters (param, first and second
 * moments m, v) for their gradient: m = b1 m + (1 - b1) g, v = b2 v +
 * (1 - b2) g^2, param -= lr (m / bc1) / (sqrt(v / bc2) + eps), with the bias
 * corrections bc = 1 - beta^t and the step's learning rate from the caller.
 * The betas are float64 so that 1 - beta is formed before rounding to fp32
 * (as torch.optim.Adam does with its Python floats).
 * The update of the registration loop's per-problem parameters in one
 * launch. No counterpart in the reference: its Registration module
 * (src/nanodrr/registration/registration.py:9-84) leaves the optimisation
 * loop to the caller.
 */
int ndrr_adam_step(float* param, const float* grad, float* m, float* v, int n, float lr,
                   double beta1, double beta2, float eps, float bias_correction1,
                   float bias_correction2, void* stream);

/*
 * Device-side state of the registration loop, so that several iterations can
 * be captured in one HIP graph and replayed (each launch reads the iteration
 * from device memory instead of taking it as an argument). No counterpart in
 * the reference (its Registration module leaves the loop to the caller).
 *   ndrr_lattice_offsets: offsets[b] in [-1/2, 1/2), a counter-based hash of
 *     (seed, *iteration, b): the per-problem sample-lattice jitter.
 *   ndrr_adam_step_scheduled: ndrr_adam_step with the learning rate and bias
 *     corrections read from row *iteration (clamped to [0, rows)) of
 *     schedule[rows][4] = {lr of column 0, lr of column 1, 1 - beta1^t,
 *     1 - beta2^t}; lr_column (0 or 1) picks the parameter group's rate.
 *   ndrr_counter_advance: *counter += 1.
 */
int ndrr_lattice_offsets(const int* iteration, uint32_t seed, float* offsets, int B, void* stream);
int ndrr_adam_step_scheduled(float* param, const float* grad, float* m, float* v, int n,
                             const float* schedule, int rows, int lr_column, const int* iteration,
                             double beta1, double beta2, float eps, void* stream);
int ndrr_counter_advance(int* counter, void* stream);

/*
 * The same loop state in fewer launches (each result bit for bit the separate
 * entries'; no counterpart in the reference, as above):
 *   ndrr_adam_step_scheduled_pair: ndrr_adam_step_scheduled on two parameter
 *     groups of n values each (group k with learning-rate column k) followed
 *     by ndrr_counter_advance on *iteration, in one launch of one workgroup
 *     (every thread reads the counter before a barrier; one thread advances
 *     it after); n <= 2^20.
 *   ndrr_pose_compose_lattice: ndrr_pose_compose and ndrr_lattice_offsets
 *     (offsets[b] for the same B problems) in one launch.
 */
int ndrr_adam_step_scheduled_pair(float* param0, const float* grad0, float* m0, float* v0,
                                  float* param1, const float* grad1, float* m1, float* v1, int n,
                                  const float* schedule, int rows, int* iteration, double beta1,
                                  double beta2, float eps, void* stream);
int ndrr_pose_compose_lattice(const double* prefix, const double* suffix, const float* rot,
                              const float* xyz, int64_t param_stride, float* M, int B,
                              const int* iteration, uint32_t seed, float* offsets, void* stream);

/*
 * Zero-normalised cross-correlation of B image pairs of N pixels (fp32):
 * score[b] = mean(xh * yh), xh = (x - mean x) / sqrt(var x + eps) (population
 * variance), yh likewise; grad_x (optional, [B][N]) = d score[b] / d x.
 * Two launches over (4096-pixel chunk, image): chunk moments into
 * `workspace` (ndrr_zncc_workspace_floats(B, N) floats), then their merge and
 * the gradient. Deterministic (fixed-order reductions). Replaces
 * NormalizedCrossCorrelation2d.forward, src/nanodrr/metrics/ncc.py:16-37.
 */
int64_t ndrr_zncc_workspace_floats(int B, int N);
int ndrr_zncc(const float* x, const float* y, int B, int N, float eps, float* workspace,
              float* score, float* grad_x, void* stream);

/*
 * Local (patch) ZNCC of B image pairs of H x W pixels (fp32): score[b] = the
 * mean over every patch x patch window (stride 1, (H - patch + 1)(W - patch + 1)
 * windows) of the window's ZNCC as ndrr_zncc defines it (population variance
 * + eps per window); grad_x (optional, [B][H][W]) = d score[b] / d x. 1 <=
 * patch <= min(H, W, 32). Two launches: per-window moments and gradient terms
 * into `workspace` (ndrr_patch_zncc_workspace_floats floats, 16-byte
 * aligned), then per-pixel sums over the windows and the score.
 * Deterministic. Replaces NormalizedCrossCorrelation2d with a patch size,
 * src/nanodrr/metrics/ncc.py:16-37,83-88 (unfolded windows standardised one
 * by one), without materialising the windows.
 */
int64_t ndrr_patch_zncc_workspace_floats(int B, int H, int W, int patch);
int ndrr_patch_zncc(const float* x, const float* y, int B, int H, int W, int patch, float eps,
                    float* workspace, float* score, float* grad_x, void* stream);

/* ---------------------------------------------------------------------------
 * Pose-proposal CNN: 3x3 convolutions on the bf16 matrix cores
 * ------------------------------------------------------------------------- */

/*
 * No counterpart in the reference (the learned initialiser of the
 * registration loop). Activations are NHWC bf16 ([N][H][W][C]), weights
 * [Cout][3][3][Cin] bf16, bias fp32, accumulation fp32; padding 1, stride 1
 * or 2 (Ho = (H - 1) / stride + 1). Cout must be a multiple of 32 and Cin 1
 * or a multiple of 32.
 *   ndrr_conv3x3_fwd: y = relu?(conv(x, w) + bias) (bias may be NULL);
 *   ndrr_conv3x3_transpose_weights: wt[3][3][Cin][Cout] from w, for
 *   ndrr_conv3x3_bwd_data: dx = conv^T(dy, w), times [x > 0] when relu_mask
 *     (the ReLU of the layer that produced x); Cin a multiple of 32;
 *   ndrr_conv3x3_bwd_weight: dw[Cout][3][3][Cin] fp32 = sum over pixels of
 *     dy x-patch and, when dbias is not NULL, dbias[Cout] fp32 = sum over
 *     pixels of dy, through a 16-byte aligned workspace of
 *     ndrr_conv3x3_bwd_weight_workspace_floats floats (split partials added
 *     in a fixed order: deterministic); Cout a multiple of 64, or 32 when
 *     Cin == 1. The environment variable NDRR_WGRAD_BLOCKS (64-65536;
 *     default 512, 1024 when Cin == 1) sets the number of workgroups the
 *     pixels are split over.
 */
int ndrr_conv3x3_fwd(const void* x, const void* w, const float* bias, void* y,
                     int N, int H, int W, int Cin, int Cout, int stride, int relu,
                     void* stream);
int ndrr_conv3x3_transpose_weights(const void* w, void* wt, int Cin, int Cout, void* stream);
int ndrr_conv3x3_bwd_data(const void* dy, const void* wt, const void* x, void* dx,
                          int N, int H, int W, int Cin, int Cout, int stride,
                          int relu_mask, void* stream);
size_t ndrr_conv3x3_bwd_weight_workspace_floats(int N, int H, int W, int Cin, int Cout,
                                                int stride);
int ndrr_conv3x3_bwd_weight(const void* dy, const void* x, float* workspace, float* dw,
                            float* dbias, int N, int H, int W, int Cin, int Cout,
                            int stride, void* stream);

/*
 * The pose-proposal CNN's head (no counterpart in the reference): 2 x 2
 * average pooling of the last feature map feat[N][FH][FW][C] (bf16, C a
 * multiple of 64), a fully connected layer of J = 256 units with ReLU and an
 * output layer of O <= 8 units, on the bf16 matrix cores with fp32
 * accumulation. The pooled features are P[N][K], K = (FH/2)(FW/2)C, with
 * k = (py (FW/2) + px) C + c, and their transpose Pt[K][Np], Np = N rounded up
 * to a multiple of 32 (zero columns past N).
 *   ndrr_head_cast_weights: the fp32 first-layer weights w1[J][K] as bf16, in
 *     both orientations (w1_bf16[J][K], w1t_bf16[K][J]).
 *   ndrr_head_pool: P and Pt from feat.
 *   ndrr_head_fwd: H[N][J] = relu(b1 + P w1^T) (fp32) and out[N][O] = b2 + H w2^T
 *     (fp32); workspace of ndrr_head_workspace_floats floats (k-slice
 *     partials, added in a fixed order); K a multiple of 1024.
 *   ndrr_head_bwd: for dout[N][O]: dh[N][J] (fp32) = [H > 0] dout w2, as bf16 in
 *     dH_bf16[N][J] and dHt_bf16[J][Np]; dw1[J][K] = dh^T P, db1, dw2[O][J],
 *     db2 (fp32); dfeat (bf16, feat's shape) = [feat > 0] (dh w1) spread over
 *     the pooling windows / 4 (the ReLU mask of the layer that made feat).
 * Deterministic (no atomics).
 */
size_t ndrr_head_workspace_floats(int N, int FH, int FW, int C, int J);
int ndrr_head_cast_weights(const float* w1, void* w1_bf16, void* w1t_bf16, int J, int K, void* stream);
int ndrr_head_pool(const void* feat, void* P, void* Pt, int N, int FH, int FW, int C, void* stream);
int ndrr_head_fwd(const void* P, const void* w1_bf16, const float* b1, const float* w2, const float* b2,
                  float* workspace, float* H, float* out, int N, int K, int J, int O, void* stream);
int ndrr_head_bwd(const float* dout, const float* H, const float* w2, const void* w1t_bf16, const void* Pt,
                  const void* feat, float* dh, void* dH_bf16, void* dHt_bf16, float* dw1, float* db1,
                  float* dw2, float* db2, void* dfeat, int N, int FH, int FW, int C, int J, int O,
                  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NANODRR_AMD_H */
