// Fused differentiable DRR ray march for AMD Instinct MI355X (gfx950).
//
// One thread marches one ray (or one chunk of its samples). A workgroup covers
// a 32 x 8 detector tile as four wave64s of 8 x 8 rays per chunk group, so the
// rays sharing a CU's L1 are spatially coherent and their gathers land in few
// cache lines; workgroups are remapped so each XCD (own L2) runs neighbouring
// tiles. Per ray the kernel clips the sample range to the samples whose 2x2x2
// footprint can touch the volume (skipped samples contribute exactly zero, so
// the clip changes no bit of the result), then marches: per sample the
// trilinear stencil (eight zero-padded corners, from the linear volume or one
// packed record) and a nearest-label route, with several samples' loads in
// flight per lane.
//
// The arithmetic of the march and of its adjoint follows nanodrr's fused
// kernel (src/nanodrr/drr/_fused.py:133-300); the launch layout, clip,
// reductions and compensated accumulation are this library's own.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <mutex>

#include "nanodrr_amd.h"
#include "se3_pose.h"

namespace {

constexpr int kWave = 64;        // rays per wave: an 8 x 8 detector patch (or 64 consecutive rays)
// Tuning knobs (compile-time; the defaults are the measured best on MI355X).
#ifndef NDRR_TILE_WAVES_X
#define NDRR_TILE_WAVES_X 4
#endif
#ifndef NDRR_TILE_WAVES_Y
#define NDRR_TILE_WAVES_Y 1
#endif
// Chunks: wave groups per workgroup marching the same rays; group c takes the
// c-th equal part of every ray's clipped sample range. Per kernel and layout.
#ifndef NDRR_FWD_CHUNKS_LINEAR
#define NDRR_FWD_CHUNKS_LINEAR 1
#endif
#ifndef NDRR_FWD_CHUNKS_STENCIL
#define NDRR_FWD_CHUNKS_STENCIL 1
#endif
#ifndef NDRR_FWD_CHUNKS_QUAD
#define NDRR_FWD_CHUNKS_QUAD 2
#endif
#ifndef NDRR_FWD_CHUNKS_COEF
#define NDRR_FWD_CHUNKS_COEF 2
#endif
// Single-channel renders (no label map) carry 7 sums per ray instead of up to
// 28 and hold fewer registers, so more chunks fit; quad layout only.
#ifndef NDRR_FWD_CHUNKS_QUAD_1CLASS
#define NDRR_FWD_CHUNKS_QUAD_1CLASS 3
#endif
#ifndef NDRR_FWD_CHUNKS_COEF_1CLASS
#define NDRR_FWD_CHUNKS_COEF_1CLASS 3
#endif
// Pose-gradient partials of the backward kernel: 1 = one per tile (the waves
// of a workgroup added through LDS after a barrier), 0 = one per wave.
#ifndef NDRR_BWD_TILE_PARTIALS
#define NDRR_BWD_TILE_PARTIALS 1
#endif
#ifndef NDRR_BWD_CHUNKS_LINEAR
#define NDRR_BWD_CHUNKS_LINEAR 1
#endif
#ifndef NDRR_BWD_CHUNKS_STENCIL
#define NDRR_BWD_CHUNKS_STENCIL 1
#endif
#ifndef NDRR_BWD_CHUNKS_QUAD
#define NDRR_BWD_CHUNKS_QUAD 1
#endif
#ifndef NDRR_BWD_CHUNKS_COEF
#define NDRR_BWD_CHUNKS_COEF 1
#endif
#ifndef NDRR_IN_FLIGHT
#define NDRR_IN_FLIGHT 4
#endif
// Labelled coefficient records: no label loads, so more record loads fit in flight.
#ifndef NDRR_IN_FLIGHT_LCOEF
#define NDRR_IN_FLIGHT_LCOEF 6
#endif
#ifndef NDRR_IN_FLIGHT_QCOEF
#define NDRR_IN_FLIGHT_QCOEF 4
#endif
#ifndef NDRR_WAVES_PER_EU
#define NDRR_WAVES_PER_EU 1
#endif
#ifndef NDRR_XCD_REMAP
#define NDRR_XCD_REMAP 1
#endif
constexpr int kTwx = NDRR_TILE_WAVES_X, kTwy = NDRR_TILE_WAVES_Y;  // 8 x 8 wave patches per tile
constexpr int kTileW = 8 * kTwx, kTileH = 8 * kTwy;                // tile size in detector pixels
constexpr int kTileRays = kWave * kTwx * kTwy;
// The two coefficient layouts (plain and labelled records) march alike.
constexpr bool is_coef(int layout) { return layout == NDRR_LAYOUT_COEF || layout == NDRR_LAYOUT_LCOEF; }
// The quad grid's layouts (corner quads and their coefficient form) share records, labels and chunking.
constexpr bool is_quad(int layout) { return layout == NDRR_LAYOUT_QUAD || layout == NDRR_LAYOUT_QCOEF; }
// Layouts whose records follow a 256-byte header and hold the coefficients times a scale s.
constexpr bool has_header(int layout) { return is_coef(layout) || layout == NDRR_LAYOUT_QCOEF; }
constexpr int fwd_chunks(int layout) {
  return layout == NDRR_LAYOUT_STENCIL ? NDRR_FWD_CHUNKS_STENCIL
         : is_quad(layout)             ? NDRR_FWD_CHUNKS_QUAD
         : is_coef(layout)             ? NDRR_FWD_CHUNKS_COEF
                                       : NDRR_FWD_CHUNKS_LINEAR;
}
constexpr int bwd_chunks(int layout) {
  return layout == NDRR_LAYOUT_STENCIL ? NDRR_BWD_CHUNKS_STENCIL
         : is_quad(layout)             ? NDRR_BWD_CHUNKS_QUAD
         : is_coef(layout)             ? NDRR_BWD_CHUNKS_COEF
                                       : NDRR_BWD_CHUNKS_LINEAR;
}
constexpr int cmax(int a, int b) { return a > b ? a : b; }
// Forward chunks for CV class slots: up to four slots only, so the LDS that
// carries the chunks' partial sums (and the single-pass adjoint's, 6 per
// slot) stays small.
constexpr int fwd_chunks(int layout, int cv) {
  return cv == 1 && is_quad(layout)             ? NDRR_FWD_CHUNKS_QUAD_1CLASS
         : cv == 1 && is_coef(layout)            ? NDRR_FWD_CHUNKS_COEF_1CLASS
                                                 : (cv <= 4 ? fwd_chunks(layout) : 1);
}
constexpr int kMaxFwdChunks = cmax(cmax(cmax(fwd_chunks(NDRR_LAYOUT_LINEAR), fwd_chunks(NDRR_LAYOUT_STENCIL)),
                                        cmax(fwd_chunks(NDRR_LAYOUT_QUAD), fwd_chunks(NDRR_LAYOUT_COEF))),
                                   cmax(NDRR_FWD_CHUNKS_QUAD_1CLASS, NDRR_FWD_CHUNKS_COEF_1CLASS));
constexpr int kMaxBwdChunks = cmax(cmax(bwd_chunks(NDRR_LAYOUT_LINEAR), bwd_chunks(NDRR_LAYOUT_STENCIL)),
                                   cmax(bwd_chunks(NDRR_LAYOUT_QUAD), bwd_chunks(NDRR_LAYOUT_COEF)));
// Pose-gradient partials per tile of the backward kernel (one, or one per
// wave of the layout with the most chunks): the workspace is sized for it.
constexpr int bwd_parts_per_tile(int layout) {
  return NDRR_BWD_TILE_PARTIALS ? 1 : (kTileRays / kWave) * bwd_chunks(layout);
}
constexpr int kBwdPartsMax = NDRR_BWD_TILE_PARTIALS ? 1 : (kTileRays / kWave) * kMaxBwdChunks;
static_assert(kTileRays * cmax(kMaxFwdChunks, kMaxBwdChunks) <= 1024, "workgroups are limited to 1024 threads");
constexpr int kThreads = 256;    // threads per block of the element-wise helper kernels
constexpr float kMargin = 0.5f;  // voxels of slack around the footprint box
constexpr int kKahanMinSamples = 512;

// ---------------------------------------------------------------------------
// Volume element loads (fp32 math whatever the storage type)
// ---------------------------------------------------------------------------
template <int DT> struct Vox;
template <> struct Vox<NDRR_F32> {
  using T = float;
  __device__ static inline float ld(const T* p, int i) { return p[i]; }
};
template <> struct Vox<NDRR_BF16> {
  using T = uint16_t;
  __device__ static inline float ld(const T* p, int i) {
    return __uint_as_float(static_cast<uint32_t>(p[i]) << 16);
  }
};
template <> struct Vox<NDRR_F16> {
  using T = _Float16;
  __device__ static inline float ld(const T* p, int i) { return static_cast<float>(p[i]); }
};

struct Params {
  const void* vol;
  const int8_t* lab;
  const float* M;
  const double* w2v;  // when set, M is composed per workgroup: (w2v @ rt[b])[:3] in FP64
  const float* rt;    // [B][4][4] camera -> world poses (rt_inv)
  const float* src;
  const float* tgt;
  const float* step;
  const float* lattice;  // [B] per-pose sample-lattice offset in steps: sample s sits at t = (s + offset) / (S - 1)
  const float* go;
  float* out;
  float* gpart;  // [B][nblocks][parts per tile][12] pose-gradient partials
  float* jac;    // [B][jac_q][C][N] per-ray sums of the single-pass adjoint
  int jac_q;     // planes of jac: 6 (G, T) or 7 (G, T, V)
  float* gsrc;
  float* gtgt;
  float* gstep;
  float* gvol;
  int gvol_merge;  // merge lanes' volume-gradient atomics per cell (cell keys fit 32 bits)
  float* gvol_bricks;  // 4 x 4 x 4-voxel bricks the merged scatter adds into instead of gvol (or null)
  int nbx, nby;        // bricks along x and y
  int64_t src_bs, src_rs, tgt_bs, step_bs;
  int64_t go_bs;  // batch stride of go in the single-pass backward (0: shared by every pose)
  uint32_t vol_bytes, lab_bytes;  // sizes of the packed volume / label buffers
  uint32_t vol_skip;              // bytes before the first record (the coefficient layout's header)
  int B, N, D, H, W, S, C;
  int IW, IH, tiles_w, nblocks;
  float inv_s;   // 1 / (S - 1)
  float s_m1;    // S - 1
  float inv_lo;  // 1 / (S - 1) - inv_s, so inv_s + inv_lo carries the reciprocal to ~2^-48
};

// Ray of thread `tr` (0 .. kTileRays-1) in tile `blk` and whether it is a real
// detector pixel. Each wave holds an 8 x 8 patch of the tile; with chunks > 1
// every chunk group holds the same rays (one chunk of samples each).
__device__ inline int ray_index(const Params& p, int blk, int tr, bool& valid) {
  if (p.IW > 0) {
    const int ty = blk / p.tiles_w;
    const int tx = blk - ty * p.tiles_w;
    const int wave = tr >> 6, lane = tr & 63;
    const int px = tx * kTileW + (wave % kTwx) * 8 + (lane & 7);
    const int py = ty * kTileH + (wave / kTwx) * 8 + (lane >> 3);
    valid = (px < p.IW) && (py < p.IH);
    return py * p.IW + px;
  }
  const int n = blk * kTileRays + tr;
  valid = n < p.N;
  return n;
}

// The camera -> voxel-index affine of pose b: p.M's row, or, when p.w2v is
// set, (world_to_voxel @ rt_inv[b])[:3] composed in FP64 and rounded once to
// fp32 (the sums of k_camera_to_voxel, in the same order) by threads 0..11
// into sM. Every thread of the workgroup must call it (it holds a barrier).
__device__ inline const float* pose_M(const Params& p, int b, float* sM) {
  if (p.w2v == nullptr) return p.M + 12 * b;
  if (threadIdx.x < 12) {
    const int i = threadIdx.x / 4, j = threadIdx.x % 4;
    const float* R = p.rt + 16 * static_cast<int64_t>(b);
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) acc = fma(p.w2v[4 * i + k], static_cast<double>(R[4 * k + j]), acc);
    sM[threadIdx.x] = static_cast<float>(acc);
  }
  __syncthreads();
  return sM;
}

// (pose, tile) of this workgroup. Workgroups are dispatched round-robin over
// the 8 XCDs (each with its own L2); with the remap, XCD k runs a contiguous
// range of (pose, tile) pairs, so the tiles sharing an L2 are neighbours.
constexpr int kXcds = 8;
__device__ inline void block_coords(int& b, int& tile) {
  if constexpr (NDRR_XCD_REMAP) {
    const int ntiles = gridDim.x;
    const int total = ntiles * gridDim.y;
    const int id = blockIdx.y * ntiles + blockIdx.x;
    const int per = total / kXcds, rem = total % kXcds;
    const int xcd = id % kXcds, slot = id / kXcds;
    // XCD k owns per (+1 for k < rem) consecutive ids starting at k * per + min(k, rem)
    const int lin = xcd * per + min(xcd, rem) + slot;
    b = lin / ntiles;
    tile = lin - b * ntiles;
  } else {
    b = blockIdx.y;
    tile = blockIdx.x;
  }
}

// ---------------------------------------------------------------------------
// Double-float (hi + lo) arithmetic for the per-ray set-up. Sample positions
// are p(s) = A + k * delta with the anchor A = p(s_lo) and the per-sample
// step delta = M (tgt - src) / (S - 1) both carried to ~2^-44 relative, so a
// sample lands within ~1 ulp of its exact position. Plain fp32 (p0 + t d)
// errs by ~1e-4 voxel at the ~1000-voxel ray lengths of a cone beam, enough
// to move samples across cell faces, where the trilinear derivative jumps.
// ---------------------------------------------------------------------------
struct DF {
  float hi, lo;
};

__device__ inline DF fast_two_sum(float a, float b) {  // |a| >= |b|
#pragma clang fp contract(off)
  const float s = a + b;
  return {s, b - (s - a)};
}

__device__ inline DF two_sum(float a, float b) {
#pragma clang fp contract(off)
  const float s = a + b;
  const float bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}

__device__ inline DF two_prod(float a, float b) {
#pragma clang fp contract(off)
  const float p = a * b;
  return {p, fmaf(a, b, -p)};
}

__device__ inline DF df_add(DF a, DF b) {
#pragma clang fp contract(off)
  const DF s = two_sum(a.hi, b.hi);
  return fast_two_sum(s.hi, s.lo + (a.lo + b.lo));
}

__device__ inline DF df_mul(DF a, float b) {
  DF p = two_prod(a.hi, b);
  return fast_two_sum(p.hi, fmaf(a.lo, b, p.lo));
}

struct Ray {
  float sx, sy, sz;     // source, camera space
  float ex, ey, ez;     // tgt - src, camera space
  DF ax, ay, az;        // voxel-space position of the anchor: the ray's first clipped sample
  DF dx, dy, dz;        // voxel-space step per sample
  float step;
  float t_off;          // the lattice offset in ray-parameter units: t(s) = s / (S - 1) + t_off
  int s_lo, n_s;        // this chunk's samples [s_lo, s_lo + n_s) can touch the volume
  int k_base;           // s_lo minus the anchor's sample index
};

// Ray-parameter interval where lo <= p0 + t d <= hi (per axis). An approximate
// reciprocal is ample: the slab is widened by kMargin voxels.
__device__ inline void axis_range(float p0, float d, float lo, float hi, float& t0, float& t1) {
  const float inv = __builtin_amdgcn_rcpf(fabsf(d) < 1e-12f ? 1e-12f : d);
  const float a = (lo - p0) * inv;
  const float b = (hi - p0) * inv;
  t0 = fminf(a, b);
  t1 = fmaxf(a, b);
}

__device__ inline DF df_mul_df(DF a, DF b) {
  DF p = two_prod(a.hi, b.hi);
  return fast_two_sum(p.hi, fmaf(a.hi, b.lo, fmaf(a.lo, b.hi, p.lo)));
}

// Per-ray set-up for the samples of chunk `chunk` of `nch` equal parts of
// the clipped sample range: the chunk's first sample and count. Positions are
// taken from one anchor per ray (its first clipped sample) in every chunk, so
// a sample lands at the same bits whatever the chunking.
__device__ inline Ray setup_ray(const Params& p, const float* __restrict__ Mb, int b, int n, bool valid, int chunk,
                                int nch) {
  Ray r;
  const int nn = valid ? n : 0;
  const float* sp = p.src + b * p.src_bs + nn * p.src_rs;
  const float* tp = p.tgt + b * p.tgt_bs + 3 * static_cast<int64_t>(nn);
  r.sx = sp[0]; r.sy = sp[1]; r.sz = sp[2];
  r.step = p.step[b * p.step_bs + nn];
  const float off = p.lattice != nullptr ? p.lattice[b] : 0.0f;
  DF p0[3], d[3];
  if (r.sx == 0.0f && r.sy == 0.0f && r.sz == 0.0f) {
    // A source at the camera origin (every perspective render): p0 is M's
    // translation column and tgt - src is tgt, both exact.
    r.ex = tp[0]; r.ey = tp[1]; r.ez = tp[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float m0 = Mb[4 * i], m1 = Mb[4 * i + 1], m2 = Mb[4 * i + 2];
      p0[i] = DF{Mb[4 * i + 3], 0.0f};
      d[i] = df_add(df_add(two_prod(m0, r.ex), two_prod(m1, r.ey)), two_prod(m2, r.ez));
    }
  } else {
    const DF ex = two_sum(tp[0], -r.sx), ey = two_sum(tp[1], -r.sy), ez = two_sum(tp[2], -r.sz);
    r.ex = ex.hi; r.ey = ey.hi; r.ez = ez.hi;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float m0 = Mb[4 * i], m1 = Mb[4 * i + 1], m2 = Mb[4 * i + 2], m3 = Mb[4 * i + 3];
      p0[i] = df_add(df_add(df_add(two_prod(m0, r.sx), two_prod(m1, r.sy)), two_prod(m2, r.sz)), DF{m3, 0.0f});
      d[i] = df_add(df_add(df_mul(ex, m0), df_mul(ey, m1)), df_mul(ez, m2));
    }
  }

  // Clip to the samples whose footprint can reach the volume (fp32 is ample here).
  float ax0, ax1, ay0, ay1, az0, az1;
  axis_range(p0[0].hi, d[0].hi, -1.0f - kMargin, static_cast<float>(p.W) + kMargin, ax0, ax1);
  axis_range(p0[1].hi, d[1].hi, -1.0f - kMargin, static_cast<float>(p.H) + kMargin, ay0, ay1);
  axis_range(p0[2].hi, d[2].hi, -1.0f - kMargin, static_cast<float>(p.D) + kMargin, az0, az1);
  const float t_lo = fmaxf(fmaxf(ax0, ay0), az0);
  const float t_hi = fminf(fminf(ax1, ay1), az1);
  int s_lo = 0, n_all = 0;
  if (valid && t_lo <= t_hi) {
    // The slab is widened by kMargin voxels, far more than the rounding of
    // t_lo, t_hi: every sample whose footprint reaches the volume is kept; the
    // few extra ones add exact zeros.
    const float f_lo = fmaxf(floorf(t_lo * p.s_m1 - off), 0.0f);
    const float f_hi = fminf(ceilf(t_hi * p.s_m1 - off) + 1.0f, static_cast<float>(p.S));
    if (f_lo < f_hi) {
      s_lo = static_cast<int>(f_lo);
      n_all = static_cast<int>(f_hi) - s_lo;
    }
  }
  const int per = (n_all + nch - 1) / nch;
  r.k_base = chunk * per;
  r.s_lo = s_lo + r.k_base;
  r.n_s = max(0, min(per, n_all - r.k_base));
  const float s0 = static_cast<float>(s_lo);
  const DF inv{p.inv_s, p.inv_lo};
  DF* const an[3] = {&r.ax, &r.ay, &r.az};
  DF* const st[3] = {&r.dx, &r.dy, &r.dz};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const DF delta = df_mul_df(d[i], inv);
    *st[i] = delta;
    *an[i] = df_add(p0[i], df_mul(delta, s0));
    if (p.lattice != nullptr) *an[i] = df_add(*an[i], df_mul(delta, off));
  }
  r.t_off = off * p.inv_s;
  return r;
}

// Voxel-space position of sample s_lo + k.
__device__ inline float sample_coord(const DF& a, const DF& d, float k) {
  return fmaf(k, d.hi, a.hi) + fmaf(k, d.lo, a.lo);
}

// Cell of a sample: integer corner (ix, iy, iz) and fractions.
struct Cell {
  int ix, iy, iz;
  float fx, fy, fz;
};

// floor(x) as an int in one instruction (the compiler emits a floor and a conversion).
__device__ inline int floor_to_int(float x) {
  int i;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(i) : "v"(x));
  return i;
}

// The fraction x - floor(x) comes from v_fract_f32, which keeps it below 1 where the
// subtraction would round up to 1 (x a tiny negative number); floor_to_int saturates
// beyond the int range like the conversion of floorf.
__device__ inline Cell cell_of(float x, float y, float z) {
  Cell c;
  c.fx = __builtin_amdgcn_fractf(x); c.fy = __builtin_amdgcn_fractf(y); c.fz = __builtin_amdgcn_fractf(z);
  c.ix = floor_to_int(x); c.iy = floor_to_int(y); c.iz = floor_to_int(z);
  return c;
}

// Corner values vABC (z offset A, y offset B, x offset C), zero outside the volume.
struct Corners {
  float v000, v001, v010, v011, v100, v101, v110, v111;
};

// What the fetches read. The linear layout reads through plain pointers; the
// packed layouts read through buffer resources with 32-bit byte offsets, and
// a cell outside the packed grid gets an offset past the end of the buffer,
// for which the hardware returns zeros (no clamping, no 64-bit address math).
constexpr uint32_t kBufferFlags = 0x00020000u;  // raw buffer, 32-bit data (gfx9 family)
constexpr uint32_t kRintBias = 0x4B400000u;     // the bits of 1.5 * 2^23 (rint_bits)

struct Src {
  const void* vol;
  const int8_t* lab;
  __amdgpu_buffer_rsrc_t vrs, lrs;
  int D, H, W;
  uint32_t lab_base;  // padded class map: the border offset less the 2^22 terms of padded_label
};

__device__ inline Src make_src(const Params& p) {
  Src s;
  s.vol = p.vol;
  s.lab = p.lab;
  s.vrs = __builtin_amdgcn_make_buffer_rsrc(static_cast<char*>(const_cast<void*>(p.vol)) + p.vol_skip, 0,
                                            static_cast<int>(p.vol_bytes - p.vol_skip),
                                            static_cast<int>(kBufferFlags));
  s.lrs = __builtin_amdgcn_make_buffer_rsrc(const_cast<int8_t*>(p.lab), 0, static_cast<int>(p.lab_bytes),
                                            static_cast<int>(kBufferFlags));
  s.D = p.D; s.H = p.H; s.W = p.W;
  const uint32_t w4 = static_cast<uint32_t>(p.W + 4), hw4 = static_cast<uint32_t>(p.H + 4) * w4;
  s.lab_base = (2u * hw4 + 2u * w4 + 2u) - (1u << 22) * (hw4 + w4) - kRintBias;
  return s;
}

// Raw data of one sample's cell as loaded; decoded into Corners at use, so
// that several samples' loads can be in flight with few registers.
template <int DT, int L> struct Fetch;

// Linear layout: eight masked element loads (clamped addresses stay in bounds).
template <int DT> struct Fetch<DT, NDRR_LAYOUT_LINEAR> {
  static constexpr int kInFlight = 2;
  float v[8];
  __device__ inline void load(const Src& s, const Cell& c, bool live) {
    const int D = s.D, H = s.H, W = s.W;
    const auto* __restrict__ vol = static_cast<const typename Vox<DT>::T*>(s.vol);
    const bool bx0 = static_cast<unsigned>(c.ix) < static_cast<unsigned>(W);
    const bool bx1 = static_cast<unsigned>(c.ix + 1) < static_cast<unsigned>(W);
    const bool by0 = static_cast<unsigned>(c.iy) < static_cast<unsigned>(H);
    const bool by1 = static_cast<unsigned>(c.iy + 1) < static_cast<unsigned>(H);
    // a sample that is not live reads zeros, like a corner outside the volume
    const bool bz0 = live & (static_cast<unsigned>(c.iz) < static_cast<unsigned>(D));
    const bool bz1 = live & (static_cast<unsigned>(c.iz + 1) < static_cast<unsigned>(D));
    const int cx0 = min(max(c.ix, 0), W - 1), cx1 = min(max(c.ix + 1, 0), W - 1);
    const int cy0 = min(max(c.iy, 0), H - 1), cy1 = min(max(c.iy + 1, 0), H - 1);
    const int cz0 = min(max(c.iz, 0), D - 1), cz1 = min(max(c.iz + 1, 0), D - 1);
    const int r00 = (cz0 * H + cy0) * W, r01 = (cz0 * H + cy1) * W;
    const int r10 = (cz1 * H + cy0) * W, r11 = (cz1 * H + cy1) * W;
    const float a000 = Vox<DT>::ld(vol, r00 + cx0), a001 = Vox<DT>::ld(vol, r00 + cx1);
    const float a010 = Vox<DT>::ld(vol, r01 + cx0), a011 = Vox<DT>::ld(vol, r01 + cx1);
    const float a100 = Vox<DT>::ld(vol, r10 + cx0), a101 = Vox<DT>::ld(vol, r10 + cx1);
    const float a110 = Vox<DT>::ld(vol, r11 + cx0), a111 = Vox<DT>::ld(vol, r11 + cx1);
    v[0] = (bz0 && by0 && bx0) ? a000 : 0.0f;
    v[1] = (bz0 && by0 && bx1) ? a001 : 0.0f;
    v[2] = (bz0 && by1 && bx0) ? a010 : 0.0f;
    v[3] = (bz0 && by1 && bx1) ? a011 : 0.0f;
    v[4] = (bz1 && by0 && bx0) ? a100 : 0.0f;
    v[5] = (bz1 && by0 && bx1) ? a101 : 0.0f;
    v[6] = (bz1 && by1 && bx0) ? a110 : 0.0f;
    v[7] = (bz1 && by1 && bx1) ? a111 : 0.0f;
  }
  __device__ inline Corners corners() const { return {v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]}; }
};

// Stencil-record layout: one record per trilinear cell (x0, y0, z0) with
// x0 in [-1, W-1], y0 in [-1, H-1], z0 in [-1, D-1], holding the cell's eight
// corner values (zero outside the volume) in the order v000 v001 v010 v011
// v100 v101 v110 v111 (x fastest); records are x-fastest too. A sample then
// costs one 16-byte load (bf16/fp16) or two (fp32) instead of eight gathers.
template <int DT> struct Rec { static constexpr int kBytes = 16; };
template <> struct Rec<NDRR_F32> { static constexpr int kBytes = 32; };

__device__ inline float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ inline float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
__device__ inline float f16_lo(uint32_t w) {
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w & 0xffffu)));
}
__device__ inline float f16_hi(uint32_t w) {
  return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(w >> 16)));
}

// Whether cell (ix, iy, iz) exists in a packed grid (cells -1 .. dim-1);
// every other cell has all eight corners outside the volume.
// Bitwise rather than short-circuit conjunction, which the compiler turns
// into a divergent branch.
__device__ inline bool cell_exists(const Src& s, const Cell& c) {
  return (static_cast<unsigned>(c.ix + 1) <= static_cast<unsigned>(s.W)) &
         (static_cast<unsigned>(c.iy + 1) <= static_cast<unsigned>(s.H)) &
         (static_cast<unsigned>(c.iz + 1) <= static_cast<unsigned>(s.D));
}

// a[23:0] * b[23:0] + c, unsigned and signed (the operands' low 24 bits, sign-extended for
// the signed form): one full-rate multiply-add each. Written as instructions because the
// compiler turns the 24-bit products of these index expressions into quarter-rate 32-bit
// multiplies or 64-bit multiply-adds once it has distributed the +1 offsets over them.
__device__ inline uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t d;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
  return d;
}
__device__ inline int mad_i24(int a, int b, int c) {
  int d;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
  return d;
}

// Index of a cell's record in a grid of (D + 1) x (H + 1) x row records.
// check_common keeps (D + 1)(H + 1) and the row below 2^24 for the packed
// layouts, so the operands of an existing cell fit; a cell that does not exist
// gets an offset past the buffer whatever this returns.
__device__ inline uint32_t record_of(const Src& s, const Cell& c, int row) {
  return mad_u24(mad_u24(static_cast<uint32_t>(c.iz + 1), static_cast<uint32_t>(s.H + 1),
                         static_cast<uint32_t>(c.iy + 1)),
                 static_cast<uint32_t>(row), static_cast<uint32_t>(c.ix + 1));
}

// Byte offset of a cell's record, or one past 2 GiB (so past the end of the
// buffer, which reads zeros) when the cell does not exist or the sample is not
// live (past the ray's last sample): its record then adds nothing without a
// select per value. Bit arithmetic rather than a select, which the compiler
// would otherwise turn into a divergent branch.
__device__ inline uint32_t record_offset(const Src& s, const Cell& c, int row, uint32_t bytes, bool live) {
  const uint32_t gone = static_cast<uint32_t>(!(cell_exists(s, c) & live)) << 31;
  return (record_of(s, c, row) * bytes) | gone;
}

__device__ inline uint4 load16(__amdgpu_buffer_rsrc_t rs, uint32_t off) {
  return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
}

template <int DT> struct Fetch<DT, NDRR_LAYOUT_STENCIL> {  // bf16 / fp16: 16-byte records
  static constexpr int kInFlight = NDRR_IN_FLIGHT;
  uint4 q;
  __device__ inline void load(const Src& s, const Cell& c, bool live) {
    const uint32_t off = record_offset(s, c, s.W + 1, 16u, live);
    q = load16(s.vrs, off);
  }
  __device__ inline Corners corners() const {
    if constexpr (DT == NDRR_BF16)
      return {bf16_lo(q.x), bf16_hi(q.x), bf16_lo(q.y), bf16_hi(q.y),
              bf16_lo(q.z), bf16_hi(q.z), bf16_lo(q.w), bf16_hi(q.w)};
    else
      return {f16_lo(q.x), f16_hi(q.x), f16_lo(q.y), f16_hi(q.y),
              f16_lo(q.z), f16_hi(q.z), f16_lo(q.w), f16_hi(q.w)};
  }
};

template <> struct Fetch<NDRR_F32, NDRR_LAYOUT_STENCIL> {  // fp32: 32-byte records
  static constexpr int kInFlight = 2;
  uint4 a, b;
  __device__ inline void load(const Src& s, const Cell& c, bool live) {
    const uint32_t off = record_offset(s, c, s.W + 1, 32u, live);
    a = load16(s.vrs, off);
    b = load16(s.vrs, off + 16u);
  }
  __device__ inline Corners corners() const {
    return {__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), __uint_as_float(a.w),
            __uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w)};
  }
};

// Quad layout: one record per (x, y0, z0), x in [-1, W], y0 in [-1, H-1],
// z0 in [-1, D-1], holding the (y, z) quad v[z0+a][y0+b][x] (zero outside the
// volume) in the order (a, b) = 00, 01, 10, 11; records x-fastest. The cell
// (x0, y0, z0) is records x0 and x0 + 1, adjacent in memory: one 16-byte load
// (bf16/fp16) or two (fp32) per sample, with each voxel stored 4 times
// instead of the stencil's 8, so a 256^3 bf16 volume (136 MB) stays resident
// in the 256 MB Infinity Cache and adjacent rays share records along x.
template <int DT> struct Fetch<DT, NDRR_LAYOUT_QUAD> {  // bf16 / fp16: two 8-byte records
  static constexpr int kInFlight = NDRR_IN_FLIGHT;
  uint4 q;  // x: (z0,y0) (z0,y0+1) at x0; y: (z0+1,y0) (z0+1,y0+1) at x0; z, w: the same at x0 + 1
  __device__ inline void load(const Src& s, const Cell& c, bool live) {
    const uint32_t off = record_offset(s, c, s.W + 2, 8u, live);
    q = load16(s.vrs, off);
  }
  __device__ inline Corners corners() const {
    if constexpr (DT == NDRR_BF16)
      return {bf16_lo(q.x), bf16_lo(q.z), bf16_hi(q.x), bf16_hi(q.z),
              bf16_lo(q.y), bf16_lo(q.w), bf16_hi(q.y), bf16_hi(q.w)};
    else
      return {f16_lo(q.x), f16_lo(q.z), f16_hi(q.x), f16_hi(q.z),
              f16_lo(q.y), f16_lo(q.w), f16_hi(q.y), f16_hi(q.w)};
  }
};

template <> struct Fetch<NDRR_F32, NDRR_LAYOUT_QUAD> {  // fp32: two 16-byte records
  static constexpr int kInFlight = 2;
  uint4 a, b;  // quads at x0 and x0 + 1: (z0,y0) (z0,y0+1) (z0+1,y0) (z0+1,y0+1)
  __device__ inline void load(const Src& s, const Cell& c, bool live) {
    const uint32_t off = record_offset(s, c, s.W + 2, 16u, live);
    a = load16(s.vrs, off);
    b = load16(s.vrs, off + 16u);
  }
  __device__ inline Corners corners() const {
    return {__uint_as_float(a.x), __uint_as_float(b.x), __uint_as_float(a.y), __uint_as_float(b.y),
            __uint_as_float(a.z), __uint_as_float(b.z), __uint_as_float(a.w), __uint_as_float(b.w)};
  }
};

// Quad coefficient records (bf16 / fp16 volumes): on the quad grid, record
// (x, y0, z0) holds the bilinear coefficients of the quad at x,
//   Q(fy, fz) = a + by fy + bz fz + byz fy fz
// (a = v(z0, y0), by = v(z0, y0+1) - a, bz = v(z0+1, y0) - a,
// byz = (v(z0+1, y0+1) - v(z0+1, y0)) - by), each times a power of two s in
// fp16 (8 bytes, after a kCoefHeader-byte header). A cell is the two adjacent
// records x0 and x0 + 1 (one 16-byte load, as for quad): v = lerp_x(Q0, Q1).
// The fp16 slots feed mixed-precision multiply-adds directly, so value and
// gradient cost 14 operations with no unpacking, and the y and z differences
// are rounded relative to their own size; the records keep the quad grid's
// half-size footprint (the coefficient layout's is twice it).
template <int DT> struct Fetch<DT, NDRR_LAYOUT_QCOEF> {
  static_assert(DT != NDRR_F32, "quad coefficient records serve 16-bit volumes");
  static constexpr int kInFlight = NDRR_IN_FLIGHT_QCOEF;
  uint4 q;  // x, y: a, by, bz, byz at x0; z, w: the same at x0 + 1
  __device__ inline void load(const Src& s, const Cell& c, bool live) {
    q = load16(s.vrs, record_offset(s, c, s.W + 2, 8u, live));
  }
};

struct QuadCoef {
  float a, by, bz, byz;
};
__device__ inline QuadCoef quad_coef(uint32_t lo, uint32_t hi) { return {f16_lo(lo), f16_hi(lo), f16_lo(hi), f16_hi(hi)}; }

// Q and its fy, fz derivatives: 4 multiply-adds.
__device__ inline float quad_coef_value(const QuadCoef& k, const Cell& c, float& dy, float& dz) {
  dz = fmaf(c.fy, k.byz, k.bz);
  dy = fmaf(c.fz, k.byz, k.by);
  return fmaf(c.fz, dz, fmaf(c.fy, k.by, k.a));
}

template <int DT>
__device__ inline float qcoef_value_grad(const Fetch<DT, NDRR_LAYOUT_QCOEF>& f, const Cell& c, float& dvx, float& dvy,
                                         float& dvz) {
  float dy0, dz0, dy1, dz1;
  const float q0 = quad_coef_value(quad_coef(f.q.x, f.q.y), c, dy0, dz0);
  const float q1 = quad_coef_value(quad_coef(f.q.z, f.q.w), c, dy1, dz1);
  dvx = q1 - q0;
  dvy = fmaf(c.fx, dy1 - dy0, dy0);
  dvz = fmaf(c.fx, dz1 - dz0, dz0);
  return fmaf(c.fx, dvx, q0);
}

// a + f (b - a), rounded by one fma after the difference, so every kernel
// that inlines it rounds the same way whatever else the compiler contracts
// around it. The difference b - a is also the edge derivative the gradient
// needs: value and gradient share it (lerp_value_grad).
__device__ inline float lerp1(float a, float b, float f) { return fmaf(f, b - a, a); }

__device__ inline float lerp_value(const Corners& v, const Cell& c) {
  const float c00 = lerp1(v.v000, v.v001, c.fx);
  const float c01 = lerp1(v.v010, v.v011, c.fx);
  const float c10 = lerp1(v.v100, v.v101, c.fx);
  const float c11 = lerp1(v.v110, v.v111, c.fx);
  const float c0 = lerp1(c00, c01, c.fy);
  const float c1 = lerp1(c10, c11, c.fy);
  return lerp1(c0, c1, c.fz);
}

// Gradient of the trilinear interpolant with respect to the sample position
// (voxel units), at fractions (fx, fy, fz) of the cell.
__device__ inline void lerp_grad(const Corners& v, const Cell& c, float& dvx, float& dvy, float& dvz) {
  const float fx = c.fx, fy = c.fy, fz = c.fz;
  dvx = ((v.v001 - v.v000) * (1.0f - fy) + (v.v011 - v.v010) * fy) * (1.0f - fz) +
        ((v.v101 - v.v100) * (1.0f - fy) + (v.v111 - v.v110) * fy) * fz;
  dvy = ((v.v010 - v.v000) * (1.0f - fx) + (v.v011 - v.v001) * fx) * (1.0f - fz) +
        ((v.v110 - v.v100) * (1.0f - fx) + (v.v111 - v.v101) * fx) * fz;
  dvz = ((v.v100 - v.v000) * (1.0f - fx) + (v.v101 - v.v001) * fx) * (1.0f - fy) +
        ((v.v110 - v.v010) * (1.0f - fx) + (v.v111 - v.v011) * fx) * fy;
}

// Value and gradient together, the gradient from the interpolation's own
// differences (the x edges d_ab = v_ab1 - v_ab0, the y differences
// e_a = c_a1 - c_a0 of the x-lerped edges c_ab, and c_1 - c_0):
//   d/dz = c_1 - c_0,  d/dy = lerp_z(e_0, e_1),  d/dx = lerp_z(lerp_y(d_00, d_01), lerp_y(d_10, d_11)),
// the same derivatives as lerp_grad in about a third of the instructions. The
// value is lerp_value's, bit for bit.
__device__ inline float lerp_value_grad(const Corners& v, const Cell& c, float& dvx, float& dvy, float& dvz) {
  const float d00 = v.v001 - v.v000, d01 = v.v011 - v.v010, d10 = v.v101 - v.v100, d11 = v.v111 - v.v110;
  const float c00 = fmaf(c.fx, d00, v.v000), c01 = fmaf(c.fx, d01, v.v010);
  const float c10 = fmaf(c.fx, d10, v.v100), c11 = fmaf(c.fx, d11, v.v110);
  const float e0 = c01 - c00, e1 = c11 - c10;
  const float c0 = fmaf(c.fy, e0, c00), c1 = fmaf(c.fy, e1, c10);
  dvz = c1 - c0;
  dvy = lerp1(e0, e1, c.fz);
  dvx = lerp1(lerp1(d00, d01, c.fy), lerp1(d10, d11, c.fy), c.fz);
  return fmaf(c.fz, dvz, c0);
}

// Coefficient layout: one record per cell on the stencil's grid holding the
// coefficients of the cell's trilinear interpolant
//   v = c0 + cx fx + cy fy + cxy fx fy + cz fz + cxz fx fz + cyz fy fz + cxyz fx fy fz
// in that order (16 bytes for 16-bit records, 32 for fp32), after a
// kCoefHeader-byte header. The gradient's edge differences are stored
// themselves, so a 16-bit record rounds them relative to their own size
// rather than to the voxel values', and value and gradient cost 11 FMAs
// instead of the corner form's 22 operations. 16-bit records hold every
// coefficient times a power of two s (header: s, 1/s) chosen from the
// volume's largest magnitude so that none overflows fp16 and the differences
// stay clear of its subnormals; bf16 records keep the value c0 in bf16 and the
// seven differences in fp16, fp16 records all eight in fp16. The kernels
// multiply each ray's sums by 1/s once (exact).
constexpr uint32_t kCoefHeader = 256;

struct Coef {
  float c0, cx, cy, cxy, cz, cxz, cyz, cxyz;
};

template <int DT> __device__ inline Coef coef_of(const uint4& q) {
  if constexpr (DT == NDRR_BF16)
    return {bf16_lo(q.x), f16_hi(q.x), f16_lo(q.y), f16_hi(q.y),
            f16_lo(q.z), f16_hi(q.z), f16_lo(q.w), f16_hi(q.w)};
  else
    return {f16_lo(q.x), f16_hi(q.x), f16_lo(q.y), f16_hi(q.y),
            f16_lo(q.z), f16_hi(q.z), f16_lo(q.w), f16_hi(q.w)};
}

template <int DT> struct Fetch<DT, NDRR_LAYOUT_COEF> {  // bf16 / fp16: 16-byte records
  static constexpr int kInFlight = NDRR_IN_FLIGHT;
  uint4 q;
  __device__ inline void load(const Src& s, const Cell& c, bool live) { q = load16(s.vrs, record_offset(s, c, s.W + 1, 16u, live)); }
  __device__ inline Coef coef() const { return coef_of<DT>(q); }
};

// Labelled coefficient records (bf16 / fp16 only): the coefficient layout's
// records with a scale that keeps every slot below 2^-7, so that the two top
// exponent bits of each fp16 slot are zero, and so are the sign and top
// exponent bit of a bf16 c0 (the volume is not negative). Those two bits of
// slot j hold the class code of the cell corner j = x + 2y + 4z (the slots'
// own order): bits 13-14 of the 16-bit slot, bits 14-15 for a bf16 c0. The
// class code is the class for classes 0..3; code 3 also stands for a dropped
// voxel (-1) when there are at most three classes. A sample's nearest voxel
// is a corner of its cell (or, for a cell with no record, a zero record:
// class 0 and value 0), so its class comes with the record the sample loads
// anyway, and the march reads no label map.
constexpr uint32_t kLcoefMaskFp16 = 0x9FFF9FFFu;  // a word of two fp16 slots without their class bits
constexpr uint32_t kLcoefMaskBf16 = 0x9FFF3FFFu;  // the bf16 c0 / fp16 cx word without its class bits
template <int DT> struct Fetch<DT, NDRR_LAYOUT_LCOEF> {
  static_assert(DT != NDRR_F32, "labelled coefficient records are 16-bit");
  static constexpr int kInFlight = NDRR_IN_FLIGHT_LCOEF;
  uint4 q;
  __device__ inline void load(const Src& s, const Cell& c, bool live) { q = load16(s.vrs, record_offset(s, c, s.W + 1, 16u, live)); }
  __device__ inline Coef coef() const {
    constexpr uint32_t m0 = DT == NDRR_BF16 ? kLcoefMaskBf16 : kLcoefMaskFp16;
    return coef_of<DT>(make_uint4(q.x & m0, q.y & kLcoefMaskFp16, q.z & kLcoefMaskFp16, q.w & kLcoefMaskFp16));
  }
  // Class code of the corner at offsets (ox, oy, oz) given as xy = kRintBias * 3 + ox + 2 oy
  // and z = kRintBias + oz (nearest_corner; mod 2^32): the corner's slot is ox + 2 oy in the
  // record's first 8 bytes for oz = 0, its last 8 for oz = 1, so the code lies at bit
  // 16 (ox + 2 oy) + 13 of that 64-bit half (+ 1 for a bf16 c0). (Selects of loaded values:
  // selecting between the members themselves lets the compiler index the record through
  // scratch memory.)
  __device__ inline int label(uint32_t xy, uint32_t z) const {
    const uint32_t qx = q.x, qy = q.y, qz = q.z, qw = q.w;
    const bool hz = z != kRintBias;
    const uint64_t half = hz ? (static_cast<uint64_t>(qw) << 32 | qz) : (static_cast<uint64_t>(qy) << 32 | qx);
    uint32_t sh = (xy - 3u * kRintBias) * 16u + 13u;
    if constexpr (DT == NDRR_BF16) sh += (xy == 3u * kRintBias && !hz) ? 1u : 0u;
    return static_cast<int>((half >> sh) & 3u);
  }
};

// 1/s of the coefficient records (1 for the other layouts), read from the header.
template <int L> __device__ inline float record_unscale(const Params& p) {
  if constexpr (has_header(L)) return static_cast<const float*>(p.vol)[2];
  return 1.0f;
}

template <> struct Fetch<NDRR_F32, NDRR_LAYOUT_COEF> {  // fp32: 32-byte records
  static constexpr int kInFlight = 2;
  uint4 a, b;
  __device__ inline void load(const Src& s, const Cell& c, bool live) {
    const uint32_t off = record_offset(s, c, s.W + 1, 32u, live);
    a = load16(s.vrs, off);
    b = load16(s.vrs, off + 16u);
  }
  __device__ inline Coef coef() const {
    return {__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z), __uint_as_float(a.w),
            __uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w)};
  }
};

// Value of the interpolant, Horner-style in fz, then fy and fx: 7 FMAs.
__device__ inline float coef_value(const Coef& k, const Cell& c) {
  const float u = fmaf(c.fy, k.cxy, k.cx);                  // cx + fy cxy
  const float txz = fmaf(c.fx, k.cxyz, k.cyz);              // cyz + fx cxyz
  const float dz = fmaf(c.fy, txz, fmaf(c.fx, k.cxz, k.cz));  // d/dfz
  return fmaf(c.fz, dz, fmaf(c.fy, k.cy, fmaf(c.fx, u, k.c0)));
}

// Value and gradient (voxel units): 11 FMAs.
__device__ inline float coef_value_grad(const Coef& k, const Cell& c, float& dvx, float& dvy, float& dvz) {
  const float u = fmaf(c.fy, k.cxy, k.cx);      // cx + fy cxy
  const float tyz = fmaf(c.fy, k.cxyz, k.cxz);  // cxz + fy cxyz
  const float txz = fmaf(c.fx, k.cxyz, k.cyz);  // cyz + fx cxyz
  dvx = fmaf(c.fz, tyz, u);
  dvy = fmaf(c.fz, txz, fmaf(c.fx, k.cxy, k.cy));
  dvz = fmaf(c.fy, txz, fmaf(c.fx, k.cxz, k.cz));
  return fmaf(c.fz, dvz, fmaf(c.fy, k.cy, fmaf(c.fx, u, k.c0)));
}

// Interpolation of a fetched cell: the corner layouts interpolate their
// corners (the lerp order of _fused.py:214-220), the coefficient layout
// evaluates its polynomial. interp_grad is the two-pass adjoint's gradient.
template <int DT, int L> __device__ inline float interp_value(const Fetch<DT, L>& f, const Cell& c) {
  return lerp_value(f.corners(), c);
}
template <int DT> __device__ inline float interp_value(const Fetch<DT, NDRR_LAYOUT_COEF>& f, const Cell& c) {
  return coef_value(f.coef(), c);
}
template <int DT, int L>
__device__ inline float interp_value_grad(const Fetch<DT, L>& f, const Cell& c, float& dvx, float& dvy, float& dvz) {
  return lerp_value_grad(f.corners(), c, dvx, dvy, dvz);
}
template <int DT>
__device__ inline float interp_value_grad(const Fetch<DT, NDRR_LAYOUT_COEF>& f, const Cell& c, float& dvx, float& dvy,
                                          float& dvz) {
  return coef_value_grad(f.coef(), c, dvx, dvy, dvz);
}
template <int DT, int L>
__device__ inline void interp_grad(const Fetch<DT, L>& f, const Cell& c, float& dvx, float& dvy, float& dvz) {
  lerp_grad(f.corners(), c, dvx, dvy, dvz);
}
template <int DT>
__device__ inline void interp_grad(const Fetch<DT, NDRR_LAYOUT_COEF>& f, const Cell& c, float& dvx, float& dvy,
                                   float& dvz) {
  coef_value_grad(f.coef(), c, dvx, dvy, dvz);
}
template <int DT> __device__ inline float interp_value(const Fetch<DT, NDRR_LAYOUT_LCOEF>& f, const Cell& c) {
  return coef_value(f.coef(), c);
}
template <int DT> __device__ inline float interp_value(const Fetch<DT, NDRR_LAYOUT_QCOEF>& f, const Cell& c) {
  float dvx, dvy, dvz;  // the derivatives share the value's products; unused ones are dropped
  return qcoef_value_grad(f, c, dvx, dvy, dvz);
}
template <int DT>
__device__ inline float interp_value_grad(const Fetch<DT, NDRR_LAYOUT_QCOEF>& f, const Cell& c, float& dvx,
                                          float& dvy, float& dvz) {
  return qcoef_value_grad(f, c, dvx, dvy, dvz);
}
template <int DT>
__device__ inline void interp_grad(const Fetch<DT, NDRR_LAYOUT_QCOEF>& f, const Cell& c, float& dvx, float& dvy,
                                   float& dvz) {
  qcoef_value_grad(f, c, dvx, dvy, dvz);
}
template <int DT>
__device__ inline float interp_value_grad(const Fetch<DT, NDRR_LAYOUT_LCOEF>& f, const Cell& c, float& dvx,
                                          float& dvy, float& dvz) {
  return coef_value_grad(f.coef(), c, dvx, dvy, dvz);
}
template <int DT>
__device__ inline void interp_grad(const Fetch<DT, NDRR_LAYOUT_LCOEF>& f, const Cell& c, float& dvx, float& dvy,
                                   float& dvz) {
  coef_value_grad(f.coef(), c, dvx, dvy, dvz);
}

// Nearest label (round half to even, zero padded: outside the volume is class 0).
__device__ inline int nearest_label(const int8_t* __restrict__ lab, int D, int H, int W,
                                    float x, float y, float z, bool live) {
  const int jx = static_cast<int>(rintf(x));
  const int jy = static_cast<int>(rintf(y));
  const int jz = static_cast<int>(rintf(z));
  const bool inb = live & (static_cast<unsigned>(jx) < static_cast<unsigned>(W)) &
                   (static_cast<unsigned>(jy) < static_cast<unsigned>(H)) &
                   (static_cast<unsigned>(jz) < static_cast<unsigned>(D));
  const int cx = min(max(jx, 0), W - 1), cy = min(max(jy, 0), H - 1), cz = min(max(jz, 0), D - 1);
  const int l = static_cast<int>(lab[(cz * H + cy) * W + cx]);
  return inb ? l : 0;
}

// Nearest label from label quads (the quad layout's labels: one uint32 per
// quad record holding the int8 classes lab[z0+a][y0+b][x] in byte 2a + b, zero
// outside the volume). The nearest voxel, rint(p) (round half to even, as
// grid_sample's nearest mode), is the corner of the sample's cell at offset 1
// along each axis where rint(p) > floor(p). Branch-free.
__device__ inline int quad_label(const Src& s, const Cell& c, float x, float y, float z, bool live) {
  const uint32_t off = record_offset(s, c, s.W + 2, 4u, live);
  const uint2 w = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(s.lrs, off, 0, 0));
  // rint(p) > floor(p) exactly when rint rounds p up, i.e. rint(p) > p
  const bool ox = rintf(x) > x;
  const bool oy = rintf(y) > y;
  const bool oz = rintf(z) > z;
  const uint32_t word = ox ? w.y : w.x;
  const uint32_t shift = (oz ? 16u : 0u) | (oy ? 8u : 0u);
  return static_cast<int8_t>(static_cast<uint8_t>(word >> shift));
}

// rint(x) (round half to even, the default rounding mode) for |x| < 2^22, in the low bits:
// x + 1.5 * 2^23 lies in [2^23, 2^24), where the spacing of floats is 1, so the sum is x
// rounded to an integer, whose bits are kRintBias + rint(x): bits 0..23 hold 2^22 + rint(x).
__device__ inline int rint_bits(float x) { return __float_as_int(x + 12582912.0f); }

// The corner of cell c nearest to sample (x, y, z): the offsets o = rint(p)
// (round half to even) - floor(p) along each axis, 0 or 1, from rint_bits
// (|p| < 2^22: the sample lies in the volume's box or its cell has no record,
// whose zero record gives class 0 whatever corner is named), returned as
// kRintBias * 3 + ox + 2 oy, with kRintBias + oz in z (both mod 2^32; the bias
// is removed where the offsets are used).
__device__ inline uint32_t nearest_corner(const Cell& c, float x, float y, float z, uint32_t& oz) {
  const uint32_t bx = static_cast<uint32_t>(rint_bits(x) - c.ix), by = static_cast<uint32_t>(rint_bits(y) - c.iy);
  oz = static_cast<uint32_t>(rint_bits(z) - c.iz);
  return bx + 2u * by;
}

// Nearest label from the padded class map of the stencil and coefficient
// layouts: int8 [D + 4][H + 4][W + 4] holding the class of voxel (z, y, x) at
// (z + 2, y + 2, x + 2) and class 0 around it. The nearest voxel is rint(p)
// (round half to even, grid_sample's nearest mode). For a sample whose cell
// exists it is a corner of that cell, at most one voxel outside the volume, so
// it lands in the border (class 0, as the reference gives it): no bounds test.
// A sample whose cell does not exist, or that is not live, fetched a zero
// record (record_offset), so the class it is routed to receives exact zeros;
// its offset may point anywhere, and buffer loads past the end read zeros.
// The offset comes from rint_bits by signed 24-bit multiply-adds
// ((H + 4)(W + 4) < 2^23, checked on the host), which read the 24-bit fields
// 2^22 + rint(p) of the factors and the whole addend (kRintBias + rint(x)):
// s.lab_base removes the 2^22 (hw4 + w4) + kRintBias this adds and adds the
// border (both modulo 2^32).
__device__ inline int padded_label(const Src& s, float x, float y, float z) {
  const int w4 = s.W + 4, hw4 = (s.H + 4) * w4;
  const int off = mad_i24(rint_bits(z), hw4, mad_i24(rint_bits(y), w4, rint_bits(x)));
  return static_cast<int8_t>(
      __builtin_amdgcn_raw_buffer_load_b8(s.lrs, static_cast<uint32_t>(off) + s.lab_base, 0, 0));
}

// a += x rounded once (never contracted with the product that formed x).
__device__ inline void add_rn(float& a, float x) {
#pragma clang fp contract(off)
  a = a + x;
}

// Neumaier-compensated accumulation: sum + comp carries ~2x fp32 precision.
__device__ inline void comp_add(float& sum, float& comp, float x) {
#pragma clang fp contract(off)
  const float t = sum + x;
  comp += (fabsf(sum) >= fabsf(x)) ? ((sum - t) + x) : ((x - t) + sum);
  sum = t;
}

// One batch of the march: positions, cells, loads (and labels) of up to
// kInFlight samples k0 .. k0 + kInFlight - 1 are issued before any is used.
// Samples past the ray's range re-load its last sample and get weight 0.
template <int DT, int L, bool LABELS>
struct Batch {
  static constexpr int U = Fetch<DT, L>::kInFlight;
  Fetch<DT, L> f[U];
  Cell cell[U];
  int cls[U];
  uint32_t cz[U];  // labelled records: the nearest corner's z offset (nearest_corner)
  __device__ inline void issue(const Src& s, const Ray& r, int k0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      // A sample past the ray's last (not live) repeats the last one's position and reads zeros
      // (its record and label loads go past their buffers), so it adds nothing.
      const bool live = k0 + u < r.n_s;
      const float k = static_cast<float>(r.k_base + min(k0 + u, r.n_s - 1));
      const float x = sample_coord(r.ax, r.dx, k), y = sample_coord(r.ay, r.dy, k), z = sample_coord(r.az, r.dz, k);
      cell[u] = cell_of(x, y, z);
      f[u].load(s, cell[u], live);
      if constexpr (LABELS) {
        if constexpr (is_quad(L))
          cls[u] = quad_label(s, cell[u], x, y, z, live);
        else if constexpr (L == NDRR_LAYOUT_LINEAR)
          cls[u] = nearest_label(s.lab, s.D, s.H, s.W, x, y, z, live);
        else if constexpr (L == NDRR_LAYOUT_LCOEF)
          cls[u] = static_cast<int>(nearest_corner(cell[u], x, y, z, cz[u]));
        else
          cls[u] = padded_label(s, x, y, z);
      }
    }
  }
  // Class of sample u (labelled records: from the loaded record, once it has arrived).
  __device__ inline int label(int u) const {
    if constexpr (L == NDRR_LAYOUT_LCOEF) return f[u].label(static_cast<uint32_t>(cls[u]), cz[u]);
    else return cls[u];
  }
};

// ---------------------------------------------------------------------------
// Forward
// ---------------------------------------------------------------------------
// A workgroup of NCH wave groups marches one tile of rays; group c covers
// chunk c of every ray's clipped sample range, so a launch holds NCH x as
// many waves, each with a shorter life. The chunks' partial sums are added in
// chunk order through LDS (deterministic).
// CV == 1: single channel (no label map). CV in {2, 4, 8}: label map, runtime C <= CV.
// JAC: the single-pass adjoint's per-ray, per-class sums are accumulated in
// the same march and written to p.jac: G_c = sum grad v, T_c = sum t grad v
// over the samples routed to class c (voxel units), and with jac_q == 7 also
// V_c = sum v (the image before the step factor); chunk partials are added
// in chunk order through LDS like the image sums.
template <int DT, int L, int CV, bool KAHAN, int NCH, bool JAC>
__global__ __launch_bounds__(kTileRays * NCH) __attribute__((amdgpu_waves_per_eu(NDRR_WAVES_PER_EU)))
void k_fwd(Params p) {
  int b, tile;
  block_coords(b, tile);
  const int tr = threadIdx.x % kTileRays, chunk = threadIdx.x / kTileRays;
  bool valid;
  const int n = ray_index(p, tile, tr, valid);
  __shared__ float sM[12];
  const float* __restrict__ Mb = pose_M(p, b, sM);
  const Ray r = setup_ray(p, Mb, b, n, valid, chunk, NCH);
  const Src src = make_src(p);

  float acc[CV];
  float comp[CV];
#pragma unroll
  for (int c = 0; c < CV; ++c) { acc[c] = 0.0f; comp[c] = 0.0f; }
  constexpr int kJ = JAC ? 6 * CV : 1;
  float jg[kJ];  // per class: G x, y, z, T x, y, z
#pragma unroll
  for (int q = 0; q < kJ; ++q) jg[q] = 0.0f;

  using B = Batch<DT, L, (CV > 1)>;
  for (int k0 = 0; k0 < r.n_s; k0 += B::U) {
    B batch;
    batch.issue(src, r, k0);
#pragma unroll
    for (int u = 0; u < B::U; ++u) {
      // A sample that is not live fetched zeros (Batch::issue): its value and gradient are 0.
      const Cell& c = batch.cell[u];
      float dvx = 0.0f, dvy = 0.0f, dvz = 0.0f;
      const float val = JAC ? interp_value_grad(batch.f[u], c, dvx, dvy, dvz) : interp_value(batch.f[u], c);
      // The t-weighted sums are kept as sums of k grad v, k = the sample's index in this chunk:
      // t = t0 + k / (S - 1) with t0 the chunk's first t, so T = t0 G + K / (S - 1) (formed
      // after the march), one multiply-add per component here.
      const float kf = static_cast<float>(k0 + u);
      // Class routing by branch: a class's sums are updated under the mask of
      // the lanes whose sample it is, so a wave pays for the classes present
      // among its lanes at this sample (mostly one or two), not for all CV.
#pragma unroll
      for (int cl = 0; cl < CV; ++cl) {
        if (CV == 1 || batch.label(u) == cl) {
          if constexpr (KAHAN) comp_add(acc[cl], comp[cl], val); else add_rn(acc[cl], val);
          if constexpr (JAC) {
            add_rn(jg[6 * cl + 0], dvx);
            add_rn(jg[6 * cl + 1], dvy);
            add_rn(jg[6 * cl + 2], dvz);
            jg[6 * cl + 3] = fmaf(kf, dvx, jg[6 * cl + 3]);
            jg[6 * cl + 4] = fmaf(kf, dvy, jg[6 * cl + 4]);
            jg[6 * cl + 5] = fmaf(kf, dvz, jg[6 * cl + 5]);
          }
        }
      }
    }
  }
  if constexpr (JAC) {
    const float t0 = fmaf(static_cast<float>(r.s_lo), p.inv_s, r.t_off);
#pragma unroll
    for (int cl = 0; cl < CV; ++cl) {
#pragma unroll
      for (int a = 0; a < 3; ++a) jg[6 * cl + 3 + a] = fmaf(t0, jg[6 * cl + a], p.inv_s * jg[6 * cl + 3 + a]);
    }
  }

  __shared__ float part[NCH > 1 ? NCH - 1 : 1][CV][kTileRays];
  constexpr bool kJpart = JAC && NCH > 1;
  __shared__ float jpart[kJpart ? NCH - 1 : 1][kJpart ? 6 * CV : 1][kJpart ? kTileRays : 1];
  if constexpr (NCH > 1) {
    if (chunk > 0) {
#pragma unroll
      for (int c = 0; c < CV; ++c) part[chunk - 1][c][tr] = KAHAN ? (acc[c] + comp[c]) : acc[c];
      if constexpr (kJpart) {
#pragma unroll
        for (int q = 0; q < 6 * CV; ++q) jpart[chunk - 1][q][tr] = jg[q];
      }
    }
    __syncthreads();
  }
  if (chunk != 0 || !valid) return;
  const float unscale = record_unscale<L>(p);
  float* o = p.out + (static_cast<int64_t>(b) * p.C) * p.N + n;
#pragma unroll
  for (int c = 0; c < CV; ++c) {
    if (c < p.C) {
      float total = acc[c], tcomp = comp[c];
#pragma unroll
      for (int k = 1; k < NCH; ++k) {
        if constexpr (KAHAN) comp_add(total, tcomp, part[k - 1][c][tr]); else total += part[k - 1][c][tr];
      }
      const float sum = (KAHAN ? (total + tcomp) : total) * unscale;
      o[static_cast<int64_t>(c) * p.N] = sum * r.step;
      if constexpr (JAC) {
        float* jp = p.jac + (static_cast<int64_t>(b) * p.jac_q * p.C + c) * p.N + n;
        const int64_t plane = static_cast<int64_t>(p.C) * p.N;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          float j = jg[6 * c + q];
          if constexpr (kJpart) {
#pragma unroll
            for (int k = 1; k < NCH; ++k) j += jpart[k - 1][6 * c + q][tr];
          }
          jp[q * plane] = j * unscale;
        }
        if (p.jac_q == 7) jp[6 * plane] = sum;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Backward (analytic adjoint)
// ---------------------------------------------------------------------------
// With q(t) = src + t e the camera-space sample, dL/dM[i][j] sums
// gp_i q_j = gp_i src_j + (gp_i t) e_j over the ray, and dL/dsrc, dL/dtgt sum
// (1 - t) M^T gp and t M^T gp: six per-ray sums G = sum gp and T = sum t gp
// give all of them (the reference's kernel keeps 19, _fused.py:168-186).
__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ inline int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}


// Ordering of a wave's own LDS accesses: its earlier LDS operations have completed before
// any lane goes on (no workgroup barrier: only this wave touches the region).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Where the merged volume-gradient scatter adds voxel (x, y, z): the brick buffer when given
// (4 x 4 x 4-voxel bricks, x fastest within a brick and among bricks: a 128-byte line holds a
// 4 x 4 x 2 block, so the voxels a wave touches at one step fill about half as many lines as
// in the [D][H][W] volume, where each (y, z) row is a line), else gvol itself.
__device__ inline float* gvol_voxel(const Params& p, int x, int y, int z) {
  if (p.gvol_bricks == nullptr) return p.gvol + (z * p.H + y) * p.W + x;
  const int brick = ((z >> 2) * p.nby + (y >> 2)) * p.nbx + (x >> 2);
  return p.gvol_bricks + brick * 64 + ((z & 3) * 16 + (y & 3) * 4 + (x & 3));
}

// Volume-gradient scatter through a wave-local LDS box of kBox^3 voxels, anchored at the cell
// of lane 36 (the middle of the wave's 8 x 8 ray patch) less 3 along each axis: each lane adds
// its sample's eight corner contributions into the box with LDS atomics where they fall inside
// it (the cells within 3 of the anchor's), else straight to memory; then each lane takes eight
// box voxels, adds every nonzero one to memory with one atomic and clears them. Adjacent rays'
// samples share corners, so a wave-step's 512 corner updates become one atomic per voxel
// touched. Every lane of the wave runs it (the merged march's wave-wide loop).
constexpr int kBox = 8;
__device__ inline void gvol_add_box(const Params& p, const Cell& c, float g, float* __restrict__ box) {
  const int lane = threadIdx.x & 63;
  const int ax = __builtin_amdgcn_readlane(c.ix, 36) - 3;
  const int ay = __builtin_amdgcn_readlane(c.iy, 36) - 3;
  const int az = __builtin_amdgcn_readlane(c.iz, 36) - 3;
  const float wx[2] = {1.0f - c.fx, c.fx};
  const float wy[2] = {1.0f - c.fy, c.fy};
  const float wz[2] = {1.0f - c.fz, c.fz};
  if (g != 0.0f) {
#pragma unroll
    for (int corner = 0; corner < 8; ++corner) {
      const int cx = corner & 1, cy = (corner >> 1) & 1, cz = corner >> 2;
      const int jx = c.ix + cx, jy = c.iy + cy, jz = c.iz + cz;
      const bool inb = static_cast<unsigned>(jx) < static_cast<unsigned>(p.W) &&
                       static_cast<unsigned>(jy) < static_cast<unsigned>(p.H) &&
                       static_cast<unsigned>(jz) < static_cast<unsigned>(p.D);
      if (!inb) continue;
      const float v = g * (wx[cx] * wy[cy] * wz[cz]);
      const unsigned bx = static_cast<unsigned>(jx - ax), by = static_cast<unsigned>(jy - ay),
                     bz = static_cast<unsigned>(jz - az);
      if ((bx | by | bz) < static_cast<unsigned>(kBox))
        atomicAdd(box + (bz * kBox + by) * kBox + bx, v);
      else
        unsafeAtomicAdd(gvol_voxel(p, jx, jy, jz), v);
    }
  }
  wave_lds_sync();
  // lane: box row (bz, by) = (lane / 8, lane % 8), voxels bx = 0 .. 7
  float4* row = reinterpret_cast<float4*>(box + kBox * lane);
  const float4 lo = row[0], hi = row[1];
  const float v8[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
  const int jy = ay + (lane & 7), jz = az + (lane >> 3);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int jx = ax + k;
    // only voxels inside the volume were added; the test also keeps the address in range
    const bool inb = static_cast<unsigned>(jx) < static_cast<unsigned>(p.W) &&
                     static_cast<unsigned>(jy) < static_cast<unsigned>(p.H) &&
                     static_cast<unsigned>(jz) < static_cast<unsigned>(p.D);
    if (inb && v8[k] != 0.0f) unsafeAtomicAdd(gvol_voxel(p, jx, jy, jz), v8[k]);
  }
  row[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  row[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  wave_lds_sync();
}

template <int DT, int L, int CV, bool GCAM, bool GVOL, int NCH>
__global__ __launch_bounds__(kTileRays * NCH) __attribute__((amdgpu_waves_per_eu(NDRR_WAVES_PER_EU)))
void k_bwd(Params p) {
  int b, tile;
  block_coords(b, tile);
  const int tr = threadIdx.x % kTileRays, chunk = threadIdx.x / kTileRays;
  bool valid;
  const int n = ray_index(p, tile, tr, valid);
  __shared__ float sM[12];
  const float* __restrict__ Mb = pose_M(p, b, sM);
  const Ray r = setup_ray(p, Mb, b, n, valid, chunk, NCH);
  const Src src = make_src(p);

  // Upstream gradient of this ray, one value per class.
  float goc[CV];
  {
    const int nn = valid ? n : 0;
    const float* g = p.go + (static_cast<int64_t>(b) * p.C) * p.N + nn;
#pragma unroll
    for (int c = 0; c < CV; ++c)
      goc[c] = (valid && c < p.C) ? g[static_cast<int64_t>(c) * p.N] : 0.0f;
  }

  float Gx = 0.f, Gy = 0.f, Gz = 0.f, Tx = 0.f, Ty = 0.f, Tz = 0.f, astep = 0.f;
  using B = Batch<DT, L, (CV > 1)>;
  // Merged volume-gradient atomics exchange values across the wave every sample: every lane then
  // runs the wave's longest loop, its extra samples not live (they repeat its last sample, or for
  // a ray that misses the volume read zeros or clamped addresses, and add nothing).
  const bool merge = GVOL && p.gvol_merge;
  const int n_loop = merge ? wave_max(r.n_s) : r.n_s;
  float* box = nullptr;  // this wave's LDS box (gvol_add_box), cleared before its first use
  if constexpr (GVOL) {
    __shared__ __attribute__((aligned(16))) float gbox[(kTileRays / kWave) * NCH][kBox * kBox * kBox];
    box = gbox[threadIdx.x / kWave];
    float4* mine = reinterpret_cast<float4*>(box + kBox * (threadIdx.x & 63));
    mine[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    mine[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    wave_lds_sync();
  }
  for (int k0 = 0; k0 < n_loop; k0 += B::U) {
    B batch;
    batch.issue(src, r, k0);
#pragma unroll
    for (int u = 0; u < B::U; ++u) {
      const bool live = k0 + u < r.n_s;
      const Cell& c = batch.cell[u];
      float gg;
      if constexpr (CV == 1) {
        gg = goc[0];
      } else {
        gg = 0.0f;
#pragma unroll
        for (int cl = 0; cl < CV; ++cl) gg = (batch.label(u) == cl) ? goc[cl] : gg;
      }
      gg = live ? gg : 0.0f;
      const float g = gg * r.step;
      const float fx = c.fx, fy = c.fy, fz = c.fz;
      float dvx, dvy, dvz;
      interp_grad(batch.f[u], c, dvx, dvy, dvz);
      const float gx = g * dvx, gy = g * dvy, gz = g * dvz;
      const float t = fmaf(static_cast<float>(r.s_lo + k0 + u), p.inv_s, r.t_off);
      Gx += gx; Gy += gy; Gz += gz;
      Tx += gx * t; Ty += gy * t; Tz += gz * t;
      if constexpr (GCAM) astep += gg * interp_value(batch.f[u], c);
      if constexpr (GVOL) {
        if (merge) {
          gvol_add_box(p, c, g, box);
        } else if (g != 0.0f) {
          const float wx[2] = {1.0f - fx, fx};
          const float wy[2] = {1.0f - fy, fy};
          const float wz[2] = {1.0f - fz, fz};
#pragma unroll
          for (int corner = 0; corner < 8; ++corner) {
            const int cx = corner & 1, cy = (corner >> 1) & 1, cz = corner >> 2;
            const int jx = c.ix + cx, jy = c.iy + cy, jz = c.iz + cz;
            const bool inb = static_cast<unsigned>(jx) < static_cast<unsigned>(p.W) &&
                             static_cast<unsigned>(jy) < static_cast<unsigned>(p.H) &&
                             static_cast<unsigned>(jz) < static_cast<unsigned>(p.D);
            if (inb) unsafeAtomicAdd(p.gvol + (jz * p.H + jy) * p.W + jx, g * (wx[cx] * wy[cy] * wz[cz]));
          }
        }
      }
    }
  }

  if constexpr (has_header(L)) {
    const float unscale = record_unscale<L>(p);
    Gx *= unscale; Gy *= unscale; Gz *= unscale;
    Tx *= unscale; Ty *= unscale; Tz *= unscale;
    astep *= unscale;
  }

  if constexpr (GCAM) {
    // Per-ray totals over the chunks, in chunk order.
    __shared__ float part[NCH > 1 ? NCH - 1 : 1][7][kTileRays];
    if constexpr (NCH > 1) {
      if (chunk > 0) {
        part[chunk - 1][0][tr] = Gx; part[chunk - 1][1][tr] = Gy; part[chunk - 1][2][tr] = Gz;
        part[chunk - 1][3][tr] = Tx; part[chunk - 1][4][tr] = Ty; part[chunk - 1][5][tr] = Tz;
        part[chunk - 1][6][tr] = astep;
      }
      __syncthreads();
    }
    if (chunk == 0 && valid) {
      float sGx = Gx, sGy = Gy, sGz = Gz, sTx = Tx, sTy = Ty, sTz = Tz, sA = astep;
#pragma unroll
      for (int k = 1; k < NCH; ++k) {
        sGx += part[k - 1][0][tr]; sGy += part[k - 1][1][tr]; sGz += part[k - 1][2][tr];
        sTx += part[k - 1][3][tr]; sTy += part[k - 1][4][tr]; sTz += part[k - 1][5][tr];
        sA += part[k - 1][6][tr];
      }
      const float m00 = Mb[0], m01 = Mb[1], m02 = Mb[2];
      const float m10 = Mb[4], m11 = Mb[5], m12 = Mb[6];
      const float m20 = Mb[8], m21 = Mb[9], m22 = Mb[10];
      const float qGx = m00 * sGx + m10 * sGy + m20 * sGz;  // M33^T G
      const float qGy = m01 * sGx + m11 * sGy + m21 * sGz;
      const float qGz = m02 * sGx + m12 * sGy + m22 * sGz;
      const float qTx = m00 * sTx + m10 * sTy + m20 * sTz;  // M33^T T
      const float qTy = m01 * sTx + m11 * sTy + m21 * sTz;
      const float qTz = m02 * sTx + m12 * sTy + m22 * sTz;
      const int64_t ray = static_cast<int64_t>(b) * p.N + n;
      p.gsrc[3 * ray + 0] = qGx - qTx; p.gsrc[3 * ray + 1] = qGy - qTy; p.gsrc[3 * ray + 2] = qGz - qTz;
      p.gtgt[3 * ray + 0] = qTx; p.gtgt[3 * ray + 1] = qTy; p.gtgt[3 * ray + 2] = qTz;
      p.gstep[ray] = sA;
    }
  }

  // Deterministic pose-gradient partials: a butterfly within each wave, then
  // either the wave sums of the tile in wave order (one partial per tile) or
  // one partial per wave, written with no workgroup barrier; k_reduce_pose
  // adds the partials of a pose in a fixed order.
  constexpr int kWaves = kTileRays * NCH / kWave;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const float a[12] = {r.sx * Gx + r.ex * Tx, r.sy * Gx + r.ey * Tx, r.sz * Gx + r.ez * Tx, Gx,
                       r.sx * Gy + r.ex * Ty, r.sy * Gy + r.ey * Ty, r.sz * Gy + r.ez * Ty, Gy,
                       r.sx * Gz + r.ex * Tz, r.sy * Gz + r.ey * Tz, r.sz * Gz + r.ez * Tz, Gz};
  if constexpr (NDRR_BWD_TILE_PARTIALS) {
    __shared__ float red[kWaves][12];
#pragma unroll
    for (int j = 0; j < 12; ++j) {
      const float v = wave_sum(a[j]);
      if (lane == 0) red[wave][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
      const int j = threadIdx.x;
      float v = red[0][j];
#pragma unroll
      for (int w = 1; w < kWaves; ++w) v += red[w][j];
      p.gpart[(static_cast<int64_t>(b) * p.nblocks + tile) * 12 + j] = v;
    }
  } else {
    float* const dst = p.gpart + ((static_cast<int64_t>(b) * p.nblocks + tile) * kWaves + wave) * 12;
#pragma unroll
    for (int j = 0; j < 12; ++j) {
      const float v = wave_sum(a[j]);
      if (lane == j) dst[j] = v;
    }
  }
}

// Single-pass adjoint: the backward as a per-ray contraction of the sums the
// forward (k_fwd<..., JAC>) stored. With g_c = go[b][c][n],
//   G = step sum_c g_c G_c,  T = step sum_c g_c T_c,  dL/dstep = sum_c g_c V_c,
// which are the per-ray sums k_bwd forms sample by sample (the per-sample
// factor go * step moved out of the sums); the pose partials, gsrc and gtgt
// follow from G and T exactly as in k_bwd.
// Rays are taken in their linear order: a workgroup takes kJacRays
// consecutive ray blocks of kTileRays rays (grid (ceil(N / (kTileRays
// kJacRays)), B)) and each thread one ray of every block, so every plane is
// read in full cache lines and several rays' loads are in flight per thread;
// a thread adds its rays' terms in block order, then the workgroup reduces
// once. The pose partials go to gpart[b][blockIdx.x].
#ifndef NDRR_JAC_RAYS
#define NDRR_JAC_RAYS 3
#endif
constexpr int kJacRays = NDRR_JAC_RAYS;

template <int CV, bool GCAM>
__global__ __launch_bounds__(kTileRays) void k_bwd_jac(Params p) {
  const int b = blockIdx.y, blk = blockIdx.x;
  __shared__ float sM[12];
  const float* __restrict__ Mb = GCAM ? pose_M(p, b, sM) : nullptr;
  const int64_t plane = static_cast<int64_t>(p.C) * p.N;
  float a[12];
#pragma unroll
  for (int j = 0; j < 12; ++j) a[j] = 0.0f;
#pragma unroll
  for (int k = 0; k < kJacRays; ++k) {
    const int n = (blk * kJacRays + k) * kTileRays + threadIdx.x;
    const bool valid = n < p.N;
    const int nn = valid ? n : 0;
    const float* sp = p.src + b * p.src_bs + nn * p.src_rs;
    const float* tp = p.tgt + b * p.tgt_bs + 3 * static_cast<int64_t>(nn);
    const float sx = sp[0], sy = sp[1], sz = sp[2];
    const float ex = tp[0] - sx, ey = tp[1] - sy, ez = tp[2] - sz;
    const float step = p.step[b * p.step_bs + nn];
    float Gx = 0.f, Gy = 0.f, Gz = 0.f, Tx = 0.f, Ty = 0.f, Tz = 0.f, astep = 0.f;
#pragma unroll
    for (int c = 0; c < CV; ++c) {
      if (c < p.C) {
        const float g = valid ? p.go[b * p.go_bs + static_cast<int64_t>(c) * p.N + nn] : 0.0f;
        const float* jp = p.jac + (static_cast<int64_t>(b) * p.jac_q * p.C + c) * p.N + nn;
        Gx += g * jp[0]; Gy += g * jp[plane]; Gz += g * jp[2 * plane];
        Tx += g * jp[3 * plane]; Ty += g * jp[4 * plane]; Tz += g * jp[5 * plane];
        if constexpr (GCAM) astep += g * jp[6 * plane];
      }
    }
    Gx *= step; Gy *= step; Gz *= step;
    Tx *= step; Ty *= step; Tz *= step;
    if constexpr (GCAM) {
      if (valid) {
        const float m00 = Mb[0], m01 = Mb[1], m02 = Mb[2];
        const float m10 = Mb[4], m11 = Mb[5], m12 = Mb[6];
        const float m20 = Mb[8], m21 = Mb[9], m22 = Mb[10];
        const float qGx = m00 * Gx + m10 * Gy + m20 * Gz;  // M33^T G
        const float qGy = m01 * Gx + m11 * Gy + m21 * Gz;
        const float qGz = m02 * Gx + m12 * Gy + m22 * Gz;
        const float qTx = m00 * Tx + m10 * Ty + m20 * Tz;  // M33^T T
        const float qTy = m01 * Tx + m11 * Ty + m21 * Tz;
        const float qTz = m02 * Tx + m12 * Ty + m22 * Tz;
        const int64_t ray = static_cast<int64_t>(b) * p.N + n;
        p.gsrc[3 * ray + 0] = qGx - qTx; p.gsrc[3 * ray + 1] = qGy - qTy; p.gsrc[3 * ray + 2] = qGz - qTz;
        p.gtgt[3 * ray + 0] = qTx; p.gtgt[3 * ray + 1] = qTy; p.gtgt[3 * ray + 2] = qTz;
        p.gstep[ray] = astep;
      }
    }
    a[0] += sx * Gx + ex * Tx; a[1] += sy * Gx + ey * Tx; a[2] += sz * Gx + ez * Tx; a[3] += Gx;
    a[4] += sx * Gy + ex * Ty; a[5] += sy * Gy + ey * Ty; a[6] += sz * Gy + ez * Ty; a[7] += Gy;
    a[8] += sx * Gz + ex * Tz; a[9] += sy * Gz + ey * Tz; a[10] += sz * Gz + ez * Tz; a[11] += Gz;
  }
  constexpr int kWaves = kTileRays / kWave;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __shared__ float red[kWaves][12];
#pragma unroll
  for (int j = 0; j < 12; ++j) {
    const float v = wave_sum(a[j]);
    if (lane == 0) red[wave][j] = v;
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    const int j = threadIdx.x;
    float v = red[0][j];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) v += red[w][j];
    p.gpart[(static_cast<int64_t>(b) * gridDim.x + blk) * 12 + j] = v;
  }
}

// gM[b][j] = sum of the nparts partials gpart[b][part][j] (per tile, or per
// wave of every tile) in a fixed order: thread t of the block sums parts t, t + 256,
// ... (compensated, loads independent of one another), then a pairwise tree
// over the 256 thread sums in LDS. Deterministic.
constexpr int kReduceThreads = 256;
// The registration form of the reduction's output: the pose step's parameter
// gradients from scale * gM (scale -1 for a descent direction; exact).
struct Se3Out {
  const double* prefix;
  const double* suffix;
  const float* rot;
  const float* xyz;
  int64_t param_stride;
  float scale;
  float* grot;
  float* gxyz;
};

__global__ __launch_bounds__(kReduceThreads) void k_reduce_pose(const float* __restrict__ part,
                                                               float* __restrict__ gM, int nparts,
                                                               const double* __restrict__ w2v,
                                                               float* __restrict__ g_rt, Se3Out se3) {
  __shared__ float red[12][kReduceThreads];
  const int b = blockIdx.x, t = threadIdx.x;
  const float4* pb = reinterpret_cast<const float4*>(part + static_cast<int64_t>(b) * nparts * 12);
  float sum[12], comp[12];
#pragma unroll
  for (int j = 0; j < 12; ++j) { sum[j] = 0.0f; comp[j] = 0.0f; }
  for (int q = t; q < nparts; q += kReduceThreads) {
    const float4 a = pb[3 * q], c = pb[3 * q + 1], d = pb[3 * q + 2];
    const float v[12] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w, d.x, d.y, d.z, d.w};
#pragma unroll
    for (int j = 0; j < 12; ++j) comp_add(sum[j], comp[j], v[j]);
  }
#pragma unroll
  for (int j = 0; j < 12; ++j) red[j][t] = sum[j] + comp[j];
  __syncthreads();
  for (int w = kReduceThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int j = 0; j < 12; ++j) red[j][t] += red[j][t + w];
    }
    __syncthreads();
  }
  if (gM != nullptr && t < 12) gM[b * 12 + t] = red[t][0];
  if (w2v != nullptr && t < 16) {
    // the rt_inv gradient, W[:3]^T gM in FP64 from the rounded gM (k_camera_to_voxel_bwd's sums)
    const int k = t / 4, j = t % 4;
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) acc = fma(w2v[4 * i + k], static_cast<double>(red[4 * i + j][0]), acc);
    g_rt[b * 16 + t] = static_cast<float>(acc);
  }
  if (se3.grot != nullptr && t == 0) {
    // the pose step's adjoint (the sums of k_pose_compose_bwd) from the rounded, scaled gM
    const int64_t ps = se3.param_stride;
    const double w[3] = {se3.rot[b * ps], se3.rot[b * ps + 1], se3.rot[b * ps + 2]};
    const double v[3] = {se3.xyz[b * ps], se3.xyz[b * ps + 1], se3.xyz[b * ps + 2]};
    double G[3][4];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 4; ++j) G[i][j] = static_cast<double>(se3.scale * red[4 * i + j][0]);
    ndrr_se3::compose_adjoint(se3.prefix + 16 * static_cast<int64_t>(b), se3.suffix + 16 * static_cast<int64_t>(b), w,
                              v, G, se3.grot + b * ps, se3.gxyz + b * ps);
  }
}

__global__ void k_labels_to_i8(const float* __restrict__ src, int8_t* __restrict__ dst, int64_t numel, int C) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= numel) return;
  const float t = truncf(src[i]);
  dst[i] = (t >= 0.0f && t < static_cast<float>(C)) ? static_cast<int8_t>(t) : static_cast<int8_t>(-1);
}

// One thread per stencil record: gather the cell's 8 corners (zero padded)
// from the linear volume and store them as one record (raw element bits).
template <int DT>
__global__ void k_pack_stencil(const typename Vox<DT>::T* __restrict__ vol, void* __restrict__ rec,
                               int D, int H, int W, int64_t n_rec) {
  using T = typename Vox<DT>::T;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n_rec) return;
  const int cx = static_cast<int>(r % (W + 1));
  const int64_t rest = r / (W + 1);
  const int cy = static_cast<int>(rest % (H + 1));
  const int cz = static_cast<int>(rest / (H + 1));
  T v[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int x = cx - 1 + (c & 1), y = cy - 1 + ((c >> 1) & 1), z = cz - 1 + (c >> 2);
    const bool inb = static_cast<unsigned>(x) < static_cast<unsigned>(W) &&
                     static_cast<unsigned>(y) < static_cast<unsigned>(H) &&
                     static_cast<unsigned>(z) < static_cast<unsigned>(D);
    v[c] = inb ? vol[(static_cast<int64_t>(z) * H + y) * W + x] : T(0);
  }
  if constexpr (DT == NDRR_F32) {
    float4* o = static_cast<float4*>(rec) + 2 * r;
    o[0] = make_float4(v[0], v[1], v[2], v[3]);
    o[1] = make_float4(v[4], v[5], v[6], v[7]);
  } else {
    uint32_t w[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t lo = __builtin_bit_cast(uint16_t, v[2 * c]);
      const uint32_t hi = __builtin_bit_cast(uint16_t, v[2 * c + 1]);
      w[c] = lo | (hi << 16);
    }
    static_cast<uint4*>(rec)[r] = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

// One thread per quad record (x, y0, z0): v[z0+a][y0+b][x] for (a, b) = 00, 01, 10, 11.
template <int DT>
__global__ void k_pack_quad(const typename Vox<DT>::T* __restrict__ vol, void* __restrict__ rec,
                            int D, int H, int W, int64_t n_rec) {
  using T = typename Vox<DT>::T;
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n_rec) return;
  const int cx = static_cast<int>(r % (W + 2));
  const int64_t rest = r / (W + 2);
  const int cy = static_cast<int>(rest % (H + 1));
  const int cz = static_cast<int>(rest / (H + 1));
  T v[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int x = cx - 1, y = cy - 1 + (c & 1), z = cz - 1 + (c >> 1);
    const bool inb = static_cast<unsigned>(x) < static_cast<unsigned>(W) &&
                     static_cast<unsigned>(y) < static_cast<unsigned>(H) &&
                     static_cast<unsigned>(z) < static_cast<unsigned>(D);
    v[c] = inb ? vol[(static_cast<int64_t>(z) * H + y) * W + x] : T(0);
  }
  if constexpr (DT == NDRR_F32) {
    static_cast<float4*>(rec)[r] = make_float4(v[0], v[1], v[2], v[3]);
  } else {
    const uint32_t w0 = static_cast<uint32_t>(__builtin_bit_cast(uint16_t, v[0])) |
                        (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, v[1])) << 16);
    const uint32_t w1 = static_cast<uint32_t>(__builtin_bit_cast(uint16_t, v[2])) |
                        (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, v[3])) << 16);
    static_cast<uint2*>(rec)[r] = make_uint2(w0, w1);
  }
}

// fp32 -> the record type, rounded to nearest even once.
template <int DT> struct Round;
template <> struct Round<NDRR_F16> {
  __device__ static inline uint32_t bits(float f) {
    return __builtin_bit_cast(uint16_t, static_cast<_Float16>(f));
  }
};
template <> struct Round<NDRR_F32> {
  __device__ static inline uint32_t bits(float f) { return __float_as_uint(f); }
};
template <> struct Round<NDRR_BF16> {
  // the upper half of the fp32 pattern, rounded to nearest even on the dropped half
  // (NaN stays a quiet NaN, infinities stay infinite)
  __device__ static inline uint32_t bits(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0xffffu) != 0u ? 0x40u : 0u);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
  }
};
// The coefficient layout's header: [0] the volume's largest |value| (fp32
// bits, an atomic maximum), [1] the scale s of the 16-bit records, [2] 1/s,
// [3] the record type; kCoefHeader bytes in all.
__global__ void k_coef_header_init(uint32_t* __restrict__ hdr) {
  if (threadIdx.x < kCoefHeader / 4) hdr[threadIdx.x] = 0u;
}

template <int DS>
__global__ __launch_bounds__(kThreads) void k_coef_absmax(const typename Vox<DS>::T* __restrict__ vol, int64_t n,
                                                          uint32_t* __restrict__ hdr) {
  __shared__ uint32_t red[kThreads / 64];
  uint32_t m = 0u;  // |value| as fp32 bits: non-negative floats order like their bit patterns
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kThreads)
    m = max(m, __float_as_uint(Vox<DS>::ld(vol, static_cast<int>(i))) & 0x7fffffffu);
  for (int o = 32; o > 0; o >>= 1) m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), o)));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; ++w) m = max(m, red[w]);
    atomicMax(hdr, m);
  }
}

// The records' scale: 1 for fp32; else the largest power of two s with
// 8 max|v| s <= 65504 (fp16's largest value; |coefficient| <= 8 max|v|),
// within [2^-120, 2^120]; 1 for a zero or non-finite maximum.
template <int DR> __device__ inline float coef_scale(uint32_t absmax_bits) {
  if constexpr (DR == NDRR_F32) return 1.0f;
  const float m = __uint_as_float(absmax_bits);
  if (!(m > 0.0f) || !(m <= 3.0e38f)) return 1.0f;
  const float t = fminf(fmaxf(8188.0f / m, 0x1p-120f), 0x1p120f);
  return __uint_as_float(__float_as_uint(t) & 0x7f800000u);  // 2^floor(log2 t)
}

// One thread per coefficient record (cell x0, y0, z0 on the stencil's grid):
// the cell's eight zero-padded corners (as fp32) become the coefficients of
// its trilinear interpolant, in fp32 in a fixed order, times the header's
// scale s, each rounded once to its slot's type (fp32; bf16 for c0 and fp16
// for the rest; fp16).
// The cell (x0, y0, z0) = (cx - 1, cy - 1, cz - 1) of record r on the stencil's grid.
__device__ inline void record_cell(int64_t r, int H, int W, int& cx, int& cy, int& cz) {
  cx = static_cast<int>(r % (W + 1));
  const int64_t rest = r / (W + 1);
  cy = static_cast<int>(rest % (H + 1));
  cz = static_cast<int>(rest / (H + 1));
}

// The coefficients of record (cx, cy, cz)'s interpolant from its eight
// zero-padded corners (as fp32), in fp32 in a fixed order.
template <int DS>
__device__ inline void cell_coefficients(const typename Vox<DS>::T* __restrict__ vol, int D, int H, int W, int cx,
                                         int cy, int cz, float c[8]) {
#pragma clang fp contract(off)
  float v[8];  // v[4a + 2b + c] = corner (z0 + a, y0 + b, x0 + c)
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int x = cx - 1 + (k & 1), y = cy - 1 + ((k >> 1) & 1), z = cz - 1 + (k >> 2);
    const bool inb = static_cast<unsigned>(x) < static_cast<unsigned>(W) &&
                     static_cast<unsigned>(y) < static_cast<unsigned>(H) &&
                     static_cast<unsigned>(z) < static_cast<unsigned>(D);
    v[k] = inb ? Vox<DS>::ld(vol, static_cast<int>((static_cast<int64_t>(z) * H + y) * W + x)) : 0.0f;
  }
  const float dx00 = v[1] - v[0], dx01 = v[3] - v[2], dx10 = v[5] - v[4], dx11 = v[7] - v[6];
  const float cyv = v[2] - v[0];
  c[0] = v[0];                  // c0
  c[1] = dx00;                  // cx
  c[2] = cyv;                   // cy
  c[3] = dx01 - dx00;           // cxy
  c[4] = v[4] - v[0];           // cz
  c[5] = dx10 - dx00;           // cxz
  c[6] = (v[6] - v[4]) - cyv;   // cyz
  c[7] = (dx11 - dx10) - (dx01 - dx00);  // cxyz
}

template <int DS, int DR>
__global__ void k_pack_coef(const typename Vox<DS>::T* __restrict__ vol, uint32_t* __restrict__ hdr,
                            void* __restrict__ rec, int D, int H, int W, int64_t n_rec) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const float scale = coef_scale<DR>(hdr[0]);
  if (r == 0) {
    hdr[1] = __float_as_uint(scale);
    hdr[2] = __float_as_uint(1.0f / scale);  // exact: a power of two
    hdr[3] = static_cast<uint32_t>(DR);
  }
  if (r >= n_rec) return;
  int cx, cy, cz;
  record_cell(r, H, W, cx, cy, cz);
  float c[8];
  cell_coefficients<DS>(vol, D, H, W, cx, cy, cz, c);
  if constexpr (DR == NDRR_F32) {
    float4* o = static_cast<float4*>(rec) + 2 * r;
    o[0] = make_float4(c[0], c[1], c[2], c[3]);
    o[1] = make_float4(c[4], c[5], c[6], c[7]);
  } else {
    uint32_t h[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) h[k] = Round<NDRR_F16>::bits(c[k] * scale);  // exact product: a power of two
    if constexpr (DR == NDRR_BF16) h[0] = Round<NDRR_BF16>::bits(c[0] * scale);
    static_cast<uint4*>(rec)[r] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16),
                                             h[6] | (h[7] << 16));
  }
}

// Labelled coefficient records. Header: [0] the largest |coefficient| over
// all records (fp32 bits, an atomic maximum), [1] s, [2] 1/s, [3] the record
// type, [4] what the pack found wrong (NDRR_LCOEF_* bits, 0 when the records
// are valid), [5] the class count they were made for.
template <int DS>
__global__ __launch_bounds__(kThreads) void k_lcoef_absmax(const typename Vox<DS>::T* __restrict__ vol, int D,
                                                           int H, int W, int64_t n_rec,
                                                           uint32_t* __restrict__ hdr) {
  __shared__ uint32_t red[kThreads / 64];
  uint32_t m = 0u;  // |coefficient| as fp32 bits: non-negative floats order like their bit patterns
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; r < n_rec;
       r += static_cast<int64_t>(gridDim.x) * kThreads) {
    int cx, cy, cz;
    record_cell(r, H, W, cx, cy, cz);
    float c[8];
    cell_coefficients<DS>(vol, D, H, W, cx, cy, cz, c);
#pragma unroll
    for (int k = 0; k < 8; ++k) m = max(m, __float_as_uint(c[k]) & 0x7fffffffu);
  }
  for (int o = 32; o > 0; o >>= 1) m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), o)));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; ++w) m = max(m, red[w]);
    atomicMax(hdr, m);
  }
}

// The labelled records' scale: the largest power of two s with
// max|coefficient| s <= 2^-7 - 2^-18, so that every slot rounds to below
// 2^-7 (fp16 exponent field <= 7); within [2^-126, 2^126]; 1 for a zero
// maximum. `fits` is false where no such s exists in that range (a
// non-finite or too large maximum).
__device__ inline float lcoef_scale(uint32_t absmax_bits, bool& fits) {
  const float m = __uint_as_float(absmax_bits);
  constexpr float kTop = 0x1p-7f - 0x1p-18f;
  fits = m <= 0x1p119f;
  if (!(m > 0.0f) || !fits) return 1.0f;
  const float t = fminf(fmaxf(kTop / m, 0x1p-126f), 0x1p126f);
  return __uint_as_float(__float_as_uint(t) & 0x7f800000u);  // 2^floor(log2 t)
}

// One thread per labelled record: the coefficient record of k_pack_coef at
// the labelled scale, with the 2-bit class code of each corner (lab & 3 in
// the volume, 0 outside) in its slot's free bits. Flags a negative value (a
// negative c0, each voxel being the c0 of one record), a class outside
// [-1, n_classes) or a dropped voxel (-1) with four classes, and a maximum
// the scale cannot bring below 2^-7.
template <int DS, int DR>
__global__ void k_pack_lcoef(const typename Vox<DS>::T* __restrict__ vol, const int8_t* __restrict__ lab,
                             int n_classes, uint32_t* __restrict__ hdr, uint4* __restrict__ rec, int D, int H,
                             int W, int64_t n_rec) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  bool fits;
  const float scale = lcoef_scale(hdr[0], fits);
  if (r == 0) {
    hdr[1] = __float_as_uint(scale);
    hdr[2] = __float_as_uint(1.0f / scale);  // exact: a power of two
    hdr[3] = static_cast<uint32_t>(DR);
    hdr[5] = static_cast<uint32_t>(n_classes);
    if (!fits) atomicOr(hdr + 4, static_cast<uint32_t>(NDRR_LCOEF_RANGE));
  }
  if (r >= n_rec) return;
  int cx, cy, cz;
  record_cell(r, H, W, cx, cy, cz);
  float c[8];
  cell_coefficients<DS>(vol, D, H, W, cx, cy, cz, c);
  uint32_t bad = c[0] < 0.0f ? static_cast<uint32_t>(NDRR_LCOEF_NEGATIVE) : 0u;
  uint32_t h[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    h[k] = Round<NDRR_F16>::bits(c[k] * scale);  // exact product: a power of two
    const int x = cx - 1 + (k & 1), y = cy - 1 + ((k >> 1) & 1), z = cz - 1 + (k >> 2);
    const bool inb = static_cast<unsigned>(x) < static_cast<unsigned>(W) &&
                     static_cast<unsigned>(y) < static_cast<unsigned>(H) &&
                     static_cast<unsigned>(z) < static_cast<unsigned>(D);
    const int l = inb ? lab[(static_cast<int64_t>(z) * H + y) * W + x] : 0;
    if (l < -1 || l >= n_classes || (l == -1 && n_classes == 4)) bad |= NDRR_LCOEF_CLASS;
    const uint32_t code = static_cast<uint32_t>(l) & 3u;
    if (k == 0 && DR == NDRR_BF16) {
      const uint32_t b = Round<NDRR_BF16>::bits(fabsf(c[0]) * scale);  // bits 14-15 zero below 2
      h[0] = b | (code << 14);
    } else {
      if (h[k] & 0x6000u) bad |= NDRR_LCOEF_RANGE;
      h[k] = (h[k] & 0x9fffu) | (code << 13);
    }
  }
  if (bad) atomicOr(hdr + 4, bad);
  rec[r] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16));
}

// One thread per quad coefficient record (x, y0, z0) on the quad grid: the
// quad's four zero-padded corners (as fp32) become a, by, bz, byz in fp32 in
// that order, times the header's scale s (the coefficient layout's, from
// max|v|: |coefficient| <= 4 max|v|), each rounded once to fp16.
template <int DS>
__global__ void k_pack_qcoef(const typename Vox<DS>::T* __restrict__ vol, uint32_t* __restrict__ hdr,
                             uint2* __restrict__ rec, int D, int H, int W, int64_t n_rec) {
#pragma clang fp contract(off)
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const float scale = coef_scale<NDRR_F16>(hdr[0]);
  if (r == 0) {
    hdr[1] = __float_as_uint(scale);
    hdr[2] = __float_as_uint(1.0f / scale);  // exact: a power of two
    hdr[3] = static_cast<uint32_t>(NDRR_F16);
  }
  if (r >= n_rec) return;
  const int cx = static_cast<int>(r % (W + 2));
  const int64_t rest = r / (W + 2);
  const int cy = static_cast<int>(rest % (H + 1));
  const int cz = static_cast<int>(rest / (H + 1));
  float v[4];  // v[2a + b] = (z0 + a, y0 + b) at x
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int x = cx - 1, y = cy - 1 + (k & 1), z = cz - 1 + (k >> 1);
    const bool inb = static_cast<unsigned>(x) < static_cast<unsigned>(W) &&
                     static_cast<unsigned>(y) < static_cast<unsigned>(H) &&
                     static_cast<unsigned>(z) < static_cast<unsigned>(D);
    v[k] = inb ? Vox<DS>::ld(vol, static_cast<int>((static_cast<int64_t>(z) * H + y) * W + x)) : 0.0f;
  }
  const float by = v[1] - v[0];
  const float c[4] = {v[0], by, v[2] - v[0], (v[3] - v[2]) - by};
  uint32_t h[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) h[k] = Round<NDRR_F16>::bits(c[k] * scale);  // exact product: a power of two
  rec[r] = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
}

// One thread per brick row (4 voxels along x): gvol[z][y][x..x+3] += the brick buffer's values
// (gvol_voxel's layout), and the row is cleared, so the buffer is zero again for the next call.
__global__ void k_gvol_unbrick(float* __restrict__ bricks, float* __restrict__ gvol, int D, int H, int W, int nbx,
                               int nby, int64_t n_rows) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n_rows) return;
  const int x4 = static_cast<int>(r % nbx);
  const int64_t rest = r / nbx;
  const int y = static_cast<int>(rest % H);
  const int z = static_cast<int>(rest / H);
  const int64_t brick = (static_cast<int64_t>(z >> 2) * nby + (y >> 2)) * nbx + x4;
  float4* row = reinterpret_cast<float4*>(bricks + brick * 64 + (z & 3) * 16 + (y & 3) * 4);
  const float4 v = *row;
  const float vs[4] = {v.x, v.y, v.z, v.w};
  float* g = gvol + (static_cast<int64_t>(z) * H + y) * W + 4 * x4;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (4 * x4 + k < W) g[k] += vs[k];
  *row = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// One thread per label-quad record (x, y0, z0): bytes lab[z0+a][y0+b][x] at 2a + b.
__global__ void k_pack_label_quads(const int8_t* __restrict__ lab, uint32_t* __restrict__ out,
                                   int D, int H, int W, int64_t n_rec) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n_rec) return;
  const int cx = static_cast<int>(r % (W + 2));
  const int64_t rest = r / (W + 2);
  const int cy = static_cast<int>(rest % (H + 1));
  const int cz = static_cast<int>(rest / (H + 1));
  uint32_t w = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int x = cx - 1, y = cy - 1 + (c & 1), z = cz - 1 + (c >> 1);
    const bool inb = static_cast<unsigned>(x) < static_cast<unsigned>(W) &&
                     static_cast<unsigned>(y) < static_cast<unsigned>(H) &&
                     static_cast<unsigned>(z) < static_cast<unsigned>(D);
    const uint32_t v = inb ? static_cast<uint8_t>(lab[(static_cast<int64_t>(z) * H + y) * W + x]) : 0u;
    w |= v << (8 * c);
  }
  out[r] = w;
}

// One thread per entry of the padded class map (stencil and coefficient layouts): int8
// [D + 4][H + 4][W + 4], lab[z][y][x] at (z + 2, y + 2, x + 2), zero elsewhere.
__global__ void k_pack_label_padded(const int8_t* __restrict__ lab, int8_t* __restrict__ out, int D, int H, int W,
                                    int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = static_cast<int>(i % (W + 4)) - 2;
  const int64_t rest = i / (W + 4);
  const int y = static_cast<int>(rest % (H + 4)) - 2;
  const int z = static_cast<int>(rest / (H + 4)) - 2;
  const bool inb = static_cast<unsigned>(x) < static_cast<unsigned>(W) &&
                   static_cast<unsigned>(y) < static_cast<unsigned>(H) &&
                   static_cast<unsigned>(z) < static_cast<unsigned>(D);
  out[i] = inb ? lab[(static_cast<int64_t>(z) * H + y) * W + x] : static_cast<int8_t>(0);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
int nblocks_of(int N, int width, int& tiles_w, int& IH) {
  if (width > 0) {
    IH = N / width;
    tiles_w = (width + kTileW - 1) / kTileW;
    return tiles_w * ((IH + kTileH - 1) / kTileH);
  }
  tiles_w = 0;
  IH = 0;
  return (N + kTileRays - 1) / kTileRays;
}

int64_t stencil_records(int D, int H, int W) {
  return static_cast<int64_t>(D + 1) * (H + 1) * (W + 1);
}

int64_t quad_records(int D, int H, int W) {
  return static_cast<int64_t>(D + 1) * (H + 1) * (W + 2);
}

int64_t packed_volume_bytes(int D, int H, int W, int dt, int layout) {
  const int64_t e = dt == NDRR_F32 ? 4 : 2;
  switch (layout) {
    case NDRR_LAYOUT_LINEAR: return static_cast<int64_t>(D) * H * W * e;
    case NDRR_LAYOUT_STENCIL: return stencil_records(D, H, W) * 8 * e;
    case NDRR_LAYOUT_QUAD: return quad_records(D, H, W) * 4 * e;
    case NDRR_LAYOUT_COEF: return kCoefHeader + stencil_records(D, H, W) * 8 * e;
    case NDRR_LAYOUT_LCOEF: return dt == NDRR_F32 ? 0 : kCoefHeader + stencil_records(D, H, W) * 16;
    case NDRR_LAYOUT_QCOEF: return kCoefHeader + quad_records(D, H, W) * 8;  // fp16 records whatever the source
    default: return 0;
  }
}

int64_t gvol_brick_floats(int D, int H, int W) {
  return static_cast<int64_t>((D + 3) / 4) * ((H + 3) / 4) * ((W + 3) / 4) * 64;
}

int64_t packed_label_bytes(int D, int H, int W, int layout) {
  if (layout == NDRR_LAYOUT_LCOEF) return 0;  // the records carry the classes
  if (is_quad(layout)) return quad_records(D, H, W) * 4;
  if (layout == NDRR_LAYOUT_STENCIL || layout == NDRR_LAYOUT_COEF)
    return static_cast<int64_t>(D + 4) * (H + 4) * (W + 4);
  return static_cast<int64_t>(D) * H * W;
}

bool valid_layout(int layout) {
  return layout == NDRR_LAYOUT_LINEAR || layout == NDRR_LAYOUT_STENCIL || layout == NDRR_LAYOUT_QUAD ||
         layout == NDRR_LAYOUT_COEF || layout == NDRR_LAYOUT_LCOEF || layout == NDRR_LAYOUT_QCOEF;
}

// The packed layouts' record index is formed with 24-bit multiplies (record_of), and the
// padded class map's offset with signed ones (padded_label).
bool packed_index_fits(int D, int H, int W) {
  return static_cast<int64_t>(D + 1) * (H + 1) < (int64_t{1} << 24) && W + 2 < (1 << 24) &&
         static_cast<int64_t>(H + 4) * (W + 4) < (int64_t{1} << 23);
}

bool valid_dtype(int dt) { return dt == NDRR_F32 || dt == NDRR_BF16 || dt == NDRR_F16; }

int check_common(const void* vol, int vol_dtype, int vol_layout, const int8_t* lab, const float* M,
                 const float* src, const float* tgt, const float* step, int B, int N, int D, int H, int W,
                 int S, int C, int width) {
  if (!vol || !M || !src || !tgt || !step) return NDRR_EINVAL;
  if (B <= 0 || N <= 0 || D <= 0 || H <= 0 || W <= 0 || S < 2 || width < 0) return NDRR_EINVAL;
  if (width > 0 && N % width != 0) return NDRR_EINVAL;
  if (!valid_layout(vol_layout)) return NDRR_EINVAL;
  if (!valid_dtype(vol_dtype)) return NDRR_EDTYPE;
  if (C < 1 || C > NDRR_MAX_CLASSES) return NDRR_ECLASSES;
  if ((vol_layout == NDRR_LAYOUT_LCOEF || vol_layout == NDRR_LAYOUT_QCOEF) && vol_dtype == NDRR_F32) return NDRR_EDTYPE;
  if (vol_layout == NDRR_LAYOUT_LCOEF && C > 4) return NDRR_ECLASSES;
  if (C > 1 && !lab && vol_layout != NDRR_LAYOUT_LCOEF) return NDRR_EINVAL;
  if (B > 65535) return NDRR_ERANGE;
  if (vol_layout != NDRR_LAYOUT_LINEAR && (packed_volume_bytes(D, H, W, vol_dtype, vol_layout) >= (int64_t{1} << 31) ||
                                           packed_label_bytes(D, H, W, vol_layout) >= (int64_t{1} << 31) ||
                                           !packed_index_fits(D, H, W)))
    return NDRR_ERANGE;
  const int64_t vox = static_cast<int64_t>(D) * H * W;
  if (vox >= (int64_t{1} << 31) || stencil_records(D, H, W) >= (int64_t{1} << 31) ||
      quad_records(D, H, W) >= (int64_t{1} << 31) ||
      static_cast<int64_t>(C) * B * N >= (int64_t{1} << 31) || 3 * static_cast<int64_t>(B) * N >= (int64_t{1} << 31))
    return NDRR_ERANGE;
  return NDRR_OK;
}

Params make_params(const void* vol, const int8_t* lab, const float* M, const float* src, const float* tgt,
                   const float* step, int B, int N, int D, int H, int W, int S, int C, int width,
                   int64_t src_bs, int64_t src_rs, int64_t tgt_bs, int64_t step_bs) {
  Params p{};
  p.vol = vol; p.lab = lab; p.M = M; p.src = src; p.tgt = tgt; p.step = step;
  p.src_bs = src_bs; p.src_rs = src_rs; p.tgt_bs = tgt_bs; p.step_bs = step_bs;
  p.B = B; p.N = N; p.D = D; p.H = H; p.W = W; p.S = S; p.C = C;
  p.IW = width;
  p.nblocks = nblocks_of(N, width, p.tiles_w, p.IH);
  p.inv_s = 1.0f / static_cast<float>(S - 1);
  p.s_m1 = static_cast<float>(S - 1);
  p.inv_lo = static_cast<float>(1.0 / static_cast<double>(S - 1) - static_cast<double>(p.inv_s));
  return p;
}

int launch_status() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? NDRR_OK : static_cast<int>(e);
}

template <int DT, int L, int CV>
void launch_fwd3(const Params& p, dim3 grid, hipStream_t st) {
  constexpr int nch = fwd_chunks(L, CV);
  const dim3 block(kTileRays * nch);
  if (p.jac) {
    if (p.S >= kKahanMinSamples)
      hipLaunchKernelGGL((k_fwd<DT, L, CV, true, nch, true>), grid, block, 0, st, p);
    else
      hipLaunchKernelGGL((k_fwd<DT, L, CV, false, nch, true>), grid, block, 0, st, p);
    return;
  }
  if (p.S >= kKahanMinSamples)
    hipLaunchKernelGGL((k_fwd<DT, L, CV, true, nch, false>), grid, block, 0, st, p);
  else
    hipLaunchKernelGGL((k_fwd<DT, L, CV, false, nch, false>), grid, block, 0, st, p);
}

template <int DT, int L>
void launch_fwd2(const Params& p, dim3 grid, hipStream_t st) {
  if (p.C == 1) launch_fwd3<DT, L, 1>(p, grid, st);
  else if (p.C == 2) launch_fwd3<DT, L, 2>(p, grid, st);
  else if (p.C <= 4) launch_fwd3<DT, L, 4>(p, grid, st);
  else if constexpr (L != NDRR_LAYOUT_LCOEF) launch_fwd3<DT, L, 8>(p, grid, st);  // labelled records: C <= 4
}

template <int DT, int L, int CV>
void launch_bwd3(const Params& p, bool gcam, bool gvol, dim3 grid, hipStream_t st) {
  constexpr int nch = bwd_chunks(L);
  const dim3 block(kTileRays * nch);
  if (gcam && gvol) hipLaunchKernelGGL((k_bwd<DT, L, CV, true, true, nch>), grid, block, 0, st, p);
  else if (gcam) hipLaunchKernelGGL((k_bwd<DT, L, CV, true, false, nch>), grid, block, 0, st, p);
  else if (gvol) hipLaunchKernelGGL((k_bwd<DT, L, CV, false, true, nch>), grid, block, 0, st, p);
  else hipLaunchKernelGGL((k_bwd<DT, L, CV, false, false, nch>), grid, block, 0, st, p);
}

template <int DT, int L>
void launch_bwd2(const Params& p, bool gcam, bool gvol, dim3 grid, hipStream_t st) {
  if (p.C == 1) launch_bwd3<DT, L, 1>(p, gcam, gvol, grid, st);
  else if (p.C == 2) launch_bwd3<DT, L, 2>(p, gcam, gvol, grid, st);
  else if (p.C <= 4) launch_bwd3<DT, L, 4>(p, gcam, gvol, grid, st);
  else if constexpr (L != NDRR_LAYOUT_LCOEF) launch_bwd3<DT, L, 8>(p, gcam, gvol, grid, st);
}

template <int DT>
void launch_fwd1(const Params& p, int layout, dim3 grid, hipStream_t st) {
  if (layout == NDRR_LAYOUT_STENCIL) launch_fwd2<DT, NDRR_LAYOUT_STENCIL>(p, grid, st);
  else if (layout == NDRR_LAYOUT_QUAD) launch_fwd2<DT, NDRR_LAYOUT_QUAD>(p, grid, st);
  else if (layout == NDRR_LAYOUT_COEF) launch_fwd2<DT, NDRR_LAYOUT_COEF>(p, grid, st);
  else if (layout == NDRR_LAYOUT_LCOEF) {
    if constexpr (DT != NDRR_F32) launch_fwd2<DT, NDRR_LAYOUT_LCOEF>(p, grid, st);  // 16-bit volumes only
  } else if (layout == NDRR_LAYOUT_QCOEF) {
    if constexpr (DT != NDRR_F32) launch_fwd2<DT, NDRR_LAYOUT_QCOEF>(p, grid, st);
  } else launch_fwd2<DT, NDRR_LAYOUT_LINEAR>(p, grid, st);
}

template <int DT>
void launch_bwd1(const Params& p, int layout, bool gcam, bool gvol, dim3 grid, hipStream_t st) {
  if (layout == NDRR_LAYOUT_STENCIL) launch_bwd2<DT, NDRR_LAYOUT_STENCIL>(p, gcam, gvol, grid, st);
  else if (layout == NDRR_LAYOUT_QUAD) launch_bwd2<DT, NDRR_LAYOUT_QUAD>(p, gcam, gvol, grid, st);
  else if (layout == NDRR_LAYOUT_COEF) launch_bwd2<DT, NDRR_LAYOUT_COEF>(p, gcam, gvol, grid, st);
  else if (layout == NDRR_LAYOUT_LCOEF) {
    if constexpr (DT != NDRR_F32) launch_bwd2<DT, NDRR_LAYOUT_LCOEF>(p, gcam, gvol, grid, st);
  } else if (layout == NDRR_LAYOUT_QCOEF) {
    if constexpr (DT != NDRR_F32) launch_bwd2<DT, NDRR_LAYOUT_QCOEF>(p, gcam, gvol, grid, st);
  } else launch_bwd2<DT, NDRR_LAYOUT_LINEAR>(p, gcam, gvol, grid, st);
}

template <int DS>
int pack_qcoef_from(const void* vol, int D, int H, int W, void* out, int64_t n, hipStream_t st) {
  using T = typename Vox<DS>::T;
  const T* v = static_cast<const T*>(vol);
  uint32_t* hdr = static_cast<uint32_t*>(out);
  uint2* rec = reinterpret_cast<uint2*>(static_cast<char*>(out) + kCoefHeader);
  const int64_t n_vox = static_cast<int64_t>(D) * H * W;
  hipLaunchKernelGGL(k_coef_header_init, dim3(1), dim3(kCoefHeader / 4), 0, st, hdr);
  const int64_t ablocks = std::min<int64_t>((n_vox + kThreads - 1) / kThreads, 4096);
  hipLaunchKernelGGL(k_coef_absmax<DS>, dim3(static_cast<unsigned>(ablocks)), dim3(kThreads), 0, st, v, n_vox, hdr);
  hipLaunchKernelGGL(k_pack_qcoef<DS>, dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                     st, v, hdr, rec, D, H, W, n);
  return launch_status();
}

template <int DS>
int pack_lcoef_from(const void* vol, const int8_t* lab, int n_classes, int record_dtype, int D, int H, int W,
                    void* out, int64_t n, hipStream_t st) {
  using T = typename Vox<DS>::T;
  const T* v = static_cast<const T*>(vol);
  uint32_t* hdr = static_cast<uint32_t*>(out);
  uint4* rec = reinterpret_cast<uint4*>(static_cast<char*>(out) + kCoefHeader);
  hipLaunchKernelGGL(k_coef_header_init, dim3(1), dim3(kCoefHeader / 4), 0, st, hdr);
  const int64_t ablocks = std::min<int64_t>((n + kThreads - 1) / kThreads, 4096);
  hipLaunchKernelGGL(k_lcoef_absmax<DS>, dim3(static_cast<unsigned>(ablocks)), dim3(kThreads), 0, st, v, D, H, W,
                     n, hdr);
  const dim3 grid(static_cast<unsigned>((n + kThreads - 1) / kThreads));
  if (record_dtype == NDRR_BF16)
    hipLaunchKernelGGL((k_pack_lcoef<DS, NDRR_BF16>), grid, dim3(kThreads), 0, st, v, lab, n_classes, hdr, rec, D, H,
                       W, n);
  else
    hipLaunchKernelGGL((k_pack_lcoef<DS, NDRR_F16>), grid, dim3(kThreads), 0, st, v, lab, n_classes, hdr, rec, D, H,
                       W, n);
  return launch_status();
}

template <int DS>
int pack_coef_from(const void* vol, int record_dtype, int D, int H, int W, void* out, int64_t n, hipStream_t st) {
  using T = typename Vox<DS>::T;
  const T* v = static_cast<const T*>(vol);
  uint32_t* hdr = static_cast<uint32_t*>(out);
  void* rec = static_cast<char*>(out) + kCoefHeader;
  const int64_t n_vox = static_cast<int64_t>(D) * H * W;
  hipLaunchKernelGGL(k_coef_header_init, dim3(1), dim3(kCoefHeader / 4), 0, st, hdr);
  const int64_t ablocks = std::min<int64_t>((n_vox + kThreads - 1) / kThreads, 4096);
  hipLaunchKernelGGL(k_coef_absmax<DS>, dim3(static_cast<unsigned>(ablocks)), dim3(kThreads), 0, st, v, n_vox, hdr);
  const dim3 grid(static_cast<unsigned>((n + kThreads - 1) / kThreads));
  switch (record_dtype) {
    case NDRR_F32: hipLaunchKernelGGL((k_pack_coef<DS, NDRR_F32>), grid, dim3(kThreads), 0, st, v, hdr, rec, D, H, W, n); break;
    case NDRR_BF16: hipLaunchKernelGGL((k_pack_coef<DS, NDRR_BF16>), grid, dim3(kThreads), 0, st, v, hdr, rec, D, H, W, n); break;
    default: hipLaunchKernelGGL((k_pack_coef<DS, NDRR_F16>), grid, dim3(kThreads), 0, st, v, hdr, rec, D, H, W, n); break;
  }
  return launch_status();
}

}  // namespace

extern "C" {

int ndrr_version(void) { return 200; }

const char* ndrr_strerror(int code) {
  switch (code) {
    case NDRR_OK: return "success";
    case NDRR_EINVAL: return "invalid argument (null pointer, non-positive size, n_samples < 2, bad width or layout)";
    case NDRR_EDTYPE: return "unsupported volume element type";
    case NDRR_ECLASSES: return "n_classes outside [1, 8]";
    case NDRR_ERANGE: return "a size exceeds the kernel's 32-bit index range";
    case NDRR_ESCRATCH: return "the work needs more than the allocation-free entry point's fixed scratch array holds";
    default: break;
  }
  if (code > 0) return hipGetErrorString(static_cast<hipError_t>(code));
  return "unknown error";
}

int ndrr_raymarch_fwd(const void* vol, int vol_dtype, int vol_layout, const int8_t* lab, const float* M,
                      const float* src, const float* tgt, const float* step, float* out, int B, int N, int D,
                      int H, int W, int n_samples, int n_classes, int width, int64_t src_batch_stride,
                      int64_t src_ray_stride, int64_t tgt_batch_stride, int64_t step_batch_stride,
                      void* stream) {
  return ndrr_raymarch_fwd_jac(vol, vol_dtype, vol_layout, lab, M, src, tgt, step, nullptr, out, nullptr, 0, B, N, D,
                               H, W,
                               n_samples, n_classes, width, src_batch_stride, src_ray_stride, tgt_batch_stride,
                               step_batch_stride, stream);
}

// The forward with M given, or (w2v, rt) given and M composed per workgroup.
static int raymarch_fwd_any(const void* vol, int vol_dtype, int vol_layout, const int8_t* lab, const float* M,
                            const double* w2v, const float* rt, const float* src, const float* tgt,
                            const float* step, const float* lattice, float* out, float* jac, int jac_planes, int B,
                            int N, int D, int H,
                            int W, int n_samples, int n_classes, int width, int64_t src_batch_stride,
                            int64_t src_ray_stride, int64_t tgt_batch_stride, int64_t step_batch_stride,
                            void* stream) {
  if (w2v != nullptr && (rt == nullptr || M != nullptr)) return NDRR_EINVAL;
  const int rc = check_common(vol, vol_dtype, vol_layout, lab, w2v ? rt : M, src, tgt, step, B, N, D, H, W,
                              n_samples, n_classes, width);
  if (rc != NDRR_OK) return rc;
  if (!out) return NDRR_EINVAL;
  if (jac && jac_planes != 6 && jac_planes != 7) return NDRR_EINVAL;
  if (jac && static_cast<int64_t>(jac_planes) * n_classes * B * N >= (int64_t{1} << 31)) return NDRR_ERANGE;
  Params p = make_params(vol, lab, M, src, tgt, step, B, N, D, H, W, n_samples, n_classes, width,
                         src_batch_stride, src_ray_stride, tgt_batch_stride, step_batch_stride);
  p.vol_bytes = static_cast<uint32_t>(packed_volume_bytes(D, H, W, vol_dtype, vol_layout));
  p.vol_skip = has_header(vol_layout) ? kCoefHeader : 0u;
  p.lab_bytes = n_classes > 1 ? static_cast<uint32_t>(packed_label_bytes(D, H, W, vol_layout)) : 0u;
  p.w2v = w2v;
  p.rt = rt;
  p.lattice = lattice;
  p.out = out;
  p.jac = jac;
  p.jac_q = jac ? jac_planes : 0;
  const dim3 grid(p.nblocks, B);
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (vol_dtype) {
    case NDRR_F32: launch_fwd1<NDRR_F32>(p, vol_layout, grid, st); break;
    case NDRR_BF16: launch_fwd1<NDRR_BF16>(p, vol_layout, grid, st); break;
    default: launch_fwd1<NDRR_F16>(p, vol_layout, grid, st); break;
  }
  return launch_status();
}

int ndrr_raymarch_fwd_jac(const void* vol, int vol_dtype, int vol_layout, const int8_t* lab, const float* M,
                          const float* src, const float* tgt, const float* step, const float* lattice_offset,
                          float* out, float* jac, int jac_planes, int B, int N, int D, int H, int W, int n_samples,
                          int n_classes, int width, int64_t src_batch_stride, int64_t src_ray_stride,
                          int64_t tgt_batch_stride, int64_t step_batch_stride, void* stream) {
  return raymarch_fwd_any(vol, vol_dtype, vol_layout, lab, M, nullptr, nullptr, src, tgt, step, lattice_offset, out,
                          jac, jac_planes, B, N, D, H, W, n_samples, n_classes, width, src_batch_stride,
                          src_ray_stride, tgt_batch_stride, step_batch_stride, stream);
}

int ndrr_raymarch_fwd_jac_pose(const void* vol, int vol_dtype, int vol_layout, const int8_t* lab,
                               const double* world_to_voxel, const float* rt_inv, const float* src,
                               const float* tgt, const float* step, const float* lattice_offset, float* out,
                               float* jac, int jac_planes, int B, int N, int D, int H, int W, int n_samples,
                               int n_classes, int width, int64_t src_batch_stride, int64_t src_ray_stride,
                               int64_t tgt_batch_stride, int64_t step_batch_stride, void* stream) {
  if (world_to_voxel == nullptr) return NDRR_EINVAL;
  return raymarch_fwd_any(vol, vol_dtype, vol_layout, lab, nullptr, world_to_voxel, rt_inv, src, tgt, step,
                          lattice_offset, out, jac, jac_planes, B, N, D, H, W, n_samples, n_classes, width,
                          src_batch_stride, src_ray_stride, tgt_batch_stride, step_batch_stride, stream);
}

void ndrr_launch_tile(int* tile_width, int* tile_height, int* tile_rays) {
  if (tile_width) *tile_width = kTileW;
  if (tile_height) *tile_height = kTileH;
  if (tile_rays) *tile_rays = kTileRays;
}

size_t ndrr_bwd_workspace_floats(int B, int N, int width) {
  if (B <= 0 || N <= 0 || width < 0) return 0;
  int tw, ih;
  return static_cast<size_t>(B) * nblocks_of(N, width, tw, ih) * kBwdPartsMax * 12;
}

// The two-pass adjoint, with the merged volume-gradient scatter adding into a brick buffer when
// one is given (then added into gvol and cleared by k_gvol_unbrick).
static int raymarch_bwd_any(const float* go, const void* vol, int vol_dtype, int vol_layout, const int8_t* lab,
                            const float* M, const float* src, const float* tgt, const float* step, float* workspace,
                            float* gM, float* gsrc, float* gtgt, float* gstep, float* gvol, float* gvol_bricks,
                            int need_gcam, int B, int N, int D, int H, int W, int n_samples, int n_classes, int width,
                            int64_t src_batch_stride, int64_t src_ray_stride, int64_t tgt_batch_stride,
                            int64_t step_batch_stride, void* stream) {
  const int rc = check_common(vol, vol_dtype, vol_layout, lab, M, src, tgt, step, B, N, D, H, W, n_samples,
                              n_classes, width);
  if (rc != NDRR_OK) return rc;
  if (!go || !workspace || !gM || reinterpret_cast<uintptr_t>(workspace) % 16 != 0) return NDRR_EINVAL;
  if (need_gcam && (!gsrc || !gtgt || !gstep)) return NDRR_EINVAL;
  if (gvol_bricks != nullptr && (gvol == nullptr || reinterpret_cast<uintptr_t>(gvol_bricks) % 16 != 0))
    return NDRR_EINVAL;
  if (gvol_bricks != nullptr && gvol_brick_floats(D, H, W) >= (int64_t{1} << 31)) return NDRR_ERANGE;
  Params p = make_params(vol, lab, M, src, tgt, step, B, N, D, H, W, n_samples, n_classes, width,
                         src_batch_stride, src_ray_stride, tgt_batch_stride, step_batch_stride);
  p.vol_bytes = static_cast<uint32_t>(packed_volume_bytes(D, H, W, vol_dtype, vol_layout));
  p.vol_skip = has_header(vol_layout) ? kCoefHeader : 0u;
  p.lab_bytes = n_classes > 1 ? static_cast<uint32_t>(packed_label_bytes(D, H, W, vol_layout)) : 0u;
  p.go = go;
  p.gpart = workspace;
  p.gsrc = gsrc; p.gtgt = gtgt; p.gstep = gstep; p.gvol = gvol;
  p.gvol_merge = static_cast<uint64_t>(D + 2) * (H + 2) * (W + 2) + 64 <= (uint64_t{1} << 32);
  // the brick buffer serves the merged scatter only (the per-lane one adds into gvol directly)
  p.gvol_bricks = p.gvol_merge ? gvol_bricks : nullptr;
  p.nbx = (W + 3) / 4;
  p.nby = (H + 3) / 4;
  const dim3 grid(p.nblocks, B);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool gcam = need_gcam != 0, gv = gvol != nullptr;
  switch (vol_dtype) {
    case NDRR_F32: launch_bwd1<NDRR_F32>(p, vol_layout, gcam, gv, grid, st); break;
    case NDRR_BF16: launch_bwd1<NDRR_BF16>(p, vol_layout, gcam, gv, grid, st); break;
    default: launch_bwd1<NDRR_F16>(p, vol_layout, gcam, gv, grid, st); break;
  }
  int r = launch_status();
  if (r != NDRR_OK) return r;
  const int parts_per_tile = vol_layout == NDRR_LAYOUT_STENCIL ? bwd_parts_per_tile(NDRR_LAYOUT_STENCIL)
                             : is_quad(vol_layout)             ? bwd_parts_per_tile(NDRR_LAYOUT_QUAD)
                             : is_coef(vol_layout)             ? bwd_parts_per_tile(NDRR_LAYOUT_COEF)
                                                               : bwd_parts_per_tile(NDRR_LAYOUT_LINEAR);
  hipLaunchKernelGGL(k_reduce_pose, dim3(B), dim3(kReduceThreads), 0, st, workspace, gM,
                     p.nblocks * parts_per_tile, nullptr, nullptr, Se3Out{});
  if (p.gvol_bricks != nullptr) {
    const int64_t rows = static_cast<int64_t>(D) * H * p.nbx;
    hipLaunchKernelGGL(k_gvol_unbrick, dim3(static_cast<unsigned>((rows + kThreads - 1) / kThreads)), dim3(kThreads),
                       0, st, p.gvol_bricks, gvol, D, H, W, p.nbx, p.nby, rows);
  }
  return launch_status();
}

int ndrr_raymarch_bwd(const float* go, const void* vol, int vol_dtype, int vol_layout, const int8_t* lab,
                      const float* M, const float* src, const float* tgt, const float* step, float* workspace,
                      float* gM, float* gsrc, float* gtgt, float* gstep, float* gvol, int need_gcam, int B, int N,
                      int D, int H, int W, int n_samples, int n_classes, int width, int64_t src_batch_stride,
                      int64_t src_ray_stride, int64_t tgt_batch_stride, int64_t step_batch_stride,
                      void* stream) {
  return raymarch_bwd_any(go, vol, vol_dtype, vol_layout, lab, M, src, tgt, step, workspace, gM, gsrc, gtgt, gstep,
                          gvol, nullptr, need_gcam, B, N, D, H, W, n_samples, n_classes, width, src_batch_stride,
                          src_ray_stride, tgt_batch_stride, step_batch_stride, stream);
}

size_t ndrr_gvol_brick_floats(int D, int H, int W) {
  if (D <= 0 || H <= 0 || W <= 0) return 0;
  return static_cast<size_t>(gvol_brick_floats(D, H, W));
}

int ndrr_raymarch_bwd_bricked(const float* go, const void* vol, int vol_dtype, int vol_layout, const int8_t* lab,
                              const float* M, const float* src, const float* tgt, const float* step,
                              float* workspace, float* gM, float* gsrc, float* gtgt, float* gstep, float* gvol,
                              float* gvol_bricks, int need_gcam, int B, int N, int D, int H, int W, int n_samples,
                              int n_classes, int width, int64_t src_batch_stride, int64_t src_ray_stride,
                              int64_t tgt_batch_stride, int64_t step_batch_stride, void* stream) {
  if (!gvol || !gvol_bricks) return NDRR_EINVAL;
  return raymarch_bwd_any(go, vol, vol_dtype, vol_layout, lab, M, src, tgt, step, workspace, gM, gsrc, gtgt, gstep,
                          gvol, gvol_bricks, need_gcam, B, N, D, H, W, n_samples, n_classes, width, src_batch_stride,
                          src_ray_stride, tgt_batch_stride, step_batch_stride, stream);
}

// The contraction with M (and gM out) given, or (w2v, rt) given and the rt gradient out.
static int raymarch_bwd_jac_any(const float* go, int64_t go_batch_stride, const float* jac, int jac_planes,
                                const float* M, const double* w2v, const float* rt, const float* src,
                                const float* tgt, const float* step, float* workspace, float* gM, float* g_rt,
                                float* gsrc, float* gtgt, float* gstep, int need_gcam, int B, int N, int n_classes,
                                int width, int64_t src_batch_stride, int64_t src_ray_stride,
                                int64_t tgt_batch_stride, int64_t step_batch_stride, void* stream,
                                const Se3Out& se3 = Se3Out{}) {
  if (w2v != nullptr ? (!rt || !g_rt || M || gM) : (!M || (!gM && !se3.grot))) return NDRR_EINVAL;
  if (!go || !jac || !src || !tgt || !step || !workspace) return NDRR_EINVAL;
  if (reinterpret_cast<uintptr_t>(workspace) % 16 != 0) return NDRR_EINVAL;
  if (B <= 0 || N <= 0 || width < 0 || (width > 0 && N % width != 0)) return NDRR_EINVAL;
  if (jac_planes != 6 && jac_planes != 7) return NDRR_EINVAL;
  if (need_gcam && (jac_planes != 7 || !gsrc || !gtgt || !gstep)) return NDRR_EINVAL;
  if (go_batch_stride != 0 && go_batch_stride != static_cast<int64_t>(n_classes) * N) return NDRR_EINVAL;
  if (n_classes < 1 || n_classes > NDRR_MAX_CLASSES) return NDRR_ECLASSES;
  if (B > 65535 || static_cast<int64_t>(jac_planes) * n_classes * B * N >= (int64_t{1} << 31) ||
      3 * static_cast<int64_t>(B) * N >= (int64_t{1} << 31))
    return NDRR_ERANGE;
  Params p{};
  p.M = M; p.w2v = w2v; p.rt = rt;
  p.src = src; p.tgt = tgt; p.step = step; p.go = go; p.go_bs = go_batch_stride;
  p.jac = const_cast<float*>(jac);
  p.jac_q = jac_planes;
  p.src_bs = src_batch_stride; p.src_rs = src_ray_stride; p.tgt_bs = tgt_batch_stride; p.step_bs = step_batch_stride;
  p.B = B; p.N = N; p.C = n_classes; p.IW = width;
  p.nblocks = nblocks_of(N, width, p.tiles_w, p.IH);
  p.gpart = workspace;
  p.gsrc = gsrc; p.gtgt = gtgt; p.gstep = gstep;
  // ceil(N / (kTileRays kJacRays)) <= nblocks: the workspace of ndrr_bwd_workspace_floats holds the partials
  const int nlin = (N + kTileRays * kJacRays - 1) / (kTileRays * kJacRays);
  const dim3 grid(nlin, B), block(kTileRays);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const bool gcam = need_gcam != 0;
  const int C = n_classes;
  if (C == 1) {
    if (gcam) hipLaunchKernelGGL((k_bwd_jac<1, true>), grid, block, 0, st, p);
    else hipLaunchKernelGGL((k_bwd_jac<1, false>), grid, block, 0, st, p);
  } else if (C == 2) {
    if (gcam) hipLaunchKernelGGL((k_bwd_jac<2, true>), grid, block, 0, st, p);
    else hipLaunchKernelGGL((k_bwd_jac<2, false>), grid, block, 0, st, p);
  } else if (C <= 4) {
    if (gcam) hipLaunchKernelGGL((k_bwd_jac<4, true>), grid, block, 0, st, p);
    else hipLaunchKernelGGL((k_bwd_jac<4, false>), grid, block, 0, st, p);
  } else {
    if (gcam) hipLaunchKernelGGL((k_bwd_jac<8, true>), grid, block, 0, st, p);
    else hipLaunchKernelGGL((k_bwd_jac<8, false>), grid, block, 0, st, p);
  }
  int r = launch_status();
  if (r != NDRR_OK) return r;
  hipLaunchKernelGGL(k_reduce_pose, dim3(B), dim3(kReduceThreads), 0, st, workspace, gM, nlin, w2v, g_rt, se3);
  return launch_status();
}

int ndrr_raymarch_bwd_jac(const float* go, int64_t go_batch_stride, const float* jac, int jac_planes, const float* M,
                          const float* src, const float* tgt, const float* step, float* workspace, float* gM,
                          float* gsrc, float* gtgt, float* gstep, int need_gcam, int B, int N, int n_classes,
                          int width, int64_t src_batch_stride, int64_t src_ray_stride, int64_t tgt_batch_stride,
                          int64_t step_batch_stride, void* stream) {
  return raymarch_bwd_jac_any(go, go_batch_stride, jac, jac_planes, M, nullptr, nullptr, src, tgt, step, workspace,
                              gM, nullptr, gsrc, gtgt, gstep, need_gcam, B, N, n_classes, width, src_batch_stride,
                              src_ray_stride, tgt_batch_stride, step_batch_stride, stream);
}

int ndrr_raymarch_bwd_jac_pose(const float* go, int64_t go_batch_stride, const float* jac, int jac_planes,
                               const double* world_to_voxel, const float* rt_inv, const float* src,
                               const float* tgt, const float* step, float* workspace, float* g_rt_inv, float* gsrc,
                               float* gtgt, float* gstep, int need_gcam, int B, int N, int n_classes, int width,
                               int64_t src_batch_stride, int64_t src_ray_stride, int64_t tgt_batch_stride,
                               int64_t step_batch_stride, void* stream) {
  if (world_to_voxel == nullptr) return NDRR_EINVAL;
  return raymarch_bwd_jac_any(go, go_batch_stride, jac, jac_planes, nullptr, world_to_voxel, rt_inv, src, tgt, step,
                              workspace, nullptr, g_rt_inv, gsrc, gtgt, gstep, need_gcam, B, N, n_classes, width,
                              src_batch_stride, src_ray_stride, tgt_batch_stride, step_batch_stride, stream);
}

int ndrr_raymarch_bwd_jac_se3(const float* go, int64_t go_batch_stride, const float* jac, int jac_planes,
                              const float* M, const float* src, const float* tgt, const float* step, float* workspace,
                              const double* prefix, const double* suffix, const float* rot, const float* xyz,
                              int64_t param_stride, float grad_scale, float* grot, float* gxyz, int B, int N,
                              int n_classes, int width, int64_t src_batch_stride, int64_t src_ray_stride,
                              int64_t tgt_batch_stride, int64_t step_batch_stride, void* stream) {
  if (!prefix || !suffix || !rot || !xyz || !grot || !gxyz || param_stride < 3) return NDRR_EINVAL;
  if (grad_scale != 1.0f && grad_scale != -1.0f) return NDRR_EINVAL;
  const Se3Out se3{prefix, suffix, rot, xyz, param_stride, grad_scale, grot, gxyz};
  return raymarch_bwd_jac_any(go, go_batch_stride, jac, jac_planes, M, nullptr, nullptr, src, tgt, step, workspace,
                              nullptr, nullptr, nullptr, nullptr, nullptr, 0, B, N, n_classes, width,
                              src_batch_stride, src_ray_stride, tgt_batch_stride, step_batch_stride, stream, se3);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free contraction: the pose partials of ndrr_raymarch_bwd_jac_pose
// go to a fixed device array. In host code `scratch` names a host-side
// stand-in, not device memory: its device address is looked up per device
// with hipGetSymbolAddress and only that address is handed to kernels.
// ---------------------------------------------------------------------------
static __device__ __align__(256) float scratch[NDRR_NOALLOC_SCRATCH_FLOATS];
static_assert(kTileRays * kJacRays == NDRR_NOALLOC_RAYS_PER_PART,
              "the contraction writes one pose partial per NDRR_NOALLOC_RAYS_PER_PART rays");

namespace {
constexpr int kMaxDevices = 64;
std::atomic<float*> g_scratch[kMaxDevices];  // device address of `scratch`, per device (null until looked up)
std::mutex g_scratch_mu;

// Device address of `scratch` on the current device, looked up once per device.
int scratch_of_current_device(float** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return static_cast<int>(e);
  if (dev < 0 || dev >= kMaxDevices) return NDRR_ERANGE;
  float* p = g_scratch[dev].load(std::memory_order_acquire);
  if (p == nullptr) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    p = g_scratch[dev].load(std::memory_order_relaxed);
    if (p == nullptr) {
      void* addr = nullptr;
      e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
      if (e != hipSuccess) return static_cast<int>(e);
      if (addr == nullptr) return static_cast<int>(hipErrorInvalidSymbol);
      p = static_cast<float*>(addr);
      g_scratch[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return NDRR_OK;
}
}  // namespace

extern "C" {

int ndrr_raymarch_bwd_jac_pose_noalloc(const float* go, int64_t go_batch_stride, const float* jac, int jac_planes,
                                       const double* world_to_voxel, const float* rt_inv, const float* src,
                                       const float* tgt, const float* step, float* g_rt_inv, float* gsrc,
                                       float* gtgt, float* gstep, int need_gcam, int B, int N, int n_classes,
                                       int width, int64_t src_batch_stride, int64_t src_ray_stride,
                                       int64_t tgt_batch_stride, int64_t step_batch_stride, void* stream) {
  if (world_to_voxel == nullptr || B <= 0 || N <= 0) return NDRR_EINVAL;
  const int64_t parts = static_cast<int64_t>(B) * ((static_cast<int64_t>(N) + NDRR_NOALLOC_RAYS_PER_PART - 1) /
                                                   NDRR_NOALLOC_RAYS_PER_PART);
  if (parts * 12 > static_cast<int64_t>(NDRR_NOALLOC_SCRATCH_FLOATS)) return NDRR_ESCRATCH;
  float* ws = nullptr;
  const int rc = scratch_of_current_device(&ws);
  if (rc != NDRR_OK) return rc;
  return raymarch_bwd_jac_any(go, go_batch_stride, jac, jac_planes, nullptr, world_to_voxel, rt_inv, src, tgt, step,
                              ws, nullptr, g_rt_inv, gsrc, gtgt, gstep, need_gcam, B, N, n_classes, width,
                              src_batch_stride, src_ray_stride, tgt_batch_stride, step_batch_stride, stream);
}

size_t ndrr_packed_bytes(int D, int H, int W, int vol_dtype, int vol_layout) {
  if (D <= 0 || H <= 0 || W <= 0 || !valid_dtype(vol_dtype)) return 0;
  return static_cast<size_t>(packed_volume_bytes(D, H, W, vol_dtype, vol_layout));
}

int ndrr_pack_volume(const void* vol, int vol_dtype, int vol_layout, int D, int H, int W, void* out,
                     void* stream) {
  if (!vol || !out || D <= 0 || H <= 0 || W <= 0) return NDRR_EINVAL;
  if (!valid_dtype(vol_dtype)) return NDRR_EDTYPE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (vol_layout == NDRR_LAYOUT_LINEAR) {
    const hipError_t e = hipMemcpyAsync(out, vol, ndrr_packed_bytes(D, H, W, vol_dtype, vol_layout),
                                        hipMemcpyDeviceToDevice, st);
    return e == hipSuccess ? NDRR_OK : static_cast<int>(e);
  }
  if (vol_layout == NDRR_LAYOUT_COEF) return ndrr_pack_coef(vol, vol_dtype, vol_dtype, D, H, W, out, stream);
  if (vol_layout == NDRR_LAYOUT_QCOEF) {
    const int64_t n = quad_records(D, H, W);
    if (n >= (int64_t{1} << 31) || static_cast<int64_t>(D) * H * W >= (int64_t{1} << 31)) return NDRR_ERANGE;
    switch (vol_dtype) {
      case NDRR_F32: return pack_qcoef_from<NDRR_F32>(vol, D, H, W, out, n, st);
      case NDRR_BF16: return pack_qcoef_from<NDRR_BF16>(vol, D, H, W, out, n, st);
      default: return pack_qcoef_from<NDRR_F16>(vol, D, H, W, out, n, st);
    }
  }
  if (vol_layout != NDRR_LAYOUT_STENCIL && vol_layout != NDRR_LAYOUT_QUAD) return NDRR_EINVAL;
  const bool quad = vol_layout == NDRR_LAYOUT_QUAD;
  const int64_t n = quad ? quad_records(D, H, W) : stencil_records(D, H, W);
  if (n >= (int64_t{1} << 31)) return NDRR_ERANGE;
  const dim3 grid(static_cast<unsigned>((n + kThreads - 1) / kThreads));
  switch (vol_dtype) {
    case NDRR_F32:
      if (quad) hipLaunchKernelGGL(k_pack_quad<NDRR_F32>, grid, dim3(kThreads), 0, st,
                                   static_cast<const float*>(vol), out, D, H, W, n);
      else hipLaunchKernelGGL(k_pack_stencil<NDRR_F32>, grid, dim3(kThreads), 0, st,
                              static_cast<const float*>(vol), out, D, H, W, n);
      break;
    case NDRR_BF16:
      if (quad) hipLaunchKernelGGL(k_pack_quad<NDRR_BF16>, grid, dim3(kThreads), 0, st,
                                   static_cast<const uint16_t*>(vol), out, D, H, W, n);
      else hipLaunchKernelGGL(k_pack_stencil<NDRR_BF16>, grid, dim3(kThreads), 0, st,
                              static_cast<const uint16_t*>(vol), out, D, H, W, n);
      break;
    default:
      if (quad) hipLaunchKernelGGL(k_pack_quad<NDRR_F16>, grid, dim3(kThreads), 0, st,
                                   static_cast<const _Float16*>(vol), out, D, H, W, n);
      else hipLaunchKernelGGL(k_pack_stencil<NDRR_F16>, grid, dim3(kThreads), 0, st,
                              static_cast<const _Float16*>(vol), out, D, H, W, n);
      break;
  }
  return launch_status();
}

int ndrr_pack_coef(const void* vol, int vol_dtype, int record_dtype, int D, int H, int W, void* out, void* stream) {
  if (!vol || !out || D <= 0 || H <= 0 || W <= 0) return NDRR_EINVAL;
  if (!valid_dtype(vol_dtype) || !valid_dtype(record_dtype)) return NDRR_EDTYPE;
  const int64_t n = stencil_records(D, H, W);
  if (n >= (int64_t{1} << 31) || static_cast<int64_t>(D) * H * W >= (int64_t{1} << 31)) return NDRR_ERANGE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (vol_dtype) {
    case NDRR_F32: return pack_coef_from<NDRR_F32>(vol, record_dtype, D, H, W, out, n, st);
    case NDRR_BF16: return pack_coef_from<NDRR_BF16>(vol, record_dtype, D, H, W, out, n, st);
    default: return pack_coef_from<NDRR_F16>(vol, record_dtype, D, H, W, out, n, st);
  }
}

int ndrr_pack_coef_labels(const void* vol, int vol_dtype, const int8_t* lab, int n_classes, int record_dtype, int D,
                          int H, int W, void* out, void* stream) {
  if (!vol || !lab || !out || D <= 0 || H <= 0 || W <= 0) return NDRR_EINVAL;
  if (!valid_dtype(vol_dtype) || (record_dtype != NDRR_BF16 && record_dtype != NDRR_F16)) return NDRR_EDTYPE;
  if (n_classes < 1 || n_classes > 4) return NDRR_ECLASSES;
  const int64_t n = stencil_records(D, H, W);
  if (n >= (int64_t{1} << 31) || static_cast<int64_t>(D) * H * W >= (int64_t{1} << 31)) return NDRR_ERANGE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (vol_dtype) {
    case NDRR_F32: return pack_lcoef_from<NDRR_F32>(vol, lab, n_classes, record_dtype, D, H, W, out, n, st);
    case NDRR_BF16: return pack_lcoef_from<NDRR_BF16>(vol, lab, n_classes, record_dtype, D, H, W, out, n, st);
    default: return pack_lcoef_from<NDRR_F16>(vol, lab, n_classes, record_dtype, D, H, W, out, n, st);
  }
}

size_t ndrr_packed_label_bytes(int D, int H, int W, int vol_layout) {
  if (D <= 0 || H <= 0 || W <= 0) return 0;
  if (!valid_layout(vol_layout)) return 0;
  return static_cast<size_t>(packed_label_bytes(D, H, W, vol_layout));
}

int ndrr_pack_labels(const int8_t* lab, int vol_layout, int D, int H, int W, void* out, void* stream) {
  if (!lab || !out || D <= 0 || H <= 0 || W <= 0) return NDRR_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (vol_layout == NDRR_LAYOUT_LINEAR) {
    const hipError_t e = hipMemcpyAsync(out, lab, static_cast<size_t>(D) * H * W, hipMemcpyDeviceToDevice, st);
    return e == hipSuccess ? NDRR_OK : static_cast<int>(e);
  }
  if (vol_layout == NDRR_LAYOUT_STENCIL || vol_layout == NDRR_LAYOUT_COEF) {
    const int64_t n = packed_label_bytes(D, H, W, vol_layout);
    if (n >= (int64_t{1} << 31) || !packed_index_fits(D, H, W)) return NDRR_ERANGE;
    hipLaunchKernelGGL(k_pack_label_padded, dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, st, lab, static_cast<int8_t*>(out), D, H, W, n);
    return launch_status();
  }
  if (!is_quad(vol_layout)) return NDRR_EINVAL;
  const int64_t n = quad_records(D, H, W);
  if (n >= (int64_t{1} << 31)) return NDRR_ERANGE;
  hipLaunchKernelGGL(k_pack_label_quads, dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, st, lab, static_cast<uint32_t*>(out), D, H, W, n);
  return launch_status();
}

int ndrr_labels_to_i8(const float* src, int8_t* dst, int64_t numel, int n_classes, void* stream) {
  if (!src || !dst || numel < 0) return NDRR_EINVAL;
  if (n_classes < 1 || n_classes > 127) return NDRR_ECLASSES;
  if (numel == 0) return NDRR_OK;
  const int64_t blocks = (numel + kThreads - 1) / kThreads;
  if (blocks > 0x7fffffff) return NDRR_ERANGE;
  hipLaunchKernelGGL(k_labels_to_i8, dim3(static_cast<unsigned>(blocks)), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), src, dst, numel, n_classes);
  return launch_status();
}

}  // extern "C"
