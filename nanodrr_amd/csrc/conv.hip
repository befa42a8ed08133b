// 3x3 convolutions of the pose-proposal CNN on AMD Instinct MI355X (gfx950):
// implicit GEMMs on the bf16 matrix cores (v_mfma_f32_16x16x32_bf16), bf16
// activations and weights in NHWC / [Cout][3][3][Cin], fp32 accumulation.
//
//   forward      y[m][co]  = relu(b[co] + sum_{tap,ci} x[m @ tap][ci] w[co][tap][ci])
//   input grad   dx[m][ci] = [x[m][ci] > 0] sum_{tap,co} dy[m' (tap)][co] w[co][tap][ci]
//   weight grad  dw[co][tap][ci] = sum_m dy[m][co] x[m @ tap][ci]
//
// with padding 1 and stride 1 or 2 (out = (in - 1) / stride + 1). The GEMM
// rows are pixels, the columns channels and K runs over (tap, channel) in
// chunks of 32 channels, so every operand row is 16-byte pieces of eight
// consecutive channels (a tap of a pixel, or a weight row). The forward and
// the input gradient stage each K step's tiles through LDS, double-buffered
// (k_igemm); the first layer's forward loads its operands directly. The weight grad
// stages 32-pixel slabs of dy and x through LDS, transposed, so that its K
// (the pixels) is contiguous; it splits the pixels over workgroups and sums
// the split partials in a fixed order (deterministic). The first layer (one
// input channel, K = 9 taps) packs its taps into one padded K chunk.
//
// MFMA operand layout (16x16x32 bf16, wave64), checked on the device: lane l
// holds A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16], j < 8, and
// D[4 (l / 16) + r][l % 16] in its four accumulators r.
//
// The forward and input-gradient GEMMs put the channels on the MFMA rows (A,
// from the weights) and the pixels on its columns (B), and pair 16-row groups
// with a row order that gives each lane eight consecutive channels of one
// pixel: row i of group 2h + e is channel 32 h + 8 (i / 4) + 4 e + i % 4, so
// lane l ends with channels 32 h + 8 (l / 16) + 0..7 of pixel l % 16 and
// writes them with one 16-byte store.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>

#include "nanodrr_amd.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kThreads = 256;      // four waves
constexpr int kPixelsPerWave = 32; // two 16-row MFMA tiles
constexpr int kCin1Groups = 8;     // 32-pixel groups per wave in the one-channel forward
// Weight-grad workgroups aimed at (the split size follows): two per CU, all
// resident at once, for the MFMA kernel (more would only add split partials
// to sum); four per CU for the memory-bound one-channel layer.
constexpr int kWgradBlocks = 512;
constexpr int kWgradBlocksCin1 = 1024;

// LDS buffers of the weight-grad slabs (1 or 2). The kernel is held to two
// waves per SIMD (register-bound); one buffer then lets two workgroups share
// a CU, which measured as fast as two buffers with one.
#ifndef NDRR_WGRAD_BUFFERS
#define NDRR_WGRAD_BUFFERS 1
#endif

// The halo kernel (k_conv_halo) for the stride-1 layers of 64-pixel rows and 64 channels (1),
// or the tiled and halo-per-tap kernels there too (0).
#ifndef NDRR_CONV_HALO
#define NDRR_CONV_HALO 1
#endif

struct ConvP {
  const __bf16* x;    // [N][H][W][Cin]
  const __bf16* w;    // [Cout][9][Cin]      (forward)
  const __bf16* wt;   // [9][Cin][Cout]      (input grad)
  const float* bias;  // [Cout]
  const __bf16* dy;   // [N][Ho][Wo][Cout]
  __bf16* y;          // [N][Ho][Wo][Cout]   (forward)
  __bf16* dx;         // [N][H][W][Cin]      (input grad)
  float* part;        // [splits][9 Cout Cin + Cout] (weight grad, then bias grad)
  int bias_grad;      // weight grad: also sum dy per output channel
  int N, H, W, Cin, Ho, Wo, Cout, S, relu, mask;
  int split;          // output pixels per weight-grad workgroup
};

__device__ inline bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int j = 0; j < 8; ++j) z[j] = static_cast<__bf16>(0.0f);
  return z;
}

__device__ inline bf16x8 load8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }

__device__ inline f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Channel (within a block's channel range) of the A row that lane l loads in row group c.
__device__ inline int pair_row(int l, int c) {
  return 32 * (c >> 1) + 8 * ((l & 15) >> 2) + 4 * (c & 1) + (l & 3);
}

// ---------------------------------------------------------------------------
// Forward of the one-channel first layer: K = the nine taps of the single
// input channel, zero-padded to 32, so one MFMA per tile and no K loop.
// Block: 4 waves, each over kCin1Groups groups of 32 output pixels; 16 * COG
// output channels.
// ---------------------------------------------------------------------------
template <int COG>
__global__ __launch_bounds__(kThreads) void k_conv_fwd(ConvP p) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int M = p.N * p.Ho * p.Wo;
  const int co0 = blockIdx.y * 16 * COG;
  const int kq = 8 * (l / 16);
  // weights and bias once per wave; the wave then walks kCin1Groups groups of 32 pixels
  bf16x8 b[COG];
#pragma unroll
  for (int c = 0; c < COG; ++c) {
    const int co = co0 + pair_row(l, c);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int tap = kq + j;
      b[c][j] = tap < 9 ? p.w[co * 9 + tap] : static_cast<__bf16>(0.0f);
    }
  }
  float bv[COG / 2][8];
#pragma unroll
  for (int h = 0; h < COG / 2; ++h)
#pragma unroll
    for (int j = 0; j < 8; ++j) bv[h][j] = p.bias ? p.bias[co0 + 32 * h + kq + j] : 0.0f;
  for (int it = 0; it < kCin1Groups; ++it) {
    const int m0 = ((blockIdx.x * 4 + wave) * kCin1Groups + it) * kPixelsPerWave;
    if (m0 >= M) break;
    bf16x8 a[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int m = m0 + 16 * g + (l % 16);
      const bool pv = m < M;
      const int mm = pv ? m : 0;
      const int ox = mm % p.Wo, oy = (mm / p.Wo) % p.Ho, n = mm / (p.Wo * p.Ho);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int tap = kq + j;
        float v = 0.0f;
        if (tap < 9 && pv) {
          const int iy = oy * p.S + tap / 3 - 1, ix = ox * p.S + tap % 3 - 1;
          if (iy >= 0 && iy < p.H && ix >= 0 && ix < p.W)
            v = static_cast<float>(p.x[(static_cast<int64_t>(n) * p.H + iy) * p.W + ix]);
        }
        a[g][j] = static_cast<__bf16>(v);
      }
    }
    f32x4 acc[2][COG];
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int c = 0; c < COG; ++c) acc[g][c] = mfma(b[c], a[g], f32x4{0.0f, 0.0f, 0.0f, 0.0f});
    // bias, ReLU, bf16, one 16-byte NHWC store of eight channels per pixel
#pragma unroll
    for (int h = 0; h < COG / 2; ++h) {
      const int cb = co0 + 32 * h + kq;
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const int m = m0 + 16 * g + (l % 16);
        if (m < M) {
          bf16x8 v;
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            float u = acc[g][2 * h + j / 4][j % 4] + bv[h][j];
            if (p.relu) u = fmaxf(u, 0.0f);
            v[j] = static_cast<__bf16>(u);
          }
          *reinterpret_cast<bf16x8*>(p.y + static_cast<int64_t>(m) * p.Cout + cb) = v;
        }
      }
    }
  }
}

// Forward of the one-channel first layer at stride 1 when every output row holds whole
// 32-pixel groups (Wo % 32 == 0). A group is 32 consecutive pixels of one row, so its taps
// are the 3 x 34 input values around it: the wave loads them once, coalesced (lane c reads
// column ox0 - 1 + c of each of the three rows, zero outside the image), into its own LDS
// strip, and each lane reads its operand's taps from there. The wave walks its groups along
// the rows, advancing the row and image counters instead of dividing per pixel. Same
// operands, products and epilogue as k_conv_fwd: bitwise the same output.
constexpr int kStrip = 40;  // 34 columns used

// LDS written by some lanes of a wave and then read by others: keep the compiler from moving
// the accesses across this point and wait for the wave's LDS operations.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int COG>
__global__ __launch_bounds__(kThreads) void k_conv_fwd_rows(ConvP p) {
  __shared__ __bf16 strip[4][3][kStrip];
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int co0 = blockIdx.y * 16 * COG;
  const int kq = 8 * (l / 16);
  bf16x8 b[COG];
#pragma unroll
  for (int c = 0; c < COG; ++c) {
    const int co = co0 + pair_row(l, c);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int tap = kq + j;
      b[c][j] = tap < 9 ? p.w[co * 9 + tap] : static_cast<__bf16>(0.0f);
    }
  }
  float bv[COG / 2][8];
#pragma unroll
  for (int h = 0; h < COG / 2; ++h)
#pragma unroll
    for (int j = 0; j < 8; ++j) bv[h][j] = p.bias ? p.bias[co0 + 32 * h + kq + j] : 0.0f;
  const int segs = p.Wo / 32, rows = p.N * p.Ho;
  const int g0 = (blockIdx.x * 4 + wave) * kCin1Groups;  // this wave's first group
  int row = g0 / segs, seg = g0 - (g0 / segs) * segs;
  int n = row / p.Ho, oy = row - (row / p.Ho) * p.Ho;
  for (int it = 0; it < kCin1Groups && row < rows; ++it) {
    const int ox0 = 32 * seg;
    if (l < 34) {
      const int ix = ox0 - 1 + l;
      const bool inx = ix >= 0 && ix < p.W;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const int iy = oy - 1 + r;
        strip[wave][r][l] = (inx && iy >= 0 && iy < p.H) ? p.x[(static_cast<int64_t>(n) * p.H + iy) * p.W + ix]
                                                         : static_cast<__bf16>(0.0f);
      }
    }
    wave_lds_sync();
    bf16x8 a[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int q = 16 * g + (l & 15);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int tap = kq + j;
        a[g][j] = tap < 9 ? strip[wave][tap / 3][q + tap % 3] : static_cast<__bf16>(0.0f);
      }
    }
    wave_lds_sync();  // every lane has read the strip before the next group rewrites it
    f32x4 acc[2][COG];
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int c = 0; c < COG; ++c) acc[g][c] = mfma(b[c], a[g], f32x4{0.0f, 0.0f, 0.0f, 0.0f});
    const int64_t m0 = static_cast<int64_t>(row) * p.Wo + ox0;
#pragma unroll
    for (int h = 0; h < COG / 2; ++h) {
      const int cb = co0 + 32 * h + kq;
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const int64_t m = m0 + 16 * g + (l & 15);
        bf16x8 v;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          float u = acc[g][2 * h + j / 4][j % 4] + bv[h][j];
          if (p.relu) u = fmaxf(u, 0.0f);
          v[j] = static_cast<__bf16>(u);
        }
        *reinterpret_cast<bf16x8*>(p.y + m * p.Cout + cb) = v;
      }
    }
    if (++seg == segs) {
      seg = 0;
      ++row;
      if (++oy == p.Ho) {
        oy = 0;
        ++n;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Tiled implicit GEMM: the forward (MODE 0) and the input gradient (MODE 1)
// of layers with Cin a multiple of 32. Block: four waves as WM x (4 / WM);
// wave tile 64 pixels x 16 CG channels (4 x CG MFMA tiles), block tile
// BM = 64 WM pixels x BN = 16 CG (4 / WM) channels. K runs in steps of 32
// channels at one tap. Each step's operands are staged through LDS (A: BN
// weight rows, B: BM activation rows, 32 K elements each in 80-byte rows, so
// the 16-byte reads of 16 consecutive rows hit distinct banks), double
// buffered: the next step's global loads are in flight during this step's
// MFMAs, and one barrier ends each step.
//   MODE 0: rows are output pixels, columns output channels, the B row of a
//           pixel at tap (ky, kx) is x at (oy S + ky - 1, ox S + kx - 1).
//   MODE 1: rows are input pixels of one parity class (blockIdx.z) for stride
//           2, columns input channels, K over the class's taps and the output
//           channels; the B row is dy at ((iy + 1 - ky) / S, (ix + 1 - kx) / S)
//           when that divides exactly and lies inside.
// ---------------------------------------------------------------------------
constexpr int kKRow = 40;  // 32 K elements + 8 of padding

template <int MODE, int WM, int CG>
__global__ __launch_bounds__(kThreads) void k_igemm(ConvP p) {
  constexpr int WN = 4 / WM;
  constexpr int BM = 64 * WM, BN = 16 * CG * WN;
  constexpr int BQ = BM * 4 / kThreads;                   // 16-byte B chunks per thread
  constexpr int AQ = (BN * 4 + kThreads - 1) / kThreads;  // 16-byte A chunks per thread
  __shared__ __attribute__((aligned(16))) __bf16 sA[2][BN][kKRow];
  __shared__ __attribute__((aligned(16))) __bf16 sB[2][BM][kKRow];
  const int t = threadIdx.x, wave = t >> 6, l = t & 63;
  const int wm = wave % WM, wn = wave / WM;
  const int ch0 = blockIdx.y * BN;
  const bool cls = MODE == 1 && p.S == 2;
  const int py = cls ? static_cast<int>(blockIdx.z) >> 1 : 0;
  const int px = cls ? static_cast<int>(blockIdx.z) & 1 : 0;
  const int Hr = MODE == 0 ? p.Ho : (cls ? (p.H - py + 1) / 2 : p.H);
  const int Wr = MODE == 0 ? p.Wo : (cls ? (p.W - px + 1) / 2 : p.W);
  const int M = p.N * Hr * Wr;
  const int m0 = blockIdx.x * BM;
  if (m0 >= M) return;  // the whole block is past this class (before any barrier)
  const int Ck = MODE == 0 ? p.Cin : p.Cout;
  const int KC = Ck / 32;
  const int ky0 = cls ? (py == 0 ? 1 : 0) : 0, kx0 = cls ? (px == 0 ? 1 : 0) : 0;
  const int ks = cls ? 2 : 1;
  const int nkx = (3 - kx0 + ks - 1) / ks, nky = (3 - ky0 + ks - 1) / ks;
  const int T = nky * nkx * KC;

  // this thread's B rows t / 4 + 64 i and 16-byte part t % 4 (also its A rows)
  const int part = t & 3;
  int rn[BQ], ry[BQ], rx[BQ];
  bool rv[BQ];
#pragma unroll
  for (int i = 0; i < BQ; ++i) {
    const int m = m0 + t / 4 + 64 * i;
    rv[i] = m < M;
    const int mm = rv[i] ? m : 0;
    const int cx = mm % Wr, cy = (mm / Wr) % Hr;
    rn[i] = mm / (Wr * Hr);
    rx[i] = cls ? 2 * cx + px : cx;
    ry[i] = cls ? 2 * cy + py : cy;
  }
  bf16x8 rb[BQ], ra[AQ];
  auto gload = [&](int s) {
    const int ti = s / KC, c0 = (s - ti * KC) * 32 + 8 * part;
    const int ky = ky0 + ks * (ti / nkx), kx = kx0 + ks * (ti % nkx);
    const int tap = 3 * ky + kx;
#pragma unroll
    for (int i = 0; i < BQ; ++i) {
      bool ok;
      int64_t off;
      if constexpr (MODE == 0) {
        const int iy = ry[i] * p.S + ky - 1, ix = rx[i] * p.S + kx - 1;
        ok = rv[i] && iy >= 0 && iy < p.H && ix >= 0 && ix < p.W;
        off = ((static_cast<int64_t>(rn[i]) * p.H + iy) * p.W + ix) * p.Cin + c0;
      } else {
        const int ty = ry[i] + 1 - ky, tx = rx[i] + 1 - kx;
        const int oy = ty / p.S, ox = tx / p.S;
        ok = rv[i] && ty >= 0 && tx >= 0 && ty == oy * p.S && tx == ox * p.S && oy < p.Ho && ox < p.Wo;
        off = ((static_cast<int64_t>(rn[i]) * p.Ho + oy) * p.Wo + ox) * p.Cout + c0;
      }
      rb[i] = ok ? load8((MODE == 0 ? p.x : p.dy) + off) : zero8();
    }
#pragma unroll
    for (int i = 0; i < AQ; ++i) {
      const int r = t / 4 + 64 * i;
      if (r < BN) {
        if constexpr (MODE == 0)
          ra[i] = load8(p.w + (static_cast<int64_t>(ch0 + r) * 9 + tap) * p.Cin + c0);
        else
          ra[i] = load8(p.wt + (static_cast<int64_t>(tap) * p.Cin + ch0 + r) * p.Cout + c0);
      }
    }
  };
  auto lstore = [&](int b) {
#pragma unroll
    for (int i = 0; i < BQ; ++i) *reinterpret_cast<bf16x8*>(&sB[b][t / 4 + 64 * i][8 * part]) = rb[i];
#pragma unroll
    for (int i = 0; i < AQ; ++i)
      if (t / 4 + 64 * i < BN) *reinterpret_cast<bf16x8*>(&sA[b][t / 4 + 64 * i][8 * part]) = ra[i];
  };

  f32x4 acc[4][CG];
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int c = 0; c < CG; ++c) acc[g][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int kq = 8 * (l >> 4);
  gload(0);
  lstore(0);
  __syncthreads();
  for (int s = 0; s < T; ++s) {
    const int b = s & 1;
    if (s + 1 < T) gload(s + 1);
    bf16x8 bx[4], ax[CG];
#pragma unroll
    for (int g = 0; g < 4; ++g) bx[g] = *reinterpret_cast<const bf16x8*>(&sB[b][wm * 64 + 16 * g + (l & 15)][kq]);
#pragma unroll
    for (int c = 0; c < CG; ++c)
      ax[c] = *reinterpret_cast<const bf16x8*>(&sA[b][wn * 16 * CG + pair_row(l, c)][kq]);
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int c = 0; c < CG; ++c) acc[g][c] = mfma(ax[c], bx[g], acc[g][c]);
    if (s + 1 < T) lstore(b ^ 1);
    __syncthreads();
  }

  // lane l: channels cb .. cb + 7 (cb = ch0 + 16 CG wn + 32 h + 8 (l / 16)) of row 16 g + l % 16
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int m = m0 + wm * 64 + 16 * g + (l & 15);
    if (m >= M) continue;
    int64_t at;
    if constexpr (MODE == 0) {
      at = static_cast<int64_t>(m) * p.Cout;
    } else {
      const int cx = m % Wr, cy = (m / Wr) % Hr, n = m / (Wr * Hr);
      const int ix = cls ? 2 * cx + px : cx, iy = cls ? 2 * cy + py : cy;
      at = ((static_cast<int64_t>(n) * p.H + iy) * p.W + ix) * p.Cin;
    }
#pragma unroll
    for (int h = 0; h < CG / 2; ++h) {
      const int cb = ch0 + 16 * CG * wn + 32 * h + kq;
      bf16x8 v;
      if constexpr (MODE == 0) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          float u = acc[g][2 * h + j / 4][j % 4] + (p.bias ? p.bias[cb + j] : 0.0f);
          if (p.relu) u = fmaxf(u, 0.0f);
          v[j] = static_cast<__bf16>(u);
        }
        *reinterpret_cast<bf16x8*>(p.y + at + cb) = v;
      } else {
        const bf16x8 xm = p.mask ? load8(p.x + at + cb) : zero8();
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          float u = acc[g][2 * h + j / 4][j % 4];
          if (p.mask && !(static_cast<float>(xm[j]) > 0.0f)) u = 0.0f;
          v[j] = static_cast<__bf16>(u);
        }
        *reinterpret_cast<bf16x8*>(p.dx + at + cb) = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Input gradient at stride 2 from a staged dy tile. A workgroup covers an input tile of
// 2 kDgTY x 2 kDgTX pixels of one image and BN = 16 CG input channels: the four parity
// classes (py, px) of that tile, kDgTY x kDgTX pixels each, one class per wave. Every tap
// of class (py, px) reads dy at (cy + (py + 1 - ky) / 2, cx + (px + 1 - kx) / 2) for class
// pixel (cy, cx), so the whole workgroup needs the (kDgTY + 1) x (kDgTX + 1) dy pixels from
// the tile's corner: they are staged in LDS once (all Cout channels, zero outside dy), in
// rows of Cout + 8 channels so that the 16-byte reads of 16 consecutive pixels hit distinct
// banks, and each wave then runs its class's taps (1, 2, 2 or 4) x Cout / 32 K steps from
// there, with the weights read from L2. The K order (taps row-major, then channel chunks),
// the products and the epilogue are k_igemm<1>'s: bitwise the same dx, without re-reading
// dy per tap and class.
// ---------------------------------------------------------------------------
constexpr int kDgTY = 4, kDgTX = 16;
constexpr int kDgTile = (kDgTY + 1) * (kDgTX + 1);  // dy pixels staged per workgroup

template <int CG>
__global__ __launch_bounds__(kThreads) void k_dgrad_s2(ConvP p) {
  extern __shared__ __attribute__((aligned(16))) __bf16 sdy[];  // [kDgTile][Cout + 8]
  constexpr int TY = kDgTY, TILE = kDgTile;
  const int t = threadIdx.x, wave = t >> 6, l = t & 63;
  const int row = p.Cout + 8;
  const int ch0 = blockIdx.y * 16 * CG;
  const int Hc = (p.H + 1) / 2, Wc = (p.W + 1) / 2;  // class rows / columns (class 0's, the largest)
  const int tiles_x = (Wc + kDgTX - 1) / kDgTX, tiles_y = (Hc + TY - 1) / TY;
  const int n = blockIdx.x / (tiles_x * tiles_y);
  const int rest = blockIdx.x - n * tiles_x * tiles_y;
  const int cy0 = (rest / tiles_x) * TY, cx0 = (rest % tiles_x) * kDgTX;
  // stage dy[n][cy0 .. cy0 + kDgTY][cx0 .. cx0 + kDgTX][:] (16-byte pieces, zero outside)
  const int pieces = p.Cout / 8;
  for (int e = t; e < TILE * pieces; e += kThreads) {
    const int pi = e / pieces, c = e - pi * pieces;
    const int oy = cy0 + pi / (kDgTX + 1), ox = cx0 + pi % (kDgTX + 1);
    const bf16x8 v = (oy < p.Ho && ox < p.Wo)
                         ? load8(p.dy + ((static_cast<int64_t>(n) * p.Ho + oy) * p.Wo + ox) * p.Cout + 8 * c)
                         : zero8();
    *reinterpret_cast<bf16x8*>(&sdy[pi * row + 8 * c]) = v;
  }
  const int py = wave >> 1, px = wave & 1;
  const int ky0 = py == 0 ? 1 : 0, kx0 = px == 0 ? 1 : 0;
  const int nky = py == 0 ? 1 : 2, nkx = px == 0 ? 1 : 2;
  const int kq = 8 * (l >> 4), tx = l & 15;
  // The ReLU mask of the 32-channel layer is loaded before the barrier and the products (its
  // latency hidden behind them); wider layers load it in the epilogue (registers).
  constexpr bool kEarly = CG == 2;
  bf16x8 xm_early[TY];
  if constexpr (kEarly) {
#pragma unroll
    for (int g = 0; g < TY; ++g) {
      const int iy = 2 * (cy0 + g) + py, ix = 2 * (cx0 + tx) + px;
      xm_early[g] = (p.mask && iy < p.H && ix < p.W)
                        ? load8(p.x + ((static_cast<int64_t>(n) * p.H + iy) * p.W + ix) * p.Cin + ch0 + kq)
                        : zero8();
    }
  }
  __syncthreads();
  f32x4 acc[TY][CG];
#pragma unroll
  for (int g = 0; g < TY; ++g)
#pragma unroll
    for (int c = 0; c < CG; ++c) acc[g][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int a = 0; a < nky; ++a) {
    const int ky = ky0 + 2 * a, dyo = (py + 1 - ky) / 2;
    for (int b = 0; b < nkx; ++b) {
      const int kx = kx0 + 2 * b, dxo = (px + 1 - kx) / 2;
      const int tap = 3 * ky + kx;
      for (int kc = 0; kc < p.Cout; kc += 32) {
        bf16x8 bx[TY], ax[CG];
#pragma unroll
        for (int c = 0; c < CG; ++c)
          ax[c] = load8(p.wt + (static_cast<int64_t>(tap) * p.Cin + ch0 + pair_row(l, c)) * p.Cout + kc + kq);
#pragma unroll
        for (int g = 0; g < TY; ++g)
          bx[g] = *reinterpret_cast<const bf16x8*>(&sdy[((g + dyo) * (kDgTX + 1) + tx + dxo) * row + kc + kq]);
#pragma unroll
        for (int g = 0; g < TY; ++g)
#pragma unroll
          for (int c = 0; c < CG; ++c) acc[g][c] = mfma(ax[c], bx[g], acc[g][c]);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < TY; ++g) {
    const int iy = 2 * (cy0 + g) + py, ix = 2 * (cx0 + tx) + px;
    if (iy >= p.H || ix >= p.W) continue;
    const int64_t at = ((static_cast<int64_t>(n) * p.H + iy) * p.W + ix) * p.Cin;
#pragma unroll
    for (int h = 0; h < CG / 2; ++h) {
      const int cb = ch0 + 32 * h + kq;
      bf16x8 xm;
      if constexpr (kEarly) xm = xm_early[g];
      else xm = p.mask ? load8(p.x + at + cb) : zero8();
      bf16x8 v;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        float u = acc[g][2 * h + j / 4][j % 4];
        if (p.mask && !(static_cast<float>(xm[j]) > 0.0f)) u = 0.0f;
        v[j] = static_cast<__bf16>(u);
      }
      *reinterpret_cast<bf16x8*>(p.dx + at + cb) = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Stride-1 input gradient from a staged dy halo, for layers with up to 64 output channels (the
// forward and wider layers measured faster on k_igemm, whose weights staged in LDS serve four
// waves at once). A workgroup covers a kHaloT x kHaloT tile of input-gradient pixels of one
// image and BN = 16 CG input channels; wave w takes its rows 4w .. 4w + 3 (64 pixels: row group
// g, column l % 16). Every tap of the tile reads dy at (y + 1 - ky, x + 1 - kx), inside the
// (kHaloT + 2)^2 pixels around the tile, so K runs over 32-channel chunks of dy, each chunk's halo
// staged once in LDS (80-byte rows: the 16-byte reads of 16 consecutive pixels hit distinct
// banks) and its nine taps read from there, the next chunk's halo loaded into registers
// meanwhile (double-buffered LDS). The weights come from L1/L2, the next tap's loaded ahead.
// Epilogue as k_igemm<1>'s.
// ---------------------------------------------------------------------------
constexpr int kHaloT = 16;
constexpr int kHaloS = kHaloT + 2;                      // halo side
constexpr int kHaloPix = kHaloS * kHaloS;               // 324 pixels
constexpr int kHaloQ = (kHaloPix * 4 + kThreads - 1) / kThreads;  // 16-byte pieces per thread

template <int CG>
__global__ __launch_bounds__(kThreads) void k_dgrad_halo_s1(ConvP p) {
  __shared__ __attribute__((aligned(16))) __bf16 sh[2][kHaloPix][kKRow];
  const int t = threadIdx.x, wave = t >> 6, l = t & 63;
  const int ch0 = blockIdx.y * 16 * CG;
  const int tiles_x = (p.W + kHaloT - 1) / kHaloT, tiles_y = (p.H + kHaloT - 1) / kHaloT;
  const int n = blockIdx.x / (tiles_x * tiles_y);
  const int rest = blockIdx.x - n * tiles_x * tiles_y;
  const int y0 = (rest / tiles_x) * kHaloT, x0 = (rest % tiles_x) * kHaloT;
  const int K = p.Cout;  // channels of the staged source, dy
  const __bf16* src = p.dy;
  const int KC = K / 32;
  // this thread's halo pieces: pixel e / 4, 16-byte part e % 4, e = t + 256 i
  bf16x8 hv[kHaloQ];
  auto hload = [&](int kc) {
#pragma unroll
    for (int i = 0; i < kHaloQ; ++i) {
      const int e = t + kThreads * i;
      const int pi = e >> 2, part = e & 3;
      const int sy = y0 - 1 + pi / kHaloS, sx = x0 - 1 + pi % kHaloS;
      const bool ok = e < kHaloPix * 4 && sy >= 0 && sy < p.H && sx >= 0 && sx < p.W;
      hv[i] = ok ? load8(src + ((static_cast<int64_t>(n) * p.H + sy) * p.W + sx) * K + 32 * kc + 8 * part) : zero8();
    }
  };
  auto hstore = [&](int b) {
#pragma unroll
    for (int i = 0; i < kHaloQ; ++i) {
      const int e = t + kThreads * i;
      if (e < kHaloPix * 4) *reinterpret_cast<bf16x8*>(&sh[b][e >> 2][8 * (e & 3)]) = hv[i];
    }
  };
  const int kq = 8 * (l >> 4), tx = l & 15;
  // weights of (tap, chunk): wt[tap][ch][k]
  auto wload = [&](int tap, int kc, bf16x8 (&a)[CG]) {
#pragma unroll
    for (int c = 0; c < CG; ++c)
      a[c] = load8(p.wt + (static_cast<int64_t>(tap) * p.Cin + ch0 + pair_row(l, c)) * p.Cout + 32 * kc + kq);
  };
  f32x4 acc[4][CG];
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int c = 0; c < CG; ++c) acc[g][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  hload(0);
  hstore(0);
  bf16x8 ax[CG], an[CG];
  wload(0, 0, ax);
  __syncthreads();
  for (int kc = 0; kc < KC; ++kc) {
    const int b = kc & 1;
    if (kc + 1 < KC) hload(kc + 1);
#pragma unroll 1
    for (int tap = 0; tap < 9; ++tap) {  // rolled: one tap's weights in flight ahead, not nine
      // the next (tap, chunk)'s weights ahead
      if (tap < 8) wload(tap + 1, kc, an);
      else if (kc + 1 < KC) wload(0, kc + 1, an);
      const int oy = 2 - tap / 3, ox = 2 - tap % 3;  // dy at (y + 1 - ky, x + 1 - kx)
      bf16x8 bx[4];
#pragma unroll
      for (int g = 0; g < 4; ++g)
        bx[g] = *reinterpret_cast<const bf16x8*>(&sh[b][(4 * wave + g + oy) * kHaloS + tx + ox][kq]);
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int c = 0; c < CG; ++c) acc[g][c] = mfma(ax[c], bx[g], acc[g][c]);
#pragma unroll
      for (int c = 0; c < CG; ++c) ax[c] = an[c];
    }
    if (kc + 1 < KC) {
      hstore(b ^ 1);
      __syncthreads();
    }
  }
  // lane l: channels cb .. cb + 7 (cb = ch0 + 32 h + kq) of pixel (y0 + 4 wave + g, x0 + tx)
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int y = y0 + 4 * wave + g, x = x0 + tx;
    if (y >= p.H || x >= p.W) continue;
    const int64_t pix = (static_cast<int64_t>(n) * p.H + y) * p.W + x;
#pragma unroll
    for (int h = 0; h < CG / 2; ++h) {
      const int cb = ch0 + 32 * h + kq;
      const bf16x8 xm = p.mask ? load8(p.x + pix * p.Cin + cb) : zero8();
      bf16x8 v;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        float u = acc[g][2 * h + j / 4][j % 4];
        if (p.mask && !(static_cast<float>(xm[j]) > 0.0f)) u = 0.0f;
        v[j] = static_cast<__bf16>(u);
      }
      *reinterpret_cast<bf16x8*>(p.dx + pix * p.Cin + cb) = v;
    }
  }
}

void launch_dgrad_halo_s1(const ConvP& p, hipStream_t st) {
  const unsigned tiles = static_cast<unsigned>(p.N * ((p.H + kHaloT - 1) / kHaloT) * ((p.W + kHaloT - 1) / kHaloT));
  if (p.Cin % 64 == 0) hipLaunchKernelGGL((k_dgrad_halo_s1<4>), dim3(tiles, p.Cin / 64), dim3(kThreads), 0, st, p);
  else hipLaunchKernelGGL((k_dgrad_halo_s1<2>), dim3(tiles, p.Cin / 32), dim3(kThreads), 0, st, p);
}

// ---------------------------------------------------------------------------
// Halo implicit GEMM for 64-pixel rows and 64 streamed channels at stride 1: the forward
// (MODE 0, Cin = 64) and the input gradient (MODE 1, Cout = 64) of a 64 x 64 layer. A
// workgroup keeps its 64 output columns' weights for all nine taps in LDS (the panel: row
// 16 c + i holds the channel that pair_row gives lane i of row group c, 9 * 64 + 8 elements
// long, so a lane group's 16-byte reads of 16 rows hit distinct banks) and runs over tiles
// of four image rows, one row per wave. Each tile's source rows (the four and one above and
// below, 66 pixels with the padding columns) are staged in LDS once, with whole-line loads,
// in pixel rows of 64 + 8 channels (16-byte reads of 16 consecutive pixels hit distinct
// banks); the next tile's are loaded into registers while this tile's products run. Every
// K step then reads both operands from LDS: the activations are read from memory about 1.5
// times instead of once per tap. The K order (taps row-major, then 32-channel chunks), the
// products and the epilogues are k_igemm's: the results are bitwise the same.
//   MODE 0: tap (ky, kx) of output pixel (y, x) reads x at (y + ky - 1, x + kx - 1).
//   MODE 1: tap (ky, kx) of input pixel (y, x) reads dy at (y + 1 - ky, x + 1 - kx).
// ---------------------------------------------------------------------------
constexpr int kHaloCols = 66;                  // 64 pixels and a padding column each side
constexpr int kHaloPitch = 72;                 // 64 channels + 8, in elements
constexpr int kHaloPieces = 6 * kHaloCols * 8; // 16-byte pieces of a tile's source rows
constexpr int kHaloLoads = (kHaloPieces + kThreads - 1) / kThreads;
constexpr int kPanelRow = 9 * 64 + 8;

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_conv_halo(ConvP p) {
  constexpr int CG = 4;
  __shared__ __attribute__((aligned(16))) __bf16 panel[16 * CG * kPanelRow];
  __shared__ __attribute__((aligned(16))) __bf16 halo[6 * kHaloCols * kHaloPitch];
  const int t = threadIdx.x, wave = t >> 6, l = t & 63;
  const int ch0 = blockIdx.y * 16 * CG;
  for (int e = t; e < 16 * CG * 72; e += kThreads) {  // 72 pieces of 8 per panel row
    const int r = e / 72, q = e - r * 72;
    const int c = r >> 4, i = r & 15;
    const int ch = ch0 + 32 * (c >> 1) + 8 * (i >> 2) + 4 * (c & 1) + (i & 3);
    const __bf16* src;
    if constexpr (MODE == 0) {
      src = p.w + static_cast<int64_t>(ch) * 576 + 8 * q;  // w[ch][tap][ci]
    } else {
      const int tap = q >> 3, part = q & 7;
      src = p.wt + (static_cast<int64_t>(tap) * p.Cin + ch) * 64 + 8 * part;  // wt[tap][ch][co]
    }
    *reinterpret_cast<bf16x8*>(panel + r * kPanelRow + 8 * q) = load8(src);
  }
  // rows: output pixels (MODE 0) or input pixels (MODE 1) on an Hr x 64 grid; source Hr x 64 x 64
  const int Hr = MODE == 0 ? p.Ho : p.H;
  const int quads = (Hr + 3) / 4;
  const int tiles = p.N * quads;
  const __bf16* const srcp = MODE == 0 ? p.x : p.dy;
  bf16x8 hv[kHaloLoads];
  auto hload = [&](int tile) {  // the tile's source rows r0 - 1 .. r0 + 4, columns -1 .. 64
    const int n = tile / quads, r0 = (tile - n * quads) * 4;
#pragma unroll
    for (int k = 0; k < kHaloLoads; ++k) {
      const int e = t + kThreads * k;
      const int hr = e / (kHaloCols * 8), rem = e - hr * (kHaloCols * 8);
      const int hc = rem >> 3, part = rem & 7;
      const int sy = r0 - 1 + hr, sx = hc - 1;
      const bool ok = e < kHaloPieces && tile < tiles && sy >= 0 && sy < Hr && sx >= 0 && sx < 64;
      hv[k] = ok ? load8(srcp + ((static_cast<int64_t>(n) * Hr + sy) * 64 + sx) * 64 + 8 * part) : zero8();
    }
  };
  auto hstore = [&]() {
#pragma unroll
    for (int k = 0; k < kHaloLoads; ++k) {
      const int e = t + kThreads * k;
      if (e < kHaloPieces) {
        const int hr = e / (kHaloCols * 8), rem = e - hr * (kHaloCols * 8);
        *reinterpret_cast<bf16x8*>(halo + (hr * kHaloCols + (rem >> 3)) * kHaloPitch + 8 * (rem & 7)) = hv[k];
      }
    }
  };
  const int kq = 8 * (l >> 4), px = l & 15;
  const __bf16* const arow = panel + px * kPanelRow + kq;
  int tile = blockIdx.x;
  hload(tile);
  for (; tile < tiles; tile += gridDim.x) {
    __syncthreads();  // the panel is staged, and every wave is done with the last tile's rows
    hstore();
    __syncthreads();
    hload(tile + gridDim.x);  // in flight during this tile's products
    f32x4 acc[4][CG];
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int c = 0; c < CG; ++c) acc[g][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // K step s = (tap, chunk) = (s / 2, s % 2); the next step's operands are read from LDS
    // while this step's products run
    auto operands = [&](int s, bf16x8 (&ax)[CG], bf16x8 (&bx)[4]) {
      const int tap = s >> 1, kc = s & 1;
      const int ky = tap / 3, kx = tap - 3 * ky;
      const int hr = MODE == 0 ? wave + ky : wave + 2 - ky;
      const int hc0 = MODE == 0 ? px + kx : px + 2 - kx;  // halo column of group 0's pixel
      const __bf16* const brow = halo + (hr * kHaloCols + hc0) * kHaloPitch + kq + 32 * kc;
#pragma unroll
      for (int c = 0; c < CG; ++c) ax[c] = *reinterpret_cast<const bf16x8*>(arow + 16 * c * kPanelRow + 32 * s);
#pragma unroll
      for (int g = 0; g < 4; ++g) bx[g] = *reinterpret_cast<const bf16x8*>(brow + 16 * g * kHaloPitch);
    };
    // two operand sets in turn: even steps use set 0, odd steps set 1
    bf16x8 ax[2][CG], bx[2][4];
    operands(0, ax[0], bx[0]);
#pragma unroll 1
    for (int s = 0; s < 18; s += 2) {
      operands(s + 1, ax[1], bx[1]);
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int c = 0; c < CG; ++c) acc[g][c] = mfma(ax[0][c], bx[0][g], acc[g][c]);
      operands(s + 2 < 18 ? s + 2 : s + 1, ax[0], bx[0]);
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int c = 0; c < CG; ++c) acc[g][c] = mfma(ax[1][c], bx[1][g], acc[g][c]);
    }
    // lane l: channels cb .. cb + 7 (cb = ch0 + 32 h + kq) of pixel 16 g + l % 16 of row r0 + wave
    const int n = tile / quads, row = (tile - n * quads) * 4 + wave;
    if (row < Hr) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int64_t m = (static_cast<int64_t>(n) * Hr + row) * 64 + 16 * g + px;
        const int64_t at = m * (MODE == 0 ? p.Cout : p.Cin);
#pragma unroll
        for (int h = 0; h < CG / 2; ++h) {
          const int cb = ch0 + 32 * h + kq;
          bf16x8 v;
          if constexpr (MODE == 0) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              float u = acc[g][2 * h + j / 4][j % 4] + (p.bias ? p.bias[cb + j] : 0.0f);
              if (p.relu) u = fmaxf(u, 0.0f);
              v[j] = static_cast<__bf16>(u);
            }
            *reinterpret_cast<bf16x8*>(p.y + at + cb) = v;
          } else {
            const bf16x8 xm = p.mask ? load8(p.x + at + cb) : zero8();
#pragma unroll
            for (int j = 0; j < 8; ++j) {
              float u = acc[g][2 * h + j / 4][j % 4];
              if (p.mask && !(static_cast<float>(xm[j]) > 0.0f)) u = 0.0f;
              v[j] = static_cast<__bf16>(u);
            }
            *reinterpret_cast<bf16x8*>(p.dx + at + cb) = v;
          }
        }
      }
    }
  }
}

// The halo kernel where it applies (stride 1, 64-pixel rows, 64 streamed channels, 64-channel
// column blocks); returns false elsewhere.
template <int MODE>
bool try_halo(const ConvP& p, hipStream_t st) {
  const int K = MODE == 0 ? p.Cin : p.Cout, cols = MODE == 0 ? p.Cout : p.Cin;
  const int wr = MODE == 0 ? p.Wo : p.W;
  if (!NDRR_CONV_HALO || p.S != 1 || K != 64 || wr != 64 || cols % 64 != 0) return false;
  const int hr = MODE == 0 ? p.Ho : p.H;
  const int64_t tiles = static_cast<int64_t>(p.N) * ((hr + 3) / 4);
  const int ycols = cols / 64;
  const int64_t want = 256 / ycols > 0 ? 256 / ycols : 1;  // one workgroup per CU (LDS-bound)
  const unsigned blocks = static_cast<unsigned>(tiles < want ? tiles : want);
  hipLaunchKernelGGL((k_conv_halo<MODE>), dim3(blocks, ycols), dim3(kThreads), 0, st, p);
  return true;
}

template <int MODE>
void launch_igemm(const ConvP& p, int rows, int cols, int classes, hipStream_t st) {
  if (cols % 128 == 0) {
    hipLaunchKernelGGL((k_igemm<MODE, 2, 4>), dim3((rows + 127) / 128, cols / 128, classes), dim3(kThreads), 0, st, p);
  } else if (cols % 64 == 0) {
    hipLaunchKernelGGL((k_igemm<MODE, 4, 4>), dim3((rows + 255) / 256, cols / 64, classes), dim3(kThreads), 0, st, p);
  } else {
    hipLaunchKernelGGL((k_igemm<MODE, 4, 2>), dim3((rows + 255) / 256, cols / 32, classes), dim3(kThreads), 0, st, p);
  }
}

// ---------------------------------------------------------------------------
// Weight gradients.
// ---------------------------------------------------------------------------
constexpr int kLdsRow = 48;  // one-channel layer: 32 pixels + 16 of padding, rows stay 16-byte aligned

// Weight gradient for Cin % 32 == 0, Cout % 64 == 0, from LDS transpose
// reads. Block (split, CI = 16 CIT input channels, 64 output channels) over
// p.split output pixels, in slabs of 32 pixels: the dy slab [pixel][64] and
// the nine tap slabs of x [tap][pixel][CI] are staged in their NHWC order
// with 16-byte stores, and the MFMA operands (8 consecutive pixels of one
// channel per lane) come from pairs of 64-bit transpose reads: within each
// 16-lane group the read gives lane i element i % 4 of the 8-byte pieces
// that lanes 4 j + i / 4 address (j = 0..3, checked on the device), so lane
// (i, b) addressing pixel row kb + 8 b + i / 4, channels 4 (i % 4)..+3 of a
// 16-channel block, receives pixels kb + 8 b + 0..3 of channel i. Rows keep
// their 32-byte channel blocks XOR-swizzled by the row so that these reads
// hit distinct banks (a 64-lane 8-byte read in two 32-lane halves, modelled
// on the CPU). Wave w takes all four 16-channel output tiles for the
// (tap, input tile) units w, w + 4, ... (nine units per wave when CIT = 4),
// so the dy operands are read once per wave and slab. With bias_grad, wave w
// also sums dy of output tile w against a ones operand. The next slab's
// global loads are in flight during this slab's MFMAs (NB = 2: two LDS
// buffers and one barrier per slab; NB = 1: one buffer and two barriers).
__device__ inline int swz(int row, int blk, int row_bytes) {
  return row_bytes == 64 ? blk ^ ((row >> 3) & 1) : blk ^ (((row >> 1) & 1) | ((row >> 2) & 2));
}

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));

// Eight consecutive pixel rows (kb + 8 (l / 16) + 0..7) of channel 16 blk + l % 16 from a slab of
// 32 rows of row_bytes bytes at `base`.
__device__ inline bf16x8 read_tr(const char* base, int row_bytes, int blk, int l) {
  const int i = l & 15, b = l >> 4;
  s16x4 h[2];
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    const int row = 4 * kb + 8 * b + (i >> 2);
    const char* a = base + row * row_bytes + 32 * swz(row, blk, row_bytes) + 8 * (i & 3);
    h[kb] = __builtin_amdgcn_ds_read_tr16_b64_v4i16((__attribute__((address_space(3))) s16x4*)(a));
  }
  const s16x8 v = __builtin_shufflevector(h[0], h[1], 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_bit_cast(bf16x8, v);
}

template <int CIT, int NB>
__global__ __launch_bounds__(kThreads, 2) void k_conv_bwd_weight(ConvP p) {
  constexpr int CI = 16 * CIT, XRB = 2 * CI;  // x row bytes (64 or 128)
  constexpr int UNITS = 9 * CIT, U = (UNITS + 3) / 4;
  constexpr int XQ = (9 * 32 * CI / 8 + kThreads - 1) / kThreads;  // 16-byte x chunks per thread
  constexpr int DYB = 32 * 128, XB = 9 * 32 * XRB;                // slab bytes
  __shared__ __attribute__((aligned(16))) char lds[NB][DYB + XB];
  const int t = threadIdx.x, wave = t >> 6, l = t & 63;
  const int split = blockIdx.x;
  const int ci0 = blockIdx.y * CI, co0 = blockIdx.z * 64;
  const int M = p.N * p.Ho * p.Wo;
  const int p_begin = split * p.split, p_end = min(M, p_begin + p.split);
  const int slabs = (p_end - p_begin + 31) / 32;
  const bool bias = p.bias_grad && blockIdx.y == 0;

  // this thread's chunks: dy (pixel t / 8, piece t % 8); x (pixel, piece) fixed over the taps
  const int dpx = t >> 3, dpc = t & 7;
  const int xpx = CIT == 4 ? t >> 3 : (t & 127) >> 2, xpc = CIT == 4 ? t & 7 : t & 3;
  bf16x8 rdy, rx[XQ];
  auto gload = [&](int s) {
    const int pb = p_begin + 32 * s;
    {
      const int m = pb + dpx;
      rdy = m < p_end ? load8(p.dy + static_cast<int64_t>(m) * p.Cout + co0 + 8 * dpc) : zero8();
    }
    const int m = pb + xpx;
    const bool in_range = m < p_end;
    const int mm = in_range ? m : 0;
    const int ox = mm % p.Wo, oy = (mm / p.Wo) % p.Ho, n = mm / (p.Wo * p.Ho);
#pragma unroll
    for (int i = 0; i < XQ; ++i) {
      const int tap = CIT == 4 ? i : 2 * i + (t >> 7);
      const int iy = oy * p.S + tap / 3 - 1, ix = ox * p.S + tap % 3 - 1;
      rx[i] = in_range && tap < 9 && iy >= 0 && iy < p.H && ix >= 0 && ix < p.W
                  ? load8(p.x + ((static_cast<int64_t>(n) * p.H + iy) * p.W + ix) * p.Cin + ci0 + 8 * xpc)
                  : zero8();
    }
  };
  auto lstore = [&](int b) {
    char* base = lds[b];
    *reinterpret_cast<bf16x8*>(base + dpx * 128 + 32 * swz(dpx, dpc >> 1, 128) + 16 * (dpc & 1)) = rdy;
#pragma unroll
    for (int i = 0; i < XQ; ++i) {
      const int tap = CIT == 4 ? i : 2 * i + (t >> 7);
      if (tap < 9)
        *reinterpret_cast<bf16x8*>(base + DYB + tap * 32 * XRB + xpx * XRB + 32 * swz(xpx, xpc >> 1, XRB) +
                                   16 * (xpc & 1)) = rx[i];
    }
  };

  f32x4 acc[U][4];
  f32x4 accb = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k < U; ++k)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[k][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  bf16x8 ones;
#pragma unroll
  for (int j = 0; j < 8; ++j) ones[j] = static_cast<__bf16>(1.0f);

  if (slabs > 0) {
    gload(0);
    lstore(0);
  }
  __syncthreads();
  for (int s = 0; s < slabs; ++s) {
    const int b = NB == 2 ? (s & 1) : 0;
    if (s + 1 < slabs) gload(s + 1);
    const char* base = lds[b];
    bf16x8 a[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = read_tr(base, 128, c, l);
    if (bias) accb = mfma(read_tr(base, 128, wave, l), ones, accb);  // a[wave] would select per element
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int u = wave + 4 * k;
      if (u < UNITS) {
        const int tap = u / CIT, tile = u % CIT;
        const bf16x8 bx = read_tr(base + DYB + tap * 32 * XRB, XRB, tile, l);
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[k][c] = mfma(a[c], bx, acc[k][c]);
      }
    }
    if constexpr (NB == 2) {
      if (s + 1 < slabs) lstore(b ^ 1);
      __syncthreads();
    } else {
      __syncthreads();
      if (s + 1 < slabs) lstore(0);
      __syncthreads();
    }
  }
  // partial[split][tap][co][ci], then partial[split][9 Cout Cin + co] for the bias
  float* part = p.part + static_cast<int64_t>(split) * (9 * p.Cout * p.Cin + p.Cout);
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const int u = wave + 4 * k;
    if (u >= UNITS) continue;
    const int tap = u / CIT, tile = u % CIT;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int co = co0 + 16 * c + 4 * (l >> 4) + r, ci = ci0 + 16 * tile + (l & 15);
        part[(tap * p.Cout + co) * p.Cin + ci] = acc[k][c][r];
      }
  }
  if (bias && (l & 15) == 0)
#pragma unroll
    for (int r = 0; r < 4; ++r) part[9 * p.Cout * p.Cin + co0 + 16 * wave + 4 * (l >> 4) + r] = accb[r];
}

// Weight gradient of the one-channel first layer (Cout == 32): a GEMM with the
// 32 output channels as rows, the nine taps (padded to 16) as columns and the
// pixels as K. Each wave stages its own 32-pixel slab of dy and of the nine
// shifted input values through LDS per step (the block's four waves take
// consecutive slabs), and the four wave partials are added in wave order;
// partial[split][tap][co].
__global__ __launch_bounds__(kThreads) void k_conv_bwd_weight_cin1(ConvP p) {
  __shared__ __attribute__((aligned(16))) __bf16 sdy[4][32][kLdsRow];  // [wave][co][pixel]
  __shared__ __attribute__((aligned(16))) __bf16 sx[4][16][kLdsRow];   // [wave][tap][pixel]
  __shared__ float red[4][32][16];
  __shared__ float bred[4][32];
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int split = blockIdx.x;
  const int M = p.N * p.Ho * p.Wo;
  const int p_begin = split * p.split, p_end = min(M, p_begin + p.split);
  const int kq = 8 * (l / 16);
  f32x4 acc[2] = {f32x4{0.0f, 0.0f, 0.0f, 0.0f}, f32x4{0.0f, 0.0f, 0.0f, 0.0f}};
  float db = 0.0f;  // lanes 0-31: this wave's sum of dy for channel l
  for (int base = p_begin; base < p_end; base += 4 * 32) {
    const int pb = base + 32 * wave;
#pragma unroll
    for (int i = 0; i < 2; ++i) {  // dy: 32 pixels x 32 channels, two 8-channel pieces per lane
      const int idx = l + 64 * i, px = idx / 4, c8 = 8 * (idx % 4);
      const int m = pb + px;
      const bf16x8 v = m < p_end ? load8(p.dy + static_cast<int64_t>(m) * p.Cout + c8) : zero8();
#pragma unroll
      for (int j = 0; j < 8; ++j) sdy[wave][c8 + j][px] = v[j];
    }
    {  // x at the nine taps of 32 pixels: lane (px, tap group of 8)
      const int px = l % 32, t0 = 8 * (l / 32);
      const int m = pb + px;
      const bool in_range = m < p_end;
      const int mm = in_range ? m : 0;
      const int ox = mm % p.Wo, oy = (mm / p.Wo) % p.Ho, n = mm / (p.Wo * p.Ho);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int tap = t0 + j;
        float v = 0.0f;
        if (in_range && tap < 9) {
          const int iy = oy * p.S + tap / 3 - 1, ix = ox * p.S + tap % 3 - 1;
          if (iy >= 0 && iy < p.H && ix >= 0 && ix < p.W)
            v = static_cast<float>(p.x[(static_cast<int64_t>(n) * p.H + iy) * p.W + ix]);
        }
        sx[wave][tap][px] = static_cast<__bf16>(v);
      }
    }
    __syncthreads();
    if (p.bias_grad && l < 32) {
#pragma unroll
      for (int q = 0; q < 32; q += 8) {
        const bf16x8 v = *reinterpret_cast<const bf16x8*>(&sdy[wave][l][q]);
#pragma unroll
        for (int j = 0; j < 8; ++j) db += static_cast<float>(v[j]);
      }
    }
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sx[wave][l % 16][kq]);
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sdy[wave][16 * g + (l % 16)][kq]);
      acc[g] = mfma(a, b, acc[g]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int r = 0; r < 4; ++r) red[wave][16 * g + 4 * (l / 16) + r][l % 16] = acc[g][r];
  if (l < 32) bred[wave][l] = db;
  __syncthreads();
  float* part = p.part + static_cast<int64_t>(split) * (9 * 32 + 32);
  for (int e = threadIdx.x; e < 32 * 9; e += kThreads) {
    const int co = e % 32, tap = e / 32;
    part[tap * 32 + co] = ((red[0][co][tap] + red[1][co][tap]) + red[2][co][tap]) + red[3][co][tap];
  }
  if (p.bias_grad && threadIdx.x < 32) {
    const int co = threadIdx.x;
    part[9 * 32 + co] = ((bred[0][co] + bred[1][co]) + bred[2][co]) + bred[3][co];
  }
}

// The same gradient at stride 1 when every output row holds whole 32-pixel groups
// (Wo % 32 == 0; splits are whole 128-pixel steps, so a group never straddles a split or a
// row). Each wave works on its own until the final sum: per group of 32 pixels it transposes
// the dy slab into LDS as above and stages the group's 3 x 34 input values as k_conv_fwd_rows
// does (lane c reads column ox0 - 1 + c of each row, coalesced), reading each tap's eight
// consecutive pixels from that strip; the next group's global loads are issued before the
// current group's LDS work and products. Same operands, products and order of summation as
// k_conv_bwd_weight_cin1: bitwise the same partials.
__global__ __launch_bounds__(kThreads) void k_conv_bwd_weight_cin1_rows(ConvP p) {
  __shared__ __attribute__((aligned(16))) __bf16 sdy[4][32][kLdsRow];  // [wave][co][pixel]
  __shared__ __bf16 strip[4][3][kStrip];
  __shared__ float red[4][32][16];
  __shared__ float bred[4][32];
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int split = blockIdx.x;
  const int M = p.N * p.Ho * p.Wo;
  const int p_begin = split * p.split, p_end = min(M, p_begin + p.split);
  const int kq = 8 * (l / 16);
  f32x4 acc[2] = {f32x4{0.0f, 0.0f, 0.0f, 0.0f}, f32x4{0.0f, 0.0f, 0.0f, 0.0f}};
  float db = 0.0f;  // lanes 0-31: this wave's sum of dy for channel l
  bf16x8 dv[2];
  __bf16 xv[3];
  // group pb's loads: two 8-channel dy pieces per lane, and (lanes < 34) the strip column
  auto load = [&](int pb) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int idx = l + 64 * i, px = idx / 4, c8 = 8 * (idx % 4);
      dv[i] = load8(p.dy + static_cast<int64_t>(pb + px) * p.Cout + c8);
    }
    const int row = pb / p.Wo, ox0 = pb - row * p.Wo;
    const int n = row / p.Ho, oy = row - n * p.Ho;
    const int ix = ox0 - 1 + l;
    const bool inx = l < 34 && ix >= 0 && ix < p.W;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int iy = oy - 1 + r;
      xv[r] = (inx && iy >= 0 && iy < p.H) ? p.x[(static_cast<int64_t>(n) * p.H + iy) * p.W + ix]
                                           : static_cast<__bf16>(0.0f);
    }
  };
  int pb = p_begin + 32 * wave;
  if (pb < p_end) load(pb);
  for (; pb < p_end; pb += 4 * 32) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int idx = l + 64 * i, px = idx / 4, c8 = 8 * (idx % 4);
#pragma unroll
      for (int j = 0; j < 8; ++j) sdy[wave][c8 + j][px] = dv[i][j];
    }
    if (l < 34) {
#pragma unroll
      for (int r = 0; r < 3; ++r) strip[wave][r][l] = xv[r];
    }
    if (pb + 4 * 32 < p_end) load(pb + 4 * 32);  // the next group's loads, in flight meanwhile
    wave_lds_sync();
    if (p.bias_grad && l < 32) {
#pragma unroll
      for (int q = 0; q < 32; q += 8) {
        const bf16x8 v = *reinterpret_cast<const bf16x8*>(&sdy[wave][l][q]);
#pragma unroll
        for (int j = 0; j < 8; ++j) db += static_cast<float>(v[j]);
      }
    }
    bf16x8 b;
    {
      const int tap = l & 15;  // B column; its K rows are pixels kq .. kq + 7
      const int ky = tap / 3, kx = tap - 3 * (tap / 3);
#pragma unroll
      for (int j = 0; j < 8; ++j) b[j] = tap < 9 ? strip[wave][ky][kq + j + kx] : static_cast<__bf16>(0.0f);
    }
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sdy[wave][16 * g + (l % 16)][kq]);
      acc[g] = mfma(a, b, acc[g]);
    }
    wave_lds_sync();  // every lane has read this group's LDS before the next group rewrites it
  }
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int r = 0; r < 4; ++r) red[wave][16 * g + 4 * (l / 16) + r][l % 16] = acc[g][r];
  if (l < 32) bred[wave][l] = db;
  __syncthreads();
  float* part = p.part + static_cast<int64_t>(split) * (9 * 32 + 32);
  for (int e = threadIdx.x; e < 32 * 9; e += kThreads) {
    const int co = e % 32, tap = e / 32;
    part[tap * 32 + co] = ((red[0][co][tap] + red[1][co][tap]) + red[2][co][tap]) + red[3][co][tap];
  }
  if (p.bias_grad && threadIdx.x < 32) {
    const int co = threadIdx.x;
    part[9 * 32 + co] = ((bred[0][co] + bred[1][co]) + bred[2][co]) + bred[3][co];
  }
}


// dw, db = the sums of the split partials, in a fixed order. A workgroup takes kSumElems
// consecutive float4 elements; its thread (q, e) adds the partials of splits q, q + G, q + 2G,
// ... (G = kSumGroups) of element e in that order, eight loads in flight at a time, and the G
// group sums meet in a fixed pairwise tree in LDS. Deterministic.
constexpr int kSumElems = 16;
constexpr int kSumGroups = kThreads / kSumElems;
__global__ __launch_bounds__(kThreads) void k_sum_splits(const float* __restrict__ part, float* __restrict__ dw,
                                                          float* __restrict__ db, int splits, int Cout, int Cin,
                                                          int64_t count) {
  __shared__ f32x4 red[kSumGroups][kSumElems];
  const int64_t stride = static_cast<int64_t>(9) * Cout * Cin + Cout;
  const int el = threadIdx.x % kSumElems, q = threadIdx.x / kSumElems;
  const int64_t e = (static_cast<int64_t>(blockIdx.x) * kSumElems + el) * 4;
  f32x4 s = {0.0f, 0.0f, 0.0f, 0.0f};
  if (e < count) {
    const float* src = part + e;
    int k = q;
    for (; k + 7 * kSumGroups < splits; k += 8 * kSumGroups) {
      f32x4 a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = *reinterpret_cast<const f32x4*>(src + (k + u * kSumGroups) * stride);
#pragma unroll
      for (int u = 0; u < 8; ++u) s += a[u];
    }
    for (; k < splits; k += kSumGroups) s += *reinterpret_cast<const f32x4*>(src + k * stride);
  }
  red[q][el] = s;
  __syncthreads();
#pragma unroll
  for (int w = kSumGroups / 2; w > 0; w >>= 1) {
    if (q < w) red[q][el] += red[q + w][el];
    __syncthreads();
  }
  if (q != 0 || e >= count) return;
  s = red[0][el];
  const int64_t nw = static_cast<int64_t>(9) * Cout * Cin;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int64_t qq = e + j;
    if (qq >= nw) {
      db[qq - nw] = s[j];
    } else {
      const int ci = static_cast<int>(qq % Cin);
      const int co = static_cast<int>((qq / Cin) % Cout);
      const int tap = static_cast<int>(qq / (static_cast<int64_t>(Cin) * Cout));
      dw[(static_cast<int64_t>(co) * 9 + tap) * Cin + ci] = s[j];
    }
  }
}

// wt[tap][ci][co] = w[co][tap][ci]
__global__ __launch_bounds__(kThreads) void k_transpose_w(const __bf16* __restrict__ w, __bf16* __restrict__ wt,
                                                           int Cout, int Cin) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (e >= static_cast<int64_t>(9) * Cout * Cin) return;
  const int co = static_cast<int>(e % Cout);
  const int ci = static_cast<int>((e / Cout) % Cin);
  const int tap = static_cast<int>(e / (static_cast<int64_t>(Cout) * Cin));
  wt[e] = w[(static_cast<int64_t>(co) * 9 + tap) * Cin + ci];
}

int status() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? NDRR_OK : static_cast<int>(e);
}

int out_size(int in, int stride) { return (in - 1) / stride + 1; }

bool shape_ok(int N, int H, int W, int Cin, int Cout, int stride) {
  if (N <= 0 || H <= 0 || W <= 0 || Cin <= 0 || Cout <= 0 || (stride != 1 && stride != 2)) return false;
  if (Cout % 32 != 0 || (Cin != 1 && Cin % 32 != 0)) return false;
  return true;
}

// Output pixels per weight-grad workgroup: enough workgroups to fill the chip
// (about kWgradBlocks over the channel tiles), in whole 128-pixel steps.
int wgrad_blocks(int Cin) {
  static const int forced = [] {
    const char* v = getenv("NDRR_WGRAD_BLOCKS");
    const int b = v ? atoi(v) : 0;
    return b >= 64 && b <= 65536 ? b : 0;
  }();
  return forced ? forced : (Cin == 1 ? kWgradBlocksCin1 : kWgradBlocks);
}

// Input channels per weight-grad workgroup: 64 when they divide Cin, else 32.
int wgrad_ci(int Cin) { return Cin % 64 == 0 ? 64 : 32; }

int wgrad_split(int M, int Cin, int Cout) {
  const int tiles = Cin == 1 ? 1 : (Cin / wgrad_ci(Cin)) * (Cout / 64);
  const int want = (wgrad_blocks(Cin) + tiles - 1) / tiles;
  int per = (M + want - 1) / want;
  per = ((per + 127) / 128) * 128;
  return per < 256 ? 256 : per;
}

bool range_ok(int N, int H, int W, int C) {
  return static_cast<int64_t>(N) * H * W * C < (int64_t{1} << 31);
}

}  // namespace

extern "C" {

int ndrr_conv3x3_fwd(const void* x, const void* w, const float* bias, void* y, int N, int H, int W, int Cin,
                     int Cout, int stride, int relu, void* stream) {
  if (!x || !w || !y || !shape_ok(N, H, W, Cin, Cout, stride)) return NDRR_EINVAL;
  const int Ho = out_size(H, stride), Wo = out_size(W, stride);
  if (!range_ok(N, H, W, Cin) || !range_ok(N, Ho, Wo, Cout)) return NDRR_ERANGE;
  ConvP p{};
  p.x = static_cast<const __bf16*>(x); p.w = static_cast<const __bf16*>(w); p.bias = bias;
  p.y = static_cast<__bf16*>(y);
  p.N = N; p.H = H; p.W = W; p.Cin = Cin; p.Ho = Ho; p.Wo = Wo; p.Cout = Cout; p.S = stride; p.relu = relu != 0;
  const int M = N * Ho * Wo;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (Cin == 1) {
    const int per_block = 4 * kCin1Groups * kPixelsPerWave;
    const int blocks = (M + per_block - 1) / per_block;
    const bool wide = Cout % 64 == 0;
    const dim3 grid(blocks, Cout / (wide ? 64 : 32));
    if (stride == 1 && Wo % 32 == 0) {  // whole 32-pixel groups per row: the staged-strip kernel
      if (wide) hipLaunchKernelGGL((k_conv_fwd_rows<4>), grid, dim3(kThreads), 0, st, p);
      else hipLaunchKernelGGL((k_conv_fwd_rows<2>), grid, dim3(kThreads), 0, st, p);
    } else if (wide) {
      hipLaunchKernelGGL((k_conv_fwd<4>), grid, dim3(kThreads), 0, st, p);
    } else {
      hipLaunchKernelGGL((k_conv_fwd<2>), grid, dim3(kThreads), 0, st, p);
    }
  } else if (!try_halo<0>(p, st)) {
    launch_igemm<0>(p, M, Cout, 1, st);
  }
  return status();
}

int ndrr_conv3x3_transpose_weights(const void* w, void* wt, int Cin, int Cout, void* stream) {
  if (!w || !wt || Cin <= 0 || Cout <= 0) return NDRR_EINVAL;
  const int64_t n = static_cast<int64_t>(9) * Cin * Cout;
  if (n >= (int64_t{1} << 31)) return NDRR_ERANGE;
  hipLaunchKernelGGL(k_transpose_w, dim3(static_cast<unsigned>((n + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), static_cast<const __bf16*>(w), static_cast<__bf16*>(wt),
                     Cout, Cin);
  return status();
}

int ndrr_conv3x3_bwd_data(const void* dy, const void* wt, const void* x, void* dx, int N, int H, int W, int Cin,
                          int Cout, int stride, int relu_mask, void* stream) {
  if (!dy || !wt || !dx || !shape_ok(N, H, W, Cin, Cout, stride) || Cin % 32 != 0) return NDRR_EINVAL;
  if (relu_mask && !x) return NDRR_EINVAL;
  const int Ho = out_size(H, stride), Wo = out_size(W, stride);
  if (!range_ok(N, H, W, Cin) || !range_ok(N, Ho, Wo, Cout)) return NDRR_ERANGE;
  ConvP p{};
  p.dy = static_cast<const __bf16*>(dy); p.wt = static_cast<const __bf16*>(wt); p.x = static_cast<const __bf16*>(x);
  p.dx = static_cast<__bf16*>(dx);
  p.N = N; p.H = H; p.W = W; p.Cin = Cin; p.Ho = Ho; p.Wo = Wo; p.Cout = Cout; p.S = stride;
  p.mask = relu_mask != 0;
  hipStream_t st = static_cast<hipStream_t>(stream);
  // the staged-tile kernel (dy read once per workgroup) up to 128 output channels; above, the
  // tiled implicit GEMM's class-per-workgroup split balances the heavier products better
  if (stride == 2 && Cout <= 128) {
    const int Hc = (H + 1) / 2, Wc = (W + 1) / 2;
    const int64_t tiles = static_cast<int64_t>(N) * ((Hc + kDgTY - 1) / kDgTY) * ((Wc + kDgTX - 1) / kDgTX);
    if (tiles > 0x7fffffff) return NDRR_ERANGE;
    const size_t lds = static_cast<size_t>(kDgTile) * (Cout + 8) * sizeof(__bf16);
    if (Cin % 64 == 0)
      hipLaunchKernelGGL((k_dgrad_s2<4>), dim3(static_cast<unsigned>(tiles), Cin / 64), dim3(kThreads), lds, st, p);
    else
      hipLaunchKernelGGL((k_dgrad_s2<2>), dim3(static_cast<unsigned>(tiles), Cin / 32), dim3(kThreads), lds, st, p);
    return status();
  }
  // the halo-staged kernel for stride 1 up to 64 output channels (each chunk of dy read once per
  // tile); beyond, the tiled implicit GEMM, whose weights staged in LDS serve four waves at once,
  // measured faster
  if (try_halo<1>(p, st)) return status();
  if (stride == 1 && Cout <= 64) {
    launch_dgrad_halo_s1(p, st);
    return status();
  }
  // pixels per class: all of them (stride 1), or the largest parity class (stride 2)
  const int M = stride == 2 ? N * ((H + 1) / 2) * ((W + 1) / 2) : N * H * W;
  launch_igemm<1>(p, M, Cin, stride == 2 ? 4 : 1, st);
  return status();
}

size_t ndrr_conv3x3_bwd_weight_workspace_floats(int N, int H, int W, int Cin, int Cout, int stride) {
  if (!shape_ok(N, H, W, Cin, Cout, stride)) return 0;
  const int64_t M = static_cast<int64_t>(N) * out_size(H, stride) * out_size(W, stride);
  if (M >= (int64_t{1} << 31)) return 0;
  const int per = wgrad_split(static_cast<int>(M), Cin, Cout);
  const int64_t splits = (M + per - 1) / per;
  return static_cast<size_t>(splits * (9 * Cout * Cin + Cout));
}

int ndrr_conv3x3_bwd_weight(const void* dy, const void* x, float* workspace, float* dw, float* dbias, int N, int H,
                            int W, int Cin, int Cout, int stride, void* stream) {
  if (!dy || !x || !workspace || !dw || !shape_ok(N, H, W, Cin, Cout, stride)) return NDRR_EINVAL;
  if ((reinterpret_cast<uintptr_t>(workspace) & 15) != 0) return NDRR_EINVAL;
  if (Cin == 1 ? Cout != 32 : Cout % 64 != 0) return NDRR_EINVAL;
  const int Ho = out_size(H, stride), Wo = out_size(W, stride);
  if (!range_ok(N, H, W, Cin) || !range_ok(N, Ho, Wo, Cout)) return NDRR_ERANGE;
  ConvP p{};
  p.dy = static_cast<const __bf16*>(dy); p.x = static_cast<const __bf16*>(x); p.part = workspace;
  p.N = N; p.H = H; p.W = W; p.Cin = Cin; p.Ho = Ho; p.Wo = Wo; p.Cout = Cout; p.S = stride;
  p.bias_grad = dbias != nullptr;
  const int M = N * Ho * Wo;
  p.split = wgrad_split(M, Cin, Cout);
  const int splits = (M + p.split - 1) / p.split;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (Cin == 1) {
    if (stride == 1 && Wo % 32 == 0)  // whole 32-pixel groups per row: the staged-strip kernel
      hipLaunchKernelGGL(k_conv_bwd_weight_cin1_rows, dim3(splits), dim3(kThreads), 0, st, p);
    else
      hipLaunchKernelGGL(k_conv_bwd_weight_cin1, dim3(splits), dim3(kThreads), 0, st, p);
  } else {
    const dim3 grid(splits, Cin / wgrad_ci(Cin), Cout / 64);
    if (wgrad_ci(Cin) == 64) hipLaunchKernelGGL((k_conv_bwd_weight<4, NDRR_WGRAD_BUFFERS>), grid, dim3(kThreads), 0, st, p);
    else hipLaunchKernelGGL((k_conv_bwd_weight<2, NDRR_WGRAD_BUFFERS>), grid, dim3(kThreads), 0, st, p);
  }
  int r = status();
  if (r != NDRR_OK) return r;
  const int64_t count = static_cast<int64_t>(9) * Cout * Cin + (dbias ? Cout : 0);
  const int64_t groups4 = (count + 4 * kSumElems - 1) / (4 * kSumElems);
  hipLaunchKernelGGL(k_sum_splits, dim3(static_cast<unsigned>(groups4)), dim3(kThreads), 0, st, workspace, dw, dbias,
                     splits, Cout, Cin, count);
  return status();
}

}  // extern "C"
