// The pose-proposal CNN's head on AMD Instinct MI355X (gfx950): 2 x 2 average
// pooling of the last feature map, a fully connected layer with ReLU and a
// small output layer, forward and backward, on the bf16 matrix cores
// (v_mfma_f32_16x16x32_bf16) with fp32 accumulation.
//
//   P[n][k]   = mean of feat[n][2py + dy][2px + dx][c] over dy, dx in {0, 1},
//               k = (py PW + px) C + c                                    (bf16)
//   H[n][j]   = relu(b1[j] + sum_k P[n][k] W1[j][k])                       (fp32)
//   out[n][o] = b2[o] + sum_j H[n][j] W2[o][j]                             (fp32)
// and, for an upstream gradient dout[n][o],
//   dH[n][j]  = [H[n][j] > 0] sum_o dout[n][o] W2[o][j]
//   dW1[j][k] = sum_n dH[n][j] P[n][k],   db1[j] = sum_n dH[n][j]
//   dW2[o][j] = sum_n dout[n][o] H[n][j], db2[o] = sum_n dout[n][o]
//   dfeat[n][y][x][c] = [feat[n][y][x][c] > 0] dH W1 [n][k(y/2, x/2, c)] / 4
// (the mask is the ReLU of the layer that produced feat).
//
// The three products are one kernel, C[m][c] = sum_k A[m][k] B[c][k] with both
// operands contiguous along k (k_head_gemm): the forward with A = P,
// B = W1 (split over k, partials added in a fixed order by k_head_fwd_finish);
// the weight gradient with A = dH^T, B = P^T (k = the batch, zero-padded to a
// multiple of 32); the input gradient with A = W1^T, B = dH, whose epilogue
// spreads each value over its pooling window and applies the ReLU mask.
// Every sum has a fixed order: results are bitwise reproducible.
//
// MFMA operand layout (16x16x32 bf16, wave64): lane l holds A[l % 16][8 (l / 16)
// + j] and B[8 (l / 16) + j][l % 16], j < 8, and D[4 (l / 16) + r][l % 16] in
// its four accumulators r (the layout the convolution kernels use).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nanodrr_amd.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kThreads = 256;
constexpr int kTileM = 32;   // rows per wave (two 16-row MFMA groups)
constexpr int kTileN = 64;   // columns per wave (four 16-column MFMA groups)
constexpr int kSplitK = 32;  // k slices of the forward product
constexpr int kPoolTile = 64;

__device__ inline f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int j = 0; j < 8; ++j) z[j] = static_cast<__bf16>(0.0f);
  return z;
}

int status() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? NDRR_OK : static_cast<int>(e);
}

// ---------------------------------------------------------------------------
// fp32 master weights -> bf16 in both orientations: w[J][K] and wt[K][J].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_head_cast(const float* __restrict__ w, __bf16* __restrict__ wb,
                                                        __bf16* __restrict__ wt, int J, int K) {
  __shared__ __bf16 tile[64][64 + 2];
  const int k0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += kThreads) {
    const int jj = e / 64, kk = e % 64;
    const int j = j0 + jj, k = k0 + kk;
    const __bf16 v = (j < J && k < K) ? static_cast<__bf16>(w[static_cast<int64_t>(j) * K + k])
                                      : static_cast<__bf16>(0.0f);
    if (j < J && k < K) wb[static_cast<int64_t>(j) * K + k] = v;
    tile[jj][kk] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += kThreads) {
    const int kk = e / 64, jj = e % 64;
    const int j = j0 + jj, k = k0 + kk;
    if (j < J && k < K) wt[static_cast<int64_t>(k) * J + j] = tile[jj][kk];
  }
}

// ---------------------------------------------------------------------------
// 2 x 2 average pooling: P[n][k] and its transpose Pt[k][np] (np < Np, zero for
// n >= N). A workgroup takes 64 values of k (64 channels of one pooled cell)
// for 64 samples; each thread pools 8 consecutive channels of one sample (four
// 16-byte loads), stores them to P with one 16-byte store and through LDS to
// Pt as 64-sample rows.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_head_pool(const __bf16* __restrict__ feat, __bf16* __restrict__ P,
                                                        __bf16* __restrict__ Pt, int N, int Np, int FH, int FW,
                                                        int C) {
  __shared__ __bf16 tile[kPoolTile][kPoolTile + 8];  // [k][n]
  const int PW = FW / 2;
  const int K = (FH / 2) * PW * C;
  const int k0 = blockIdx.x * kPoolTile, n0 = blockIdx.y * kPoolTile;
  const int cell = k0 / C, c0 = k0 % C;
  const int py = cell / PW, px = cell % PW;
  for (int item = threadIdx.x; item < kPoolTile * (kPoolTile / 8); item += kThreads) {
    const int nl = item / (kPoolTile / 8), q = item % (kPoolTile / 8);
    const int n = n0 + nl;
    float s[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = 0.0f;
    if (n < N) {
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const int y = 2 * py + (d >> 1), x = 2 * px + (d & 1);
        const bf16x8 v = *reinterpret_cast<const bf16x8*>(
            feat + ((static_cast<int64_t>(n) * FH + y) * FW + x) * C + c0 + 8 * q);
#pragma unroll
        for (int j = 0; j < 8; ++j) s[j] += static_cast<float>(v[j]);
      }
    }
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = static_cast<__bf16>(0.25f * s[j]);
    if (n < N) *reinterpret_cast<bf16x8*>(P + static_cast<int64_t>(n) * K + k0 + 8 * q) = o;
#pragma unroll
    for (int j = 0; j < 8; ++j) tile[8 * q + j][nl] = o[j];
  }
  __syncthreads();
  for (int item = threadIdx.x; item < kPoolTile * (kPoolTile / 8); item += kThreads) {
    const int kl = item / (kPoolTile / 8), q = item % (kPoolTile / 8);
    const int n = n0 + 8 * q;
    if (n >= Np) continue;  // Np is a multiple of 32: a piece lies wholly inside or outside
    bf16x8 o;
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = tile[kl][8 * q + j];
    *reinterpret_cast<bf16x8*>(Pt + static_cast<int64_t>(k0 + kl) * Np + n) = o;
  }
}

// ---------------------------------------------------------------------------
// C[m][c] = sum over k in this workgroup's slice of A[m][k] B[c][k]. A wave
// computes a 32 x 64 tile (2 x 4 MFMA tiles), four waves a 64 x 128 block;
// blockIdx.z picks the k slice. Rows beyond M and columns beyond Ncol read
// zeros and are not written. K (and the slice) is a multiple of 32; lda, ldb
// multiples of 8. The next k step's fragments are loaded before this step's
// MFMAs.
// EPI 0: C to out[z][m][c] (fp32, row length Ncol).
// EPI 1: the input gradient's epilogue: row m = k of the pooled features, column
//        c = sample n; the value / 4 goes to the pooling window's four pixels of
//        dfeat (bf16), times [feat > 0].
// ---------------------------------------------------------------------------
struct GemmP {
  const __bf16* A;
  const __bf16* B;
  int M, Ncol, K, lda, ldb, kslice;
  float* out;
  // EPI 1
  const __bf16* feat;
  __bf16* dfeat;
  int FH, FW, C;
};

template <int EPI>
__global__ __launch_bounds__(kThreads) void k_head_gemm(GemmP p) {
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int m0 = blockIdx.y * (2 * kTileM) + (wave >> 1) * kTileM;
  const int c0 = blockIdx.x * (2 * kTileN) + (wave & 1) * kTileN;
  const int kbeg = blockIdx.z * p.kslice;
  const int kend = min(p.K, kbeg + p.kslice);
  const int kq = 8 * (l >> 4);
  const __bf16* arow[2];
  bool aok[2];
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const int m = m0 + 16 * g + (l & 15);
    aok[g] = m < p.M;
    arow[g] = p.A + static_cast<int64_t>(aok[g] ? m : 0) * p.lda + kq;
  }
  const __bf16* brow[4];
  bool bok[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int c = c0 + 16 * g + (l & 15);
    bok[g] = c < p.Ncol;
    brow[g] = p.B + static_cast<int64_t>(bok[g] ? c : 0) * p.ldb + kq;
  }
  f32x4 acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  bf16x8 a[2], b[4];
  auto load = [&](int k) {
#pragma unroll
    for (int g = 0; g < 2; ++g) a[g] = aok[g] ? *reinterpret_cast<const bf16x8*>(arow[g] + k) : zero8();
#pragma unroll
    for (int g = 0; g < 4; ++g) b[g] = bok[g] ? *reinterpret_cast<const bf16x8*>(brow[g] + k) : zero8();
  };
  if (kbeg < kend) load(kbeg);
  for (int k = kbeg; k < kend; k += 32) {
    bf16x8 ca[2] = {a[0], a[1]};
    bf16x8 cb[4] = {b[0], b[1], b[2], b[3]};
    if (k + 32 < kend) load(k + 32);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma(ca[i], cb[j], acc[i][j]);
  }
  if constexpr (EPI == 0) {
    float* o = p.out + static_cast<int64_t>(blockIdx.z) * p.M * p.Ncol;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = c0 + 16 * j + (l & 15);
        if (c >= p.Ncol) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int m = m0 + 16 * i + 4 * (l >> 4) + r;
          if (m < p.M) o[static_cast<int64_t>(m) * p.Ncol + c] = acc[i][j][r];
        }
      }
  } else {
    const int PW = p.FW / 2;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int k = m0 + 16 * i + 4 * (l >> 4);  // four consecutive channels k .. k + 3 of one cell
      if (k >= p.M) continue;
      const int cell = k / p.C, ch = k % p.C;
      const int py = cell / PW, px = cell % PW;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int n = c0 + 16 * j + (l & 15);
        if (n >= p.Ncol) continue;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
          const int y = 2 * py + (d >> 1), x = 2 * px + (d & 1);
          const int64_t at = ((static_cast<int64_t>(n) * p.FH + y) * p.FW + x) * p.C + ch;
          const bf16x4 f = *reinterpret_cast<const bf16x4*>(p.feat + at);
          bf16x4 g;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            g[r] = static_cast<__bf16>(static_cast<float>(f[r]) > 0.0f ? 0.25f * acc[i][j][r] : 0.0f);
          *reinterpret_cast<bf16x4*>(p.dfeat + at) = g;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Forward finish, one workgroup per sample: thread j adds the k-slice partials
// of H[n][j] in slice order, adds the bias, applies ReLU and stores H; the six
// outputs are then block sums of H W2 (a butterfly in each wave, then the
// waves in order).
// ---------------------------------------------------------------------------
template <int J>
__global__ __launch_bounds__(J) void k_head_fwd_finish(const float* __restrict__ part, int splits, int N,
                                                       const float* __restrict__ b1, const float* __restrict__ w2,
                                                       const float* __restrict__ b2, int O, float* __restrict__ H,
                                                       float* __restrict__ out) {
  __shared__ float red[J / 64];
  const int n = blockIdx.x, j = threadIdx.x;
  float h = b1[j];
  for (int z = 0; z < splits; ++z) h += part[(static_cast<int64_t>(z) * N + n) * J + j];
  h = fmaxf(h, 0.0f);
  H[static_cast<int64_t>(n) * J + j] = h;
  for (int o = 0; o < O; ++o) {
    float v = h * w2[o * J + j];
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s, 64);
    __syncthreads();
    if ((j & 63) == 0) red[j >> 6] = v;
    __syncthreads();
    if (j == 0) {
      float t = b2[o];
#pragma unroll
      for (int w = 0; w < J / 64; ++w) t += red[w];
      out[static_cast<int64_t>(n) * O + o] = t;
    }
  }
}

// ---------------------------------------------------------------------------
// Backward through the output layer and the ReLU, one workgroup per (padded)
// sample: dH[n][j] = [H > 0] sum_o dout[n][o] W2[o][j], stored as fp32 and as
// bf16 in both orientations (dHt[j][np] with zeros for the padding samples).
// ---------------------------------------------------------------------------
template <int J>
__global__ __launch_bounds__(J) void k_head_bwd_hidden(const float* __restrict__ dout, const float* __restrict__ H,
                                                       const float* __restrict__ w2, int N, int Np, int O,
                                                       float* __restrict__ dh, __bf16* __restrict__ dHb,
                                                       __bf16* __restrict__ dHt) {
  const int n = blockIdx.x, j = threadIdx.x;
  float g = 0.0f;
  if (n < N) {
    for (int o = 0; o < O; ++o) g += dout[static_cast<int64_t>(n) * O + o] * w2[o * J + j];
    g = H[static_cast<int64_t>(n) * J + j] > 0.0f ? g : 0.0f;
    dh[static_cast<int64_t>(n) * J + j] = g;
    dHb[static_cast<int64_t>(n) * J + j] = static_cast<__bf16>(g);
  }
  dHt[static_cast<int64_t>(j) * Np + n] = static_cast<__bf16>(g);
}

// Bias and output-layer gradients, one workgroup per hidden unit j: thread t
// sums samples t, t + 256, ... of dh[n][j] and dout[n][o] H[n][j], then the
// block adds the thread sums in a fixed tree; workgroup 0 also forms db2.
constexpr int kSmallThreads = 256;
constexpr int kMaxOutputs = 8;

__global__ __launch_bounds__(kSmallThreads) void k_head_bwd_small(const float* __restrict__ dout,
                                                                  const float* __restrict__ H,
                                                                  const float* __restrict__ dh, int N, int J, int O,
                                                                  float* __restrict__ db1, float* __restrict__ dw2,
                                                                  float* __restrict__ db2) {
  __shared__ float red[1 + 2 * kMaxOutputs][kSmallThreads];
  const int j = blockIdx.x, t = threadIdx.x;
  float s[1 + 2 * kMaxOutputs];
#pragma unroll
  for (int q = 0; q < 1 + 2 * kMaxOutputs; ++q) s[q] = 0.0f;
  for (int n = t; n < N; n += kSmallThreads) {
    const float h = H[static_cast<int64_t>(n) * J + j];
    s[0] += dh[static_cast<int64_t>(n) * J + j];
#pragma unroll
    for (int o = 0; o < kMaxOutputs; ++o) {
      if (o < O) {
        const float d = dout[static_cast<int64_t>(n) * O + o];
        s[1 + o] += d * h;
        s[1 + kMaxOutputs + o] += d;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 1 + 2 * kMaxOutputs; ++q) red[q][t] = s[q];
  __syncthreads();
  for (int w = kSmallThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int q = 0; q < 1 + 2 * kMaxOutputs; ++q) red[q][t] += red[q][t + w];
    }
    __syncthreads();
  }
  if (t == 0) db1[j] = red[0][0];
  if (t < O) {
    dw2[t * J + j] = red[1 + t][0];
    if (j == 0) db2[t] = red[1 + kMaxOutputs + t][0];
  }
}

constexpr int kHidden = 256;  // J: the hidden width the kernels are built for

bool head_shape_ok(int N, int FH, int FW, int C, int J, int O) {
  return N > 0 && FH > 0 && FW > 0 && FH % 2 == 0 && FW % 2 == 0 && C > 0 && C % kPoolTile == 0 &&
         J == kHidden && O > 0 && O <= kMaxOutputs;
}

int padded(int N) { return (N + 31) / 32 * 32; }

}  // namespace

extern "C" {

size_t ndrr_head_workspace_floats(int N, int FH, int FW, int C, int J) {
  if (N <= 0 || FH <= 0 || FW <= 0 || C <= 0 || J <= 0) return 0;
  return static_cast<size_t>(kSplitK) * N * J;
}

int ndrr_head_cast_weights(const float* w1, void* w1_bf16, void* w1t_bf16, int J, int K, void* stream) {
  if (!w1 || !w1_bf16 || !w1t_bf16 || J <= 0 || K <= 0) return NDRR_EINVAL;
  if (static_cast<int64_t>(J) * K >= (int64_t{1} << 31)) return NDRR_ERANGE;
  hipLaunchKernelGGL(k_head_cast, dim3((K + 63) / 64, (J + 63) / 64), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), w1, static_cast<__bf16*>(w1_bf16),
                     static_cast<__bf16*>(w1t_bf16), J, K);
  return status();
}

int ndrr_head_pool(const void* feat, void* P, void* Pt, int N, int FH, int FW, int C, void* stream) {
  if (!feat || !P || !Pt || !head_shape_ok(N, FH, FW, C, kHidden, 1)) return NDRR_EINVAL;
  const int K = (FH / 2) * (FW / 2) * C;
  const int Np = padded(N);
  if (static_cast<int64_t>(N) * FH * FW * C >= (int64_t{1} << 31) || static_cast<int64_t>(K) * Np >= (int64_t{1} << 31))
    return NDRR_ERANGE;
  hipLaunchKernelGGL(k_head_pool, dim3(K / kPoolTile, (Np + kPoolTile - 1) / kPoolTile), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), static_cast<const __bf16*>(feat), static_cast<__bf16*>(P),
                     static_cast<__bf16*>(Pt), N, Np, FH, FW, C);
  return status();
}

int ndrr_head_fwd(const void* P, const void* w1_bf16, const float* b1, const float* w2, const float* b2,
                  float* workspace, float* H, float* out, int N, int K, int J, int O, void* stream) {
  if (!P || !w1_bf16 || !b1 || !w2 || !b2 || !workspace || !H || !out) return NDRR_EINVAL;
  if (N <= 0 || K <= 0 || K % (32 * kSplitK) != 0 || J != kHidden || O <= 0 || O > kMaxOutputs) return NDRR_EINVAL;
  if (static_cast<int64_t>(N) * K >= (int64_t{1} << 31) || N > 65535 * 2 * kTileM) return NDRR_ERANGE;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  GemmP g{};
  g.A = static_cast<const __bf16*>(P);
  g.B = static_cast<const __bf16*>(w1_bf16);
  g.M = N; g.Ncol = J; g.K = K; g.lda = K; g.ldb = K; g.kslice = K / kSplitK;
  g.out = workspace;
  const dim3 grid((J + 2 * kTileN - 1) / (2 * kTileN), (N + 2 * kTileM - 1) / (2 * kTileM), kSplitK);
  hipLaunchKernelGGL(k_head_gemm<0>, grid, dim3(kThreads), 0, st, g);
  hipLaunchKernelGGL(k_head_fwd_finish<kHidden>, dim3(N), dim3(kHidden), 0, st, workspace, kSplitK, N, b1, w2, b2, O,
                     H, out);
  return status();
}

int ndrr_head_bwd(const float* dout, const float* H, const float* w2, const void* w1t_bf16, const void* Pt,
                  const void* feat, float* dh, void* dH_bf16, void* dHt_bf16, float* dw1, float* db1, float* dw2,
                  float* db2, void* dfeat, int N, int FH, int FW, int C, int J, int O, void* stream) {
  if (!dout || !H || !w2 || !w1t_bf16 || !Pt || !feat || !dh || !dH_bf16 || !dHt_bf16 || !dw1 || !db1 || !dw2 ||
      !db2 || !dfeat || !head_shape_ok(N, FH, FW, C, J, O))
    return NDRR_EINVAL;
  const int K = (FH / 2) * (FW / 2) * C;
  const int Np = padded(N);
  if (static_cast<int64_t>(N) * FH * FW * C >= (int64_t{1} << 31) || static_cast<int64_t>(K) * Np >= (int64_t{1} << 31))
    return NDRR_ERANGE;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_head_bwd_hidden<kHidden>, dim3(Np), dim3(kHidden), 0, st, dout, H, w2, N, Np, O, dh,
                     static_cast<__bf16*>(dH_bf16), static_cast<__bf16*>(dHt_bf16));
  hipLaunchKernelGGL(k_head_bwd_small, dim3(J), dim3(kSmallThreads), 0, st, dout, H, dh, N, J, O, db1, dw2, db2);
  // dW1[j][k] = sum_n dHt[j][n] Pt[k][n]
  GemmP w{};
  w.A = static_cast<const __bf16*>(dHt_bf16);
  w.B = static_cast<const __bf16*>(Pt);
  w.M = J; w.Ncol = K; w.K = Np; w.lda = Np; w.ldb = Np; w.kslice = Np;
  w.out = dw1;
  hipLaunchKernelGGL(k_head_gemm<0>, dim3((K + 2 * kTileN - 1) / (2 * kTileN), (J + 2 * kTileM - 1) / (2 * kTileM), 1),
                     dim3(kThreads), 0, st, w);
  // dP^T[k][n] = sum_j W1t[k][j] dH[n][j], spread over the pooling windows with the ReLU mask
  GemmP d{};
  d.A = static_cast<const __bf16*>(w1t_bf16);
  d.B = static_cast<const __bf16*>(dH_bf16);
  d.M = K; d.Ncol = N; d.K = J; d.lda = J; d.ldb = J; d.kslice = J;
  d.feat = static_cast<const __bf16*>(feat);
  d.dfeat = static_cast<__bf16*>(dfeat);
  d.FH = FH; d.FW = FW; d.C = C;
  hipLaunchKernelGGL(k_head_gemm<1>, dim3((N + 2 * kTileN - 1) / (2 * kTileN), (K + 2 * kTileM - 1) / (2 * kTileM), 1),
                     dim3(kThreads), 0, st, d);
  return status();
}

}  // extern "C"
