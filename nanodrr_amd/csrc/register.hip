// Registration support kernels for AMD Instinct MI355X (gfx950):
//   * the SE(3) pose step, in FP64: exponential of the per-problem rotation
//     vector, composition with the fixed camera/pivot transforms, and its
//     analytic adjoint back to the (rotation, translation) parameters;
//   * zero-normalised cross-correlation of each rendered image with its
//     target, with the gradient: chunk moments over (chunk, image)
//     workgroups merged in a fixed order (deterministic);
//   * local (patch) ZNCC over every p x p window, with its gradient;
//   * the Adam update of the registration parameters, and the device-side
//     loop state (iteration counter, jitter offsets, schedule table) that lets
//     several iterations be captured in one HIP graph.
//
// The parameterisation follows nanodrr's Registration module
// (src/nanodrr/registration/registration.py:64-68):
//   pose = rt_inv @ pivot @ T(w, v) @ pivot_inv,  T = [R(w), R(w) v; 0, 1]
// with R(w) = exp([w]x) (convert(..., "so3_log"), src/nanodrr/geometry/se3.py:62-69,109-110)
// and the ZNCC of src/nanodrr/metrics/ncc.py:8-37 (population variance + eps).

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nanodrr_amd.h"
#include "se3_pose.h"

namespace {

constexpr int kThreads = 256;

// M = (prefix @ T @ suffix)[:3, :] with T homogeneous (last row 0 0 0 1).
__global__ void k_pose_compose(const double* __restrict__ prefix, const double* __restrict__ suffix,
                               const float* __restrict__ rot, const float* __restrict__ xyz, float* __restrict__ M,
                               int B, int64_t param_stride) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double w[3] = {rot[b * param_stride], rot[b * param_stride + 1], rot[b * param_stride + 2]};
  const double v[3] = {xyz[b * param_stride], xyz[b * param_stride + 1], xyz[b * param_stride + 2]};
  ndrr_se3::compose(prefix + 16 * static_cast<int64_t>(b), suffix + 16 * static_cast<int64_t>(b), w, v,
                    M + 12 * static_cast<int64_t>(b));
}

// Adjoint of k_pose_compose: gM (3x4 per problem) -> dL/dw, dL/dv.
__global__ void k_pose_compose_bwd(const double* __restrict__ prefix, const double* __restrict__ suffix,
                                   const float* __restrict__ rot, const float* __restrict__ xyz,
                                   const float* __restrict__ gM, float* __restrict__ grot, float* __restrict__ gxyz,
                                   int B, int64_t param_stride) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double w[3] = {rot[b * param_stride], rot[b * param_stride + 1], rot[b * param_stride + 2]};
  const double v[3] = {xyz[b * param_stride], xyz[b * param_stride + 1], xyz[b * param_stride + 2]};
  double G[3][4];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) G[i][j] = gM[12 * static_cast<int64_t>(b) + 4 * i + j];
  ndrr_se3::compose_adjoint(prefix + 16 * static_cast<int64_t>(b), suffix + 16 * static_cast<int64_t>(b), w, v, G,
                            grot + b * param_stride, gxyz + b * param_stride);
}

// ZNCC per image: xh = (x - mean x) / sqrt(var x + eps), yh likewise,
// score = mean(xh * yh); grad wrt x = (yh - score xh) / (N sigma_x).
//
// Two launches over (chunk, image) grids so that a batch of a few images
// still fills the chip. k_zncc_part reduces a chunk of kZnccChunk pixels to
// its moments (count, means, centred second moments, co-moment), two passes
// over values held in registers. k_zncc_finish merges an image's chunk
// moments (Chan, Golub and LeVeque's pairwise update, in a fixed order) and,
// per chunk, writes the gradient. Every sum has a fixed order, so results are
// bitwise reproducible.
constexpr int kZnccPer = 16;
constexpr int kZnccChunk = kThreads * kZnccPer;
constexpr int kZnccStat = 8;  // floats per chunk record (6 used)

struct Moments {
  float n, mx, my, sxx, syy, sxy;
};

__device__ inline Moments merge(const Moments& a, const Moments& b) {
  if (b.n == 0.0f) return a;
  if (a.n == 0.0f) return b;
  const float n = a.n + b.n;
  const float dx = b.mx - a.mx, dy = b.my - a.my;
  const float f = a.n * b.n / n;
  return {n,
          fmaf(dx, b.n / n, a.mx),
          fmaf(dy, b.n / n, a.my),
          a.sxx + b.sxx + dx * dx * f,
          a.syy + b.syy + dy * dy * f,
          a.sxy + b.sxy + dx * dy * f};
}

// Sum over the block: a butterfly within each wave, then the four wave sums
// in order. Every thread returns the same value.
__device__ inline float block_sum(float v, float* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6;
  __syncthreads();  // red may still be read by a previous call
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(kThreads) void k_zncc_part(const float* __restrict__ x, const float* __restrict__ y,
                                                        float* __restrict__ part, int N) {
  __shared__ float red[kThreads / 64];
  const int b = blockIdx.y, p = blockIdx.x, P = gridDim.x;
  const int base = p * kZnccChunk;
  const float* xb = x + static_cast<int64_t>(b) * N;
  const float* yb = y + static_cast<int64_t>(b) * N;
  const int n = min(kZnccChunk, N - base);
  float xv[kZnccPer], yv[kZnccPer];
  float sx = 0.0f, sy = 0.0f;
#pragma unroll
  for (int j = 0; j < kZnccPer; ++j) {
    const int e = base + j * kThreads + threadIdx.x;
    xv[j] = e < N ? xb[e] : 0.0f;
    yv[j] = e < N ? yb[e] : 0.0f;
    sx += xv[j];
    sy += yv[j];
  }
  const float inv_n = 1.0f / static_cast<float>(n);
  const float mx = block_sum(sx, red) * inv_n;
  const float my = block_sum(sy, red) * inv_n;
  float vx = 0.0f, vy = 0.0f, cxy = 0.0f;
#pragma unroll
  for (int j = 0; j < kZnccPer; ++j) {
    if (base + j * kThreads + static_cast<int>(threadIdx.x) < N) {
      const float dx = xv[j] - mx, dy = yv[j] - my;
      vx = fmaf(dx, dx, vx);
      vy = fmaf(dy, dy, vy);
      cxy = fmaf(dx, dy, cxy);
    }
  }
  vx = block_sum(vx, red);
  vy = block_sum(vy, red);
  cxy = block_sum(cxy, red);
  if (threadIdx.x == 0) {
    float* r = part + (static_cast<int64_t>(b) * P + p) * kZnccStat;
    r[0] = static_cast<float>(n);
    r[1] = mx;
    r[2] = my;
    r[3] = vx;
    r[4] = vy;
    r[5] = cxy;
  }
}

__global__ __launch_bounds__(kThreads) void k_zncc_finish(const float* __restrict__ x, const float* __restrict__ y,
                                                          const float* __restrict__ part, int P,
                                                          float* __restrict__ score, float* __restrict__ gx, int N,
                                                          float eps) {
  __shared__ float st[5];
  const int b = blockIdx.y, p = blockIdx.x;
  if (threadIdx.x < 64) {
    // lane l merges chunks l, l + 64, ... in order, then a fixed tree over the lanes
    Moments m{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (int q = threadIdx.x; q < P; q += 64) {
      const float* r = part + (static_cast<int64_t>(b) * P + q) * kZnccStat;
      m = merge(m, Moments{r[0], r[1], r[2], r[3], r[4], r[5]});
    }
    for (int o = 32; o > 0; o >>= 1) {
      const Moments u{__shfl_down(m.n, o), __shfl_down(m.mx, o), __shfl_down(m.my, o),
                      __shfl_down(m.sxx, o), __shfl_down(m.syy, o), __shfl_down(m.sxy, o)};
      if (static_cast<int>(threadIdx.x) + o < 64) m = merge(m, u);
    }
    if (threadIdx.x == 0) {
      const float inv_n = 1.0f / static_cast<float>(N);
      const float sdx = sqrtf(m.sxx * inv_n + eps), sdy = sqrtf(m.syy * inv_n + eps);
      const float sc = m.sxy * inv_n / (sdx * sdy);
      st[0] = m.mx;
      st[1] = m.my;
      st[2] = sdx;
      st[3] = sdy;
      st[4] = sc;
      if (p == 0) score[b] = sc;
    }
  }
  __syncthreads();
  if (!gx) return;
  const float mx = st[0], my = st[1], sdx = st[2], sdy = st[3], sc = st[4];
  const float g = 1.0f / (static_cast<float>(N) * sdx);
  const float* xb = x + static_cast<int64_t>(b) * N;
  const float* yb = y + static_cast<int64_t>(b) * N;
  float* gb = gx + static_cast<int64_t>(b) * N;
  const int base = p * kZnccChunk;
#pragma unroll 4
  for (int j = 0; j < kZnccPer; ++j) {
    const int e = base + j * kThreads + threadIdx.x;
    if (e < N) {
      const float xh = (xb[e] - mx) / sdx, yh = (yb[e] - my) / sdy;
      gb[e] = (yh - sc * xh) * g;
    }
  }
}

// One Adam step in place over n parameters: m = b1 m + (1 - b1) g,
// v = b2 v + (1 - b2) g^2, p -= lr (m / bc1) / (sqrt(v / bc2) + eps), where
// bc1 = 1 - b1^t and bc2 = 1 - b2^t come from the host with the step's lr.
__global__ __launch_bounds__(kThreads) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, int n, float lr,
                                                   float b1, float c1, float b2, float c2, float eps, float bc1,
                                                   float bc2) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = m[i] * b1 + c1 * gi;  // c = 1 - b, rounded from float64 (1 - 0.999f is 1.3e-5 off)
  const float vi = v[i] * b2 + c2 * (gi * gi);
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// ---------------------------------------------------------------------------
// Local (patch) ZNCC: the mean over every p x p window (stride 1) of the
// window's ZNCC, as nanodrr's patch NCC (src/nanodrr/metrics/ncc.py:21-24,
// 83-88: each unfolded window is standardised on its own, with the
// population variance + eps), and its gradient.
//
// k_patch_windows: a workgroup takes a 16 x 16 block of windows and stages the
// (16 + p - 1)^2 pixels they cover, of x and y, in LDS; each thread forms its
// window's means, then its centred moments (two passes over LDS), its score s
// and the window's gradient terms alpha = 1 / (P n sx sy) and
// beta = s / (P n sx^2) (n = p^2 pixels, P windows), written with the means as
// one float4 per window; the block's sum of s goes to a partial.
// k_patch_pixels: per 16 x 16 pixel block, the windows covering it are staged
// in LDS (zeros outside the window grid) and each pixel sums
//   dS/dx = sum over its windows of alpha (y - my) - beta (x - mx),
// the derivative of S = (1/P) sum s; block 0 of each image also adds the
// score partials in a fixed order. Deterministic.
// ---------------------------------------------------------------------------
constexpr int kPatchTile = 16;
constexpr int kPatchMax = 32;
constexpr int kPatchSpan = kPatchTile + kPatchMax - 1;

__global__ __launch_bounds__(kThreads) void k_patch_windows(const float* __restrict__ x, const float* __restrict__ y,
                                                            int H, int W, int p, float eps, float inv_pn,
                                                            float4* __restrict__ coef, float* __restrict__ part) {
  __shared__ float sx[kPatchSpan][kPatchSpan + 1];
  __shared__ float sy[kPatchSpan][kPatchSpan + 1];
  __shared__ float red[kThreads / 64];
  const int Ho = H - p + 1, Wo = W - p + 1;
  const int tiles_w = (Wo + kPatchTile - 1) / kPatchTile;
  const int b = blockIdx.y;
  const int ty = blockIdx.x / tiles_w, tx = blockIdx.x - ty * tiles_w;
  const int y0 = ty * kPatchTile, x0 = tx * kPatchTile;
  const int span = kPatchTile + p - 1;
  const float* xb = x + static_cast<int64_t>(b) * H * W;
  const float* yb = y + static_cast<int64_t>(b) * H * W;
  for (int e = threadIdx.x; e < span * span; e += kThreads) {
    const int r = e / span, c = e - r * span;
    const int gy = y0 + r, gx = x0 + c;
    const bool in = gy < H && gx < W;
    sx[r][c] = in ? xb[gy * W + gx] : 0.0f;
    sy[r][c] = in ? yb[gy * W + gx] : 0.0f;
  }
  __syncthreads();
  const int ly = threadIdx.x / kPatchTile, lx = threadIdx.x % kPatchTile;
  const int wy = y0 + ly, wx = x0 + lx;
  const bool valid = wy < Ho && wx < Wo;
  float s = 0.0f;
  if (valid) {
    const float inv_n = 1.0f / static_cast<float>(p * p);
    float ax = 0.0f, ay = 0.0f;
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < p; ++j) {
        ax += sx[ly + i][lx + j];
        ay += sy[ly + i][lx + j];
      }
    const float mx = ax * inv_n, my = ay * inv_n;
    float vx = 0.0f, vy = 0.0f, cxy = 0.0f;
    for (int i = 0; i < p; ++i)
      for (int j = 0; j < p; ++j) {
        const float dx = sx[ly + i][lx + j] - mx, dy = sy[ly + i][lx + j] - my;
        vx = fmaf(dx, dx, vx);
        vy = fmaf(dy, dy, vy);
        cxy = fmaf(dx, dy, cxy);
      }
    const float sdx = sqrtf(vx * inv_n + eps), sdy = sqrtf(vy * inv_n + eps);
    s = cxy * inv_n / (sdx * sdy);
    const float alpha = inv_pn / (sdx * sdy), beta = s * inv_pn / (sdx * sdx);
    coef[(static_cast<int64_t>(b) * Ho + wy) * Wo + wx] = make_float4(alpha, my, beta, mx);
  }
  const float total = block_sum(s, red);
  if (threadIdx.x == 0) part[static_cast<int64_t>(b) * gridDim.x + blockIdx.x] = total;
}

__global__ __launch_bounds__(kThreads) void k_patch_pixels(const float* __restrict__ x, const float* __restrict__ y,
                                                           int H, int W, int p, const float4* __restrict__ coef,
                                                           const float* __restrict__ part, int nparts, float inv_p,
                                                           float* __restrict__ score, float* __restrict__ gx) {
  __shared__ float4 sc[kPatchSpan][kPatchSpan + 1];
  __shared__ float red[kThreads / 64];
  const int b = blockIdx.y;
  if (blockIdx.x == 0) {  // the image's score: its window partials in a fixed order
    float acc = 0.0f;
    for (int q = threadIdx.x; q < nparts; q += kThreads) acc += part[static_cast<int64_t>(b) * nparts + q];
    const float total = block_sum(acc, red);
    if (threadIdx.x == 0) score[b] = total * inv_p;
  }
  if (gx == nullptr) return;
  const int Ho = H - p + 1, Wo = W - p + 1;
  const int tiles_w = (W + kPatchTile - 1) / kPatchTile;
  const int ty = blockIdx.x / tiles_w, tx = blockIdx.x - ty * tiles_w;
  const int y0 = ty * kPatchTile, x0 = tx * kPatchTile;
  const int span = kPatchTile + p - 1;
  // windows (y0 - p + 1 + r, x0 - p + 1 + c) for r, c < span
  for (int e = threadIdx.x; e < span * span; e += kThreads) {
    const int r = e / span, c = e - r * span;
    const int wy = y0 - p + 1 + r, wx = x0 - p + 1 + c;
    const bool in = wy >= 0 && wy < Ho && wx >= 0 && wx < Wo;
    sc[r][c] = in ? coef[(static_cast<int64_t>(b) * Ho + wy) * Wo + wx] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  __syncthreads();
  const int ly = threadIdx.x / kPatchTile, lx = threadIdx.x % kPatchTile;
  const int py = y0 + ly, px = x0 + lx;
  if (py >= H || px >= W) return;
  const int64_t k = (static_cast<int64_t>(b) * H + py) * W + px;
  const float xv = x[k], yv = y[k];
  float g = 0.0f;
  for (int i = 0; i < p; ++i)
    for (int j = 0; j < p; ++j) {
      const float4 c = sc[ly + i][lx + j];  // windows (py - p + 1 + i, px - p + 1 + j)
      g = fmaf(c.x, yv - c.y, g);
      g = fmaf(-c.z, xv - c.w, g);
    }
  gx[k] = g;
}

// Registration-loop state kept on the device, so that a captured HIP graph of
// several iterations replays with each iteration's own values:
//   * the iteration counter (an int the loop advances with k_advance);
//   * the lattice-jitter offsets, drawn from a counter-based hash of
//     (seed, iteration, problem) instead of a generator's state;
//   * the Adam schedule, one row per iteration [lr_rot * f, lr_xyz * f,
//     1 - beta1^t, 1 - beta2^t] (f the cosine factor, t = iteration + 1),
//     computed once on the host in float64 and rounded to fp32.
// 32-bit finaliser of Wellons' hash prospector ("lowbias32").
__device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__device__ inline float lattice_offset(uint32_t seed, int iteration, int b) {
  uint32_t h = mix32(seed ^ 0x9e3779b9u);
  h = mix32(h ^ static_cast<uint32_t>(iteration));
  h = mix32(h ^ static_cast<uint32_t>(b));
  return static_cast<float>(h >> 8) * 0x1p-24f - 0.5f;  // uniform in [-1/2, 1/2)
}

__global__ void k_lattice_offsets(const int* __restrict__ iteration, uint32_t seed, float* __restrict__ offsets,
                                  int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  offsets[b] = lattice_offset(seed, *iteration, b);
}

// k_pose_compose and k_lattice_offsets in one launch (both one thread per problem).
__global__ void k_pose_compose_lattice(const double* __restrict__ prefix, const double* __restrict__ suffix,
                                       const float* __restrict__ rot, const float* __restrict__ xyz,
                                       float* __restrict__ M, int B, int64_t param_stride,
                                       const int* __restrict__ iteration, uint32_t seed, float* __restrict__ offsets) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  offsets[b] = lattice_offset(seed, *iteration, b);
  const double w[3] = {rot[b * param_stride], rot[b * param_stride + 1], rot[b * param_stride + 2]};
  const double v[3] = {xyz[b * param_stride], xyz[b * param_stride + 1], xyz[b * param_stride + 2]};
  ndrr_se3::compose(prefix + 16 * static_cast<int64_t>(b), suffix + 16 * static_cast<int64_t>(b), w, v,
                    M + 12 * static_cast<int64_t>(b));
}

// Element i's Adam update with schedule row `it` (the same update as k_adam).
__device__ inline void adam_scheduled_one(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                          float* __restrict__ v, int i, const float* __restrict__ schedule, int it,
                                          int lr_column, float b1, float c1, float b2, float c2, float eps) {
  const float* row = schedule + 4 * static_cast<int64_t>(it);
  const float lr = row[lr_column], bc1 = row[2], bc2 = row[3];
  const float gi = g[i];
  const float mi = m[i] * b1 + c1 * gi;
  const float vi = v[i] * b2 + c2 * (gi * gi);
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ void k_adam_scheduled(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                 float* __restrict__ v, int n, const float* __restrict__ schedule, int rows,
                                 int lr_column, const int* __restrict__ iteration, float b1, float c1, float b2,
                                 float c2, float eps) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int it = min(max(*iteration, 0), rows - 1);
  adam_scheduled_one(p, g, m, v, i, schedule, it, lr_column, b1, c1, b2, c2, eps);
}

// Both parameter groups' scheduled Adam steps (group k uses learning-rate
// column k) and the iteration counter's advance, in one workgroup: every
// thread reads the counter before the barrier, one thread advances it after.
constexpr int kPairThreads = 1024;
__global__ __launch_bounds__(kPairThreads) void k_adam_scheduled_pair(
    float* __restrict__ p0, const float* __restrict__ g0, float* __restrict__ m0, float* __restrict__ v0,
    float* __restrict__ p1, const float* __restrict__ g1, float* __restrict__ m1, float* __restrict__ v1, int n,
    const float* __restrict__ schedule, int rows, int* __restrict__ iteration, float b1, float c1, float b2, float c2,
    float eps) {
  const int it_raw = *iteration;
  const int it = min(max(it_raw, 0), rows - 1);
  for (int j = threadIdx.x; j < 2 * n; j += kPairThreads) {
    if (j < n) adam_scheduled_one(p0, g0, m0, v0, j, schedule, it, 0, b1, c1, b2, c2, eps);
    else adam_scheduled_one(p1, g1, m1, v1, j - n, schedule, it, 1, b1, c1, b2, c2, eps);
  }
  __syncthreads();
  if (threadIdx.x == 0) iteration[0] = it_raw + 1;
}

__global__ void k_advance(int* __restrict__ counter) {
  if (threadIdx.x == 0) counter[0] = counter[0] + 1;
}

int status() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? NDRR_OK : static_cast<int>(e);
}

}  // namespace

extern "C" {

int ndrr_pose_compose(const double* prefix, const double* suffix, const float* rot, const float* xyz,
                      int64_t param_stride, float* M, int B, void* stream) {
  if (!prefix || !suffix || !rot || !xyz || !M || B <= 0 || param_stride < 3) return NDRR_EINVAL;
  const int blocks = (B + 63) / 64;
  hipLaunchKernelGGL(k_pose_compose, dim3(blocks), dim3(64), 0, static_cast<hipStream_t>(stream), prefix, suffix,
                     rot, xyz, M, B, param_stride);
  return status();
}

int ndrr_pose_compose_bwd(const double* prefix, const double* suffix, const float* rot, const float* xyz,
                          int64_t param_stride, const float* gM, float* grot, float* gxyz, int B, void* stream) {
  if (!prefix || !suffix || !rot || !xyz || !gM || !grot || !gxyz || B <= 0 || param_stride < 3)
    return NDRR_EINVAL;
  const int blocks = (B + 63) / 64;
  hipLaunchKernelGGL(k_pose_compose_bwd, dim3(blocks), dim3(64), 0, static_cast<hipStream_t>(stream), prefix,
                     suffix, rot, xyz, gM, grot, gxyz, B, param_stride);
  return status();
}

int ndrr_adam_step(float* param, const float* grad, float* m, float* v, int n, float lr, double beta1, double beta2,
                   float eps, float bias_correction1, float bias_correction2, void* stream) {
  if (!param || !grad || !m || !v || n <= 0 || !(bias_correction1 > 0.0f) || !(bias_correction2 > 0.0f))
    return NDRR_EINVAL;
  if (!(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0)) return NDRR_EINVAL;
  hipLaunchKernelGGL(k_adam, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, static_cast<hipStream_t>(stream),
                     param, grad, m, v, n, lr, static_cast<float>(beta1), static_cast<float>(1.0 - beta1),
                     static_cast<float>(beta2), static_cast<float>(1.0 - beta2), eps, bias_correction1,
                     bias_correction2);
  return status();
}

int64_t ndrr_zncc_workspace_floats(int B, int N) {
  if (B <= 0 || N <= 0) return 0;
  return static_cast<int64_t>(B) * ((N + kZnccChunk - 1) / kZnccChunk) * kZnccStat;
}

int ndrr_zncc(const float* x, const float* y, int B, int N, float eps, float* workspace, float* score, float* grad_x,
              void* stream) {
  if (!x || !y || !workspace || !score || B <= 0 || N <= 0 || !(eps >= 0.0f)) return NDRR_EINVAL;
  if (B > 65535 || N > 0x7fffffff - kZnccChunk) return NDRR_ERANGE;
  const int P = (N + kZnccChunk - 1) / kZnccChunk;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_zncc_part, dim3(P, B), dim3(kThreads), 0, st, x, y, workspace, N);
  hipLaunchKernelGGL(k_zncc_finish, dim3(grad_x ? P : 1, B), dim3(kThreads), 0, st, x, y, workspace, P, score,
                     grad_x, N, eps);
  return status();
}

int64_t ndrr_patch_zncc_workspace_floats(int B, int H, int W, int patch) {
  if (B <= 0 || patch < 1 || patch > kPatchMax || H < patch || W < patch) return 0;
  const int64_t Ho = H - patch + 1, Wo = W - patch + 1;
  const int64_t tiles = ((Ho + kPatchTile - 1) / kPatchTile) * ((Wo + kPatchTile - 1) / kPatchTile);
  return static_cast<int64_t>(B) * (Ho * Wo * 4 + ((tiles + 3) / 4) * 4);
}

int ndrr_patch_zncc(const float* x, const float* y, int B, int H, int W, int patch, float eps, float* workspace,
                    float* score, float* grad_x, void* stream) {
  if (!x || !y || !workspace || !score || B <= 0 || H <= 0 || W <= 0 || !(eps >= 0.0f)) return NDRR_EINVAL;
  if (patch < 1 || patch > kPatchMax || patch > H || patch > W) return NDRR_EINVAL;
  if (reinterpret_cast<uintptr_t>(workspace) % 16 != 0) return NDRR_EINVAL;
  if (B > 65535 || static_cast<int64_t>(B) * H * W >= (int64_t{1} << 31)) return NDRR_ERANGE;
  const int Ho = H - patch + 1, Wo = W - patch + 1;
  const int wtiles = ((Ho + kPatchTile - 1) / kPatchTile) * ((Wo + kPatchTile - 1) / kPatchTile);
  const int ptiles = ((H + kPatchTile - 1) / kPatchTile) * ((W + kPatchTile - 1) / kPatchTile);
  float4* coef = reinterpret_cast<float4*>(workspace);
  float* part = workspace + static_cast<int64_t>(B) * Ho * Wo * 4;
  const float n = static_cast<float>(patch * patch);
  const float P = static_cast<float>(Ho) * static_cast<float>(Wo);
  const hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_patch_windows, dim3(wtiles, B), dim3(kThreads), 0, st, x, y, H, W, patch, eps,
                     1.0f / (P * n), coef, part);
  hipLaunchKernelGGL(k_patch_pixels, dim3(grad_x ? ptiles : 1, B), dim3(kThreads), 0, st, x, y, H, W, patch,
                     coef, part, wtiles, 1.0f / P, score, grad_x);
  return status();
}

int ndrr_lattice_offsets(const int* iteration, uint32_t seed, float* offsets, int B, void* stream) {
  if (!iteration || !offsets || B <= 0) return NDRR_EINVAL;
  hipLaunchKernelGGL(k_lattice_offsets, dim3((B + 63) / 64), dim3(64), 0, static_cast<hipStream_t>(stream),
                     iteration, seed, offsets, B);
  return status();
}

int ndrr_adam_step_scheduled(float* param, const float* grad, float* m, float* v, int n, const float* schedule,
                             int rows, int lr_column, const int* iteration, double beta1, double beta2, float eps,
                             void* stream) {
  if (!param || !grad || !m || !v || !schedule || !iteration || n <= 0 || rows <= 0) return NDRR_EINVAL;
  if ((lr_column != 0 && lr_column != 1) || !(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0))
    return NDRR_EINVAL;
  hipLaunchKernelGGL(k_adam_scheduled, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), param, grad, m, v, n, schedule, rows, lr_column, iteration,
                     static_cast<float>(beta1), static_cast<float>(1.0 - beta1), static_cast<float>(beta2),
                     static_cast<float>(1.0 - beta2), eps);
  return status();
}

int ndrr_counter_advance(int* counter, void* stream) {
  if (!counter) return NDRR_EINVAL;
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), counter);
  return status();
}

int ndrr_adam_step_scheduled_pair(float* param0, const float* grad0, float* m0, float* v0, float* param1,
                                  const float* grad1, float* m1, float* v1, int n, const float* schedule, int rows,
                                  int* iteration, double beta1, double beta2, float eps, void* stream) {
  if (!param0 || !grad0 || !m0 || !v0 || !param1 || !grad1 || !m1 || !v1 || !schedule || !iteration || n <= 0 ||
      rows <= 0)
    return NDRR_EINVAL;
  if (!(beta1 >= 0.0 && beta1 < 1.0) || !(beta2 >= 0.0 && beta2 < 1.0)) return NDRR_EINVAL;
  if (n > (1 << 20)) return NDRR_ERANGE;
  hipLaunchKernelGGL(k_adam_scheduled_pair, dim3(1), dim3(kPairThreads), 0, static_cast<hipStream_t>(stream), param0,
                     grad0, m0, v0, param1, grad1, m1, v1, n, schedule, rows, iteration, static_cast<float>(beta1),
                     static_cast<float>(1.0 - beta1), static_cast<float>(beta2), static_cast<float>(1.0 - beta2), eps);
  return status();
}

int ndrr_pose_compose_lattice(const double* prefix, const double* suffix, const float* rot, const float* xyz,
                              int64_t param_stride, float* M, int B, const int* iteration, uint32_t seed,
                              float* offsets, void* stream) {
  if (!prefix || !suffix || !rot || !xyz || !M || !iteration || !offsets || B <= 0 || param_stride < 3)
    return NDRR_EINVAL;
  hipLaunchKernelGGL(k_pose_compose_lattice, dim3((B + 63) / 64), dim3(64), 0, static_cast<hipStream_t>(stream),
                     prefix, suffix, rot, xyz, M, B, param_stride, iteration, seed, offsets);
  return status();
}

}  // extern "C"
