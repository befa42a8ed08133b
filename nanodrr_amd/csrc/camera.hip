// Ray and camera set-up of a render on AMD Instinct MI355X (gfx950), one
// launch each instead of a chain of small tensor operations:
//   * the detector rays of a pinhole (or orthographic) camera and their
//     Riemann step (nanodrr's _make_tgt, _make_src and step_size,
//     src/nanodrr/drr/renderer.py:82,101-132);
//   * the camera -> voxel-index affine M = (world_to_voxel @ rt_inv)[:3] in
//     FP64, rounded once to fp32 (src/nanodrr/drr/backends.py:96), and its
//     adjoint back to rt_inv.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nanodrr_amd.h"

namespace {

constexpr int kThreads = 256;

// tgt[b][n] = sdd_b K^-1_b [u + 1/2, v + 1/2, 1] with n = v * width + u;
// src[b] = 0 (perspective, one source per pose) or tgt[b][n] - (0, 0, sdd_b)
// (orthographic, one per ray); step[b][n] = |tgt - src| / (S - 1).
__global__ __launch_bounds__(kThreads) void k_make_rays(const float* __restrict__ k_inv, int64_t k_bs,
                                                        const float* __restrict__ sdd, int64_t sdd_bs, int height,
                                                        int width, int ortho, float inv_s, float* __restrict__ tgt,
                                                        float* __restrict__ src, float* __restrict__ step) {
  const int b = blockIdx.y;
  const int n = blockIdx.x * kThreads + threadIdx.x;
  const int N = height * width;
  const float* K = k_inv + b * k_bs;
  const float d = sdd[b * sdd_bs];
  if (!ortho && n == 0) {
    src[3 * b + 0] = 0.0f;
    src[3 * b + 1] = 0.0f;
    src[3 * b + 2] = 0.0f;
  }
  if (n >= N) return;
  const float u = static_cast<float>(n % width) + 0.5f;
  const float v = static_cast<float>(n / width) + 0.5f;
  float t[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) t[i] = d * fmaf(K[3 * i], u, fmaf(K[3 * i + 1], v, K[3 * i + 2]));
  const int64_t r = static_cast<int64_t>(b) * N + n;
  tgt[3 * r + 0] = t[0];
  tgt[3 * r + 1] = t[1];
  tgt[3 * r + 2] = t[2];
  float len2;
  if (ortho) {
    src[3 * r + 0] = t[0];
    src[3 * r + 1] = t[1];
    src[3 * r + 2] = t[2] - d;
    const float dz = t[2] - (t[2] - d);
    len2 = dz * dz;
  } else {
    len2 = fmaf(t[0], t[0], fmaf(t[1], t[1], t[2] * t[2]));
  }
  step[r] = sqrtf(len2) * inv_s;
}

// M[b] = (W @ R_b)[:3] in FP64, W = world_to_voxel (4x4 float64), R_b = rt_inv[b] (4x4 fp32).
__global__ __launch_bounds__(kThreads) void k_camera_to_voxel(const double* __restrict__ w2v,
                                                              const float* __restrict__ rt, float* __restrict__ M,
                                                              int B) {
  const int e = blockIdx.x * kThreads + threadIdx.x;  // one thread per entry of M
  if (e >= 12 * B) return;
  const int b = e / 12, ij = e % 12, i = ij / 4, j = ij % 4;
  const float* R = rt + 16 * static_cast<int64_t>(b);
  double acc = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) acc = fma(w2v[4 * i + k], static_cast<double>(R[4 * k + j]), acc);
  M[e] = static_cast<float>(acc);
}

// Adjoint: g_rt[b] = W[:3]^T @ gM[b] (4x3 by 3x4) in FP64, rounded to fp32.
__global__ __launch_bounds__(kThreads) void k_camera_to_voxel_bwd(const double* __restrict__ w2v,
                                                                  const float* __restrict__ gM,
                                                                  float* __restrict__ g_rt, int B) {
  const int e = blockIdx.x * kThreads + threadIdx.x;  // one thread per entry of g_rt
  if (e >= 16 * B) return;
  const int b = e / 16, kj = e % 16, k = kj / 4, j = kj % 4;
  const float* G = gM + 12 * static_cast<int64_t>(b);
  double acc = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) acc = fma(w2v[4 * i + k], static_cast<double>(G[4 * i + j]), acc);
  g_rt[e] = static_cast<float>(acc);
}

int status() {
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? NDRR_OK : static_cast<int>(e);
}

}  // namespace

extern "C" {

int ndrr_make_rays(const float* k_inv, int64_t k_batch_stride, const float* sdd, int64_t sdd_batch_stride, int B,
                   int height, int width, int orthographic, int n_samples, float* tgt, float* src, float* step,
                   void* stream) {
  if (!k_inv || !sdd || !tgt || !src || !step || B <= 0 || height <= 0 || width <= 0 || n_samples < 2)
    return NDRR_EINVAL;
  if (k_batch_stride < 0 || sdd_batch_stride < 0) return NDRR_EINVAL;
  const int64_t N = static_cast<int64_t>(height) * width;
  if (B > 65535 || 3 * N * B >= (int64_t{1} << 31)) return NDRR_ERANGE;
  const float inv_s = static_cast<float>(1.0 / static_cast<double>(n_samples - 1));
  const dim3 grid(static_cast<unsigned>((N + kThreads - 1) / kThreads), B);
  hipLaunchKernelGGL(k_make_rays, grid, dim3(kThreads), 0, static_cast<hipStream_t>(stream), k_inv, k_batch_stride,
                     sdd, sdd_batch_stride, height, width, orthographic != 0, inv_s, tgt, src, step);
  return status();
}

int ndrr_camera_to_voxel(const double* world_to_voxel, const float* rt_inv, float* M, int B, void* stream) {
  if (!world_to_voxel || !rt_inv || !M || B <= 0) return NDRR_EINVAL;
  if (B > (1 << 26)) return NDRR_ERANGE;
  hipLaunchKernelGGL(k_camera_to_voxel, dim3((12 * B + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), world_to_voxel, rt_inv, M, B);
  return status();
}

int ndrr_camera_to_voxel_bwd(const double* world_to_voxel, const float* gM, float* g_rt_inv, int B, void* stream) {
  if (!world_to_voxel || !gM || !g_rt_inv || B <= 0) return NDRR_EINVAL;
  if (B > (1 << 26)) return NDRR_ERANGE;
  hipLaunchKernelGGL(k_camera_to_voxel_bwd, dim3((16 * B + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), world_to_voxel, gM, g_rt_inv, B);
  return status();
}

}  // extern "C"
