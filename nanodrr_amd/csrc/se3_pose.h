// The SE(3) pose step of registration in FP64, shared by the pose-step kernels
// (register.hip) and the contraction's pose reduction (raymarch.hip) so that
// both run the same operations in the same order:
//   pose = prefix @ T(w, v) @ suffix,  T = [R(w), R(w) v; 0, 1],  R(w) = exp([w]x)
// after nanodrr's Registration.pose (src/nanodrr/registration/registration.py:64-68)
// and convert(..., "so3_log") (src/nanodrr/geometry/se3.py:62-69,109-110).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ndrr_se3 {

// Rodrigues coefficients of R = I + A K + B K^2 (K = [w]x, th = |w|) and the
// derivatives a = A'(th)/th, b = B'(th)/th, by Taylor series below 1e-3.
struct Rod {
  double A, B, a, b;
};

inline __device__ Rod rodrigues(double th2) {
  Rod r;
  if (th2 < 1e-6) {
    r.A = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
    r.B = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
    r.a = -1.0 / 3.0 + th2 / 30.0 - th2 * th2 / 840.0;
    r.b = -1.0 / 12.0 + th2 / 180.0 - th2 * th2 / 6720.0;
  } else {
    const double th = sqrt(th2);
    const double s = sin(th), c = cos(th);
    r.A = s / th;
    r.B = (1.0 - c) / th2;
    r.a = (th * c - s) / (th2 * th);
    r.b = (th * s - 2.0 * (1.0 - c)) / (th2 * th2);
  }
  return r;
}

inline __device__ void hat(const double w[3], double K[3][3]) {
  K[0][0] = 0.0;   K[0][1] = -w[2]; K[0][2] = w[1];
  K[1][0] = w[2];  K[1][1] = 0.0;   K[1][2] = -w[0];
  K[2][0] = -w[1]; K[2][1] = w[0];  K[2][2] = 0.0;
}

inline __device__ void matmul3(const double X[3][3], const double Y[3][3], double Z[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Z[i][j] = X[i][0] * Y[0][j] + X[i][1] * Y[1][j] + X[i][2] * Y[2][j];
}

// T = [R, R v; 0, 1] as 3x4 rows.
inline __device__ void make_T(const double w[3], const double v[3], double R[3][3], double T[3][4]) {
  double K[3][3], K2[3][3];
  hat(w, K);
  matmul3(K, K, K2);
  const Rod c = rodrigues(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = (i == j ? 1.0 : 0.0) + c.A * K[i][j] + c.B * K2[i][j];
  for (int i = 0; i < 3; ++i) {
    T[i][0] = R[i][0]; T[i][1] = R[i][1]; T[i][2] = R[i][2];
    T[i][3] = R[i][0] * v[0] + R[i][1] * v[1] + R[i][2] * v[2];
  }
}

inline __device__ void load44(const double* p, double X[4][4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) X[i][j] = p[4 * i + j];
}

// M = (prefix @ T @ suffix)[:3, :] (rounded once to fp32) for one problem.
inline __device__ void compose(const double* prefix, const double* suffix, const double w[3], const double v[3],
                               float* M) {
  double R[3][3], T[3][4], P[4][4], S[4][4];
  make_T(w, v, R, T);
  load44(prefix, P);
  load44(suffix, S);
  double PT[3][4];  // rows 0..2 of prefix @ T
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) {
      double acc = (j == 3) ? P[i][3] : 0.0;
      for (int k = 0; k < 3; ++k) acc += P[i][k] * T[k][j];
      PT[i][j] = acc;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 4; ++j) {
      double acc = PT[i][3] * S[3][j];
      for (int k = 0; k < 3; ++k) acc += PT[i][k] * S[k][j];
      M[4 * i + j] = static_cast<float>(acc);
    }
}

// Adjoint of compose for one problem: G = dL/dM (3x4) -> dL/dw (grot[0..2]), dL/dv (gxyz[0..2]).
inline __device__ void compose_adjoint(const double* prefix, const double* suffix, const double w[3],
                                       const double v[3], const double G[3][4], float* grot, float* gxyz) {
  double R[3][3], T[3][4], P[4][4], S[4][4];
  make_T(w, v, R, T);
  load44(prefix, P);
  load44(suffix, S);
  // M = P[:3] @ T4 @ S with T4 = [T; 0 0 0 1]. dL/dT4 = P[:3]^T G S^T; only rows 0..2 of T4 vary.
  double GS[3][4];  // G @ S^T
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 4; ++k) {
      double acc = 0.0;
      for (int j = 0; j < 4; ++j) acc += G[i][j] * S[k][j];
      GS[i][k] = acc;
    }
  double GT[3][4];  // (P[:3]^T @ GS)[:3]
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 4; ++k) {
      double acc = 0.0;
      for (int i = 0; i < 3; ++i) acc += P[i][r] * GS[i][k];
      GT[r][k] = acc;
    }
  // T[:, :3] = R, T[:, 3] = R v  ->  dL/dR = GT[:, :3] + GT[:, 3] v^T,  dL/dv = R^T GT[:, 3]
  double GR[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) GR[i][j] = GT[i][j] + GT[i][3] * v[j];
  for (int j = 0; j < 3; ++j) gxyz[j] = static_cast<float>(R[0][j] * GT[0][3] + R[1][j] * GT[1][3] + R[2][j] * GT[2][3]);
  // dR/dw_k = a w_k K + A [e_k]x + b w_k K^2 + B ([e_k]x K + K [e_k]x)
  double K[3][3], K2[3][3];
  hat(w, K);
  matmul3(K, K, K2);
  const Rod c = rodrigues(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  for (int k = 0; k < 3; ++k) {
    double e[3] = {0.0, 0.0, 0.0};
    e[k] = 1.0;
    double E[3][3], EK[3][3], KE[3][3];
    hat(e, E);
    matmul3(E, K, EK);
    matmul3(K, E, KE);
    double acc = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double d = c.a * w[k] * K[i][j] + c.A * E[i][j] + c.b * w[k] * K2[i][j] + c.B * (EK[i][j] + KE[i][j]);
        acc += GR[i][j] * d;
      }
    grot[k] = static_cast<float>(acc);
  }
}

}  // namespace ndrr_se3
