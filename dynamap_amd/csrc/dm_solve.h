// FP64 small-matrix routines used on the device (batched tracking) and on the
// host (per-frame path with a host solve, as the reference does):
//   - 6x6 LDL^T solve (replaces Eigen's ldlt().solve, gnSolver.cu:439-442)
//   - closed-form SE(3) exponential of v2t(xi) (replaces Eigen's matrix
//     exp(), gnSolver.cu:443-455): R = I + a W + b W^2, V = I + b W + c W^2,
//     a = sin(t)/t, b = (1 - cos t)/t^2, c = (t - sin t)/t^3
//   - the Gauss-Newton step rule of gnSolver.cu:400-411.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace dm {

#define DM_HD __host__ __device__ __forceinline__

// LDL^T without pivoting (A symmetric positive definite, row-major 6x6):
// the factorisation, then the solve for one right-hand side.
DM_HD int ldlt6_factor(const double *A, double *L, double *D) {
  for (int k = 0; k < 36; ++k) L[k] = 0.0;
  for (int j = 0; j < 6; ++j) {
    double d = A[6 * j + j];
    for (int k = 0; k < j; ++k) d -= L[6 * j + k] * L[6 * j + k] * D[k];
    D[j] = d;
    if (!(d > 0.0)) return -1;
    L[6 * j + j] = 1.0;
    for (int i = j + 1; i < 6; ++i) {
      double a = A[6 * i + j];
      for (int k = 0; k < j; ++k) a -= L[6 * i + k] * L[6 * j + k] * D[k];
      L[6 * i + j] = a / d;
    }
  }
  return 0;
}

DM_HD void ldlt6_solve(const double *L, const double *D, const double *b, double *x) {
  double y[6];
  for (int i = 0; i < 6; ++i) {
    double a = b[i];
    for (int k = 0; k < i; ++k) a -= L[6 * i + k] * y[k];
    y[i] = a;
  }
  for (int i = 0; i < 6; ++i) y[i] /= D[i];
  for (int i = 5; i >= 0; --i) {
    double a = y[i];
    for (int k = i + 1; k < 6; ++k) a -= L[6 * k + i] * x[k];
    x[i] = a;
  }
}

DM_HD int ldlt6(const double *A, const double *b, double *x) {
  double L[36], D[6];
  if (ldlt6_factor(A, L, D) != 0) return -1;
  ldlt6_solve(L, D, b, x);
  return 0;
}

DM_HD void se3_exp(const double *xi, double *T) {
  const double w0 = xi[0], w1 = xi[1], w2 = xi[2];
  const double th2 = w0 * w0 + w1 * w1 + w2 * w2, th = sqrt(th2);
  double a, b, c;
  if (th < 1e-5) {
    a = 1.0 - th2 / 6.0 + th2 * th2 / 120.0;
    b = 0.5 - th2 / 24.0 + th2 * th2 / 720.0;
    c = 1.0 / 6.0 - th2 / 120.0 + th2 * th2 / 5040.0;
  } else {
    a = sin(th) / th;
    b = (1.0 - cos(th)) / th2;
    c = (th - sin(th)) / (th2 * th);
  }
  const double W[9] = {0, -w2, w1, w2, 0, -w0, -w1, w0, 0};
  double W2[9];
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) W2[3 * r + k] = W[3 * r] * W[k] + W[3 * r + 1] * W[3 + k] + W[3 * r + 2] * W[6 + k];
  double R[9], V[9];
  for (int k = 0; k < 9; ++k) {
    const double I = (k % 4 == 0) ? 1.0 : 0.0;
    R[k] = I + a * W[k] + b * W2[k];
    V[k] = I + b * W[k] + c * W2[k];
  }
  for (int r = 0; r < 3; ++r) {
    T[4 * r] = R[3 * r]; T[4 * r + 1] = R[3 * r + 1]; T[4 * r + 2] = R[3 * r + 2];
    T[4 * r + 3] = V[3 * r] * xi[3] + V[3 * r + 1] * xi[4] + V[3 * r + 2] * xi[5];
  }
  T[12] = 0; T[13] = 0; T[14] = 0; T[15] = 1;
}

DM_HD void mat4_mul(const double *A, const double *B, double *C) {
  double t[16];
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 4; ++k)
      t[4 * r + k] = A[4 * r] * B[k] + A[4 * r + 1] * B[4 + k] + A[4 * r + 2] * B[8 + k] + A[4 * r + 3] * B[12 + k];
  for (int k = 0; k < 16; ++k) C[k] = t[k];
}

// Rigid inverse in FP64, rounded to float (R^T, -R^T t).
DM_HD void rigid_inverse_f(const double *T, float *Ti) {
  for (int r = 0; r < 3; ++r) {
    for (int k = 0; k < 3; ++k) Ti[4 * r + k] = (float)T[4 * k + r];
    Ti[4 * r + 3] = (float)(-(T[0 * 4 + r] * T[3] + T[1 * 4 + r] * T[7] + T[2 * 4 + r] * T[11]));
  }
  Ti[12] = 0.f; Ti[13] = 0.f; Ti[14] = 0.f; Ti[15] = 1.f;
}

// Rigid inverse of a float matrix, computed in FP64 and rounded to float.
DM_HD void rigid_inverse_ff(const float *T, float *Ti) {
  double Td[16];
  for (int k = 0; k < 16; ++k) Td[k] = (double)T[k];
  rigid_inverse_f(Td, Ti);
}

// gnSolver.cu:400-411: scale by 1/max(A), add (lambda*max(A_scaled)/it) I,
// xi = LDLT^-1 (-b).  A and b are the raw sums (b = sum J r).
DM_HD int icp_step(const double *A_in, const double *b_in, int it, double lambda, double *xi) {
  double A[36], b[6];
  double mx = A_in[0];
  for (int k = 1; k < 36; ++k) if (A_in[k] > mx) mx = A_in[k];
  if (!(mx > 0.0)) { for (int k = 0; k < 6; ++k) xi[k] = 0.0; return -1; }
  const double sc = 1.0 / mx;
  for (int k = 0; k < 36; ++k) A[k] = A_in[k] * sc;
  for (int k = 0; k < 6; ++k) b[k] = -b_in[k] * sc;
  double mx2 = A[0];
  for (int k = 1; k < 36; ++k) if (A[k] > mx2) mx2 = A[k];
  const double alpha = lambda * mx2 / (double)it;
  for (int k = 0; k < 6; ++k) A[7 * k] += alpha;
  if (ldlt6(A, b, xi) != 0) { for (int k = 0; k < 6; ++k) xi[k] = 0.0; return -1; }
  return 0;
}

// gnSolver.cu:413-415 with the summed twist acc and its value prev after the
// last step that did not end a level: true (the level ends) when
// |acc - prev| <= min_inc; otherwise prev <- acc.
DM_HD bool level_converged(const double *acc, double *prev, double min_inc) {
  double c2 = 0.0;
  for (int k = 0; k < 6; ++k) {
    const double c = acc[k] - prev[k];
    c2 += c * c;
  }
  if (sqrt(c2) <= min_inc) return true;
  for (int k = 0; k < 6; ++k) prev[k] = acc[k];
  return false;
}

// Expand the 21 upper-triangular sums (row-major order) into a full 6x6.
DM_HD void sym6_from_upper(const double *u21, double *A) {
  int k = 0;
  for (int r = 0; r < 6; ++r)
    for (int c = r; c < 6; ++c) { A[6 * r + c] = u21[k]; A[6 * c + r] = u21[k]; ++k; }
}

}  // namespace dm
