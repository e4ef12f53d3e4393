// lcr_math.hpp -- FP64 small-matrix solvers used at the end of every ICP
// iteration, and the convergence state machine.  All loops have compile-time
// trip counts and are unrolled, so every array index is a constant and the
// matrices stay in registers (no private scratch memory).
//
// Solvers:
//   * horn_rotation   -- point-to-point: the rotation maximising trace(R^T S) from
//     the 3x3 cross-covariance S, as the eigenvector of Horn's 4x4 matrix with the
//     largest eigenvalue (Horn 1987).  Equals Umeyama's U diag(1,1,det) V^T, which
//     TransformationEstimationSVD computes (pcl/registration/impl/
//     transformation_estimation_svd.hpp; called via fine_registration.cpp:121).
//   * chol6_solve     -- point-to-plane: the 6x6 normal equations of
//     TransformationEstimationPointToPlaneLLS, by Cholesky.
//   * eigh3           -- normals: symmetric 3x3 eigen-decomposition by cyclic
//     Jacobi (NormalEstimation's solvePlaneParameters, normal_estimation.cpp:108).
//   * conv_step       -- DefaultConvergenceCriteria::hasConverged.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "lcr_common.hpp"

namespace lcr {

#define LCR_HD __host__ __device__ inline

// Cyclic Jacobi on a symmetric 4x4 (upper and lower kept in sync).  On exit a is
// diagonal (eigenvalues) and v's columns are the eigenvectors.
LCR_HD void jacobi4(double a[4][4], double v[4][4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) v[i][j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 24; ++sweep) {
    double off = 0.0, dg = 0.0;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      dg += a[p][p] * a[p][p];
#pragma unroll
      for (int q = p + 1; q < 4; ++q) off += a[p][q] * a[p][q];
    }
    if (off == 0.0 || off <= 1e-32 * dg) break;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
#pragma unroll
      for (int q = p + 1; q < 4; ++q) {
        const double apq = a[p][q];
        if (apq != 0.0) {
          const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
          const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
          const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const double akp = a[k][p], akq = a[k][q];
            a[k][p] = c * akp - s * akq;
            a[k][q] = s * akp + c * akq;
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const double apk = a[p][k], aqk = a[q][k];
            a[p][k] = c * apk - s * aqk;
            a[q][k] = s * apk + c * aqk;
          }
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const double vkp = v[k][p], vkq = v[k][q];
            v[k][p] = c * vkp - s * vkq;
            v[k][q] = s * vkp + c * vkq;
          }
        }
      }
    }
  }
}

// S[a][b] = cross-covariance sum_i (s_i - mu_s)_a (t_i - mu_t)_b / n, source first.
// Returns the row-major rotation R (3x3) taking the source onto the target.
LCR_HD void horn_rotation(const double S[3][3], double R[3][3]) {
  const double sxx = S[0][0], sxy = S[0][1], sxz = S[0][2];
  const double syx = S[1][0], syy = S[1][1], syz = S[1][2];
  const double szx = S[2][0], szy = S[2][1], szz = S[2][2];
  double N[4][4] = {
      {sxx + syy + szz, syz - szy, szx - sxz, sxy - syx},
      {syz - szy, sxx - syy - szz, sxy + syx, szx + sxz},
      {szx - sxz, sxy + syx, -sxx + syy - szz, syz + szy},
      {sxy - syx, szx + sxz, syz + szy, -sxx - syy + szz}};
  // scale for a well-conditioned Jacobi (the eigenvectors are scale-free)
  double m = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) m = fmax(m, fabs(N[i][j]));
  const double inv = m > 0.0 ? 1.0 / m : 1.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) N[i][j] *= inv;
  double V[4][4];
  jacobi4(N, V);
  double best = N[0][0];
  double q0 = V[0][0], q1 = V[1][0], q2 = V[2][0], q3 = V[3][0];
#pragma unroll
  for (int c = 1; c < 4; ++c) {
    if (N[c][c] > best) {
      best = N[c][c];
      q0 = V[0][c];
      q1 = V[1][c];
      q2 = V[2][c];
      q3 = V[3][c];
    }
  }
  const double nq = sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
  q0 /= nq;
  q1 /= nq;
  q2 /= nq;
  q3 /= nq;
  R[0][0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
  R[0][1] = 2.0 * (q1 * q2 - q0 * q3);
  R[0][2] = 2.0 * (q1 * q3 + q0 * q2);
  R[1][0] = 2.0 * (q2 * q1 + q0 * q3);
  R[1][1] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3;
  R[1][2] = 2.0 * (q2 * q3 - q0 * q1);
  R[2][0] = 2.0 * (q3 * q1 - q0 * q2);
  R[2][1] = 2.0 * (q3 * q2 + q0 * q1);
  R[2][2] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
}

// Point-to-point increment from the FP64 raw moments of one iteration:
// v[0] n, v[1..3] sum s, v[4..6] sum t, v[7..15] sum s_a t_b, as laid out in
// lcr_common.hpp.  T is row-major 4x4.
LCR_HD void p2p_increment(const double* v, double T[16]) {
  const double n = v[0];
  const double ms[3] = {v[1] / n, v[2] / n, v[3] / n};
  const double mt[3] = {v[4] / n, v[5] / n, v[6] / n};
  double S[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) S[a][b] = v[7 + 3 * a + b] / n - ms[a] * mt[b];
  double R[3][3];
  horn_rotation(S, R);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) T[4 * i + j] = R[i][j];
    T[4 * i + 3] = mt[i] - (R[i][0] * ms[0] + R[i][1] * ms[1] + R[i][2] * ms[2]);
  }
  T[12] = 0.0;
  T[13] = 0.0;
  T[14] = 0.0;
  T[15] = 1.0;
}

// upper-triangle index of (i, j), i <= j, in a 6x6 row-major packing
LCR_HD constexpr int tri6(int i, int j) { return i * 6 - (i * (i - 1)) / 2 + (j - i); }

// Solve M x = b for symmetric positive definite M given as its packed upper
// triangle.  Returns false when a pivot is not positive.
LCR_HD bool chol6_solve(const double* up, const double* b, double x[6]) {
  double L[6][6];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) L[i][j] = 0.0;
  bool ok = true;
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double d = up[tri6(j, j)];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
    if (!(d > 0.0)) ok = false;
    const double ljj = sqrt(d > 0.0 ? d : 1.0);
    L[j][j] = ljj;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      double s = up[tri6(j, i)];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
      L[i][j] = s / ljj;
    }
  }
  double y[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = b[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
    y[i] = s / L[i][i];
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double s = y[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) s -= L[k][i] * x[k];
    x[i] = s / L[i][i];
  }
  return ok;
}

// TransformationEstimationPointToPlaneLLS::constructTransformationMatrix:
// T = [Rz(g) Ry(b) Rx(a) | t].
LCR_HD void p2pl_matrix(const double x[6], double T[16]) {
  const double ca = cos(x[0]), sa = sin(x[0]);
  const double cb = cos(x[1]), sb = sin(x[1]);
  const double cg = cos(x[2]), sg = sin(x[2]);
  T[0] = cg * cb;
  T[1] = -sg * ca + cg * sb * sa;
  T[2] = sg * sa + cg * sb * ca;
  T[3] = x[3];
  T[4] = sg * cb;
  T[5] = cg * ca + sg * sb * sa;
  T[6] = -cg * sa + sg * sb * ca;
  T[7] = x[4];
  T[8] = -sb;
  T[9] = cb * sa;
  T[10] = cb * ca;
  T[11] = x[5];
  T[12] = 0.0;
  T[13] = 0.0;
  T[14] = 0.0;
  T[15] = 1.0;
}

// Point-to-plane increment from v[1..21] (packed A^T A) and v[22..27] (A^T b).
// Returns false if the system is not positive definite (T = identity then).
LCR_HD bool p2pl_increment(const double* v, double T[16]) {
  double x[6];
  const bool ok = chol6_solve(v + 1, v + 22, x);
  if (!ok) {
#pragma unroll
    for (int i = 0; i < 16; ++i) T[i] = (i % 5 == 0) ? 1.0 : 0.0;
    return false;
  }
  p2pl_matrix(x, T);
  return true;
}

LCR_HD void mat4_mul(const double* A, const double* B, double* C) {
  double t[16];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k) s += A[4 * i + k] * B[4 * k + j];
      t[4 * i + j] = s;
    }
#pragma unroll
  for (int i = 0; i < 16; ++i) C[i] = t[i];
}

// Symmetric 3x3 eigen-decomposition (cyclic Jacobi, FP64).  a is row-major and
// overwritten; returns eigenvalues ascending in w and eigenvectors as columns of v.
LCR_HD void eigh3(double a[3][3], double w[3], double v[3][3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) v[i][j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 24; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    const double dg = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
    if (off == 0.0 || off <= 1e-34 * dg) break;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
      for (int q = p + 1; q < 3; ++q) {
        const double apq = a[p][q];
        if (apq != 0.0) {
          const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
          const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
          const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double akp = a[k][p], akq = a[k][q];
            a[k][p] = c * akp - s * akq;
            a[k][q] = s * akp + c * akq;
          }
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double apk = a[p][k], aqk = a[q][k];
            a[p][k] = c * apk - s * aqk;
            a[q][k] = s * apk + c * aqk;
          }
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const double vkp = v[k][p], vkq = v[k][q];
            v[k][p] = c * vkp - s * vkq;
            v[k][q] = s * vkp + c * vkq;
          }
        }
      }
    }
  }
  w[0] = a[0][0];
  w[1] = a[1][1];
  w[2] = a[2][2];
  // sort ascending, permuting eigenvector columns: compare-swaps (0,1) (1,2) (0,1)
#pragma unroll
  for (int pass = 0; pass < 3; ++pass) {
    const int i = (pass == 1) ? 1 : 0;
    const int j = (pass == 1) ? 2 : 1;
    if (w[j] < w[i]) {
      const double tw = w[i];
      w[i] = w[j];
      w[j] = tw;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double tv = v[k][i];
        v[k][i] = v[k][j];
        v[k][j] = tv;
      }
    }
  }
}

// DefaultConvergenceCriteria::hasConverged (default_convergence_criteria.hpp) with
// max_iterations_similar_transforms_ = 0 and failure_after_max_iter_ = false.  The
// rotation and translation tests read the increment's coefficients as float, as
// PCL's Matrix4f does.  Returns true when converged; st->state says why.
LCR_HD bool conv_step(PairState* st, const IcpConst& c, const double* Tinc, double mse) {
  bool similar = false;
  if (st->iterations >= c.max_iterations) {
    st->state = kIterations;
    return true;
  }
  const float t00 = (float)Tinc[0], t11 = (float)Tinc[5], t22 = (float)Tinc[10];
  const float t03 = (float)Tinc[3], t13 = (float)Tinc[7], t23 = (float)Tinc[11];
  const float tr = ((t00 + t11) + t22) - 1.0f;
  const double cos_angle = 0.5 * (double)tr;
  const float tsq = (t03 * t03 + t13 * t13) + t23 * t23;
  const double translation_sqr = (double)tsq;
  if (cos_angle >= c.rotation_threshold && translation_sqr <= c.translation_threshold) {
    if (st->similar >= 0) {
      st->state = kTransform;
      return true;
    }
    similar = true;
  }
  const double dm = fabs(mse - st->prev_mse);
  if (dm < c.mse_abs) {
    if (st->similar >= 0) {
      st->state = kAbsMse;
      return true;
    }
    similar = true;
  }
  if (dm / st->prev_mse < c.mse_rel) {
    if (st->similar >= 0) {
      st->state = kRelMse;
      return true;
    }
    similar = true;
  }
  st->similar = similar ? st->similar + 1 : 0;
  st->prev_mse = mse;
  return false;
}

}  // namespace lcr
