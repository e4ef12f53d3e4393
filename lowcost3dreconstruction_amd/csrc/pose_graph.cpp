// pose_graph.cpp -- host side of the cycle driver: pose chaining with an even
// spread of the loop-closure residual, and centroid_align's initial pose for the
// top and bottom views.  Plain FP64 host arithmetic on a few 4x4 matrices; no
// device work.
//
// Replaces the in-place chaining of scripts/alignment.sh:105-113 (each pair's
// transform applied to every later capture file) and the seed step of
// alignment.sh:117-118 (pcl_tools/centroid_align.cpp:104-130).
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/lcr.h"
#include "runtime.hpp"

namespace {

void mul4(const double* A, const double* B, double* C) {
  double t[16];
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) {
      double v = 0.0;
      for (int k = 0; k < 4; ++k) v += A[4 * r + k] * B[4 * k + c];
      t[4 * r + c] = v;
    }
  memcpy(C, t, sizeof t);
}

void eye4(double* T) {
  for (int i = 0; i < 16; ++i) T[i] = (i % 5 == 0) ? 1.0 : 0.0;
}

// [w]x
void hat(const double* w, double* W) {
  W[0] = 0.0;
  W[1] = -w[2];
  W[2] = w[1];
  W[3] = w[2];
  W[4] = 0.0;
  W[5] = -w[0];
  W[6] = -w[1];
  W[7] = w[0];
  W[8] = 0.0;
}

void mul3(const double* A, const double* B, double* C) {
  double t[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) t[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
  memcpy(C, t, sizeof t);
}

// Coefficients of exp([w]x) = I + a [w]x + b [w]x^2 and of the SE(3) left
// Jacobian V = I + b [w]x + c [w]x^2, theta = |w| (series near 0).
void so3_coeffs(double th, double* a, double* b, double* c) {
  if (th < 1e-4) {
    const double t2 = th * th;
    *a = 1.0 - t2 / 6.0;
    *b = 0.5 - t2 / 24.0;
    *c = 1.0 / 6.0 - t2 / 120.0;
  } else {
    *a = sin(th) / th;
    *b = (1.0 - cos(th)) / (th * th);
    *c = (th - sin(th)) / (th * th * th);
  }
}

// Rotation vector of a rotation matrix R (row-major 3x3), angle in [0, pi].
void so3_log(const double* R, double* w) {
  const double tr = R[0] + R[4] + R[8];
  const double ct = fmin(1.0, fmax(-1.0, 0.5 * (tr - 1.0)));
  const double th = acos(ct);
  const double v[3] = {R[7] - R[5], R[2] - R[6], R[3] - R[1]};
  if (th < 1e-4) {
    // sin(th) / th ~ 1 - th^2 / 6
    const double f = 0.5 * (1.0 + th * th / 6.0);
    for (int i = 0; i < 3; ++i) w[i] = f * v[i];
    return;
  }
  if (M_PI - th > 1e-6) {
    const double f = th / (2.0 * sin(th));
    for (int i = 0; i < 3; ++i) w[i] = f * v[i];
    return;
  }
  // near pi: the axis from the symmetric part, R + I = 2 n n^T
  int k = 0;
  for (int i = 1; i < 3; ++i)
    if (R[4 * i] > R[4 * k]) k = i;
  double n[3];
  const double d = sqrt(fmax(0.0, 0.5 * (R[4 * k] + 1.0)));
  for (int i = 0; i < 3; ++i) n[i] = (i == k) ? d : 0.5 * (R[3 * i + k] + R[3 * k + i]) / (2.0 * d);
  // the antisymmetric part (2 sin(th) n) still fixes the sign short of pi itself
  const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  const double sg = (v[0] * n[0] + v[1] * n[1] + v[2] * n[2]) < 0.0 ? -1.0 : 1.0;
  for (int i = 0; i < 3; ++i) w[i] = sg * th * n[i] / nn;
}

// T^s = exp(s log T) for a rigid 4x4 T.
void se3_pow(const double* T, double s, double* out) {
  const double R[9] = {T[0], T[1], T[2], T[4], T[5], T[6], T[8], T[9], T[10]};
  const double t[3] = {T[3], T[7], T[11]};
  double w[3];
  so3_log(R, w);
  const double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double a, b, c;
  so3_coeffs(th, &a, &b, &c);
  double W[9], W2[9], V[9];
  hat(w, W);
  mul3(W, W, W2);
  for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0 ? 1.0 : 0.0) + b * W[i] + c * W2[i];
  // u = V^-1 t (V is invertible for th < 2 pi)
  double Vi[9];
  const double det = V[0] * (V[4] * V[8] - V[5] * V[7]) - V[1] * (V[3] * V[8] - V[5] * V[6]) +
                     V[2] * (V[3] * V[7] - V[4] * V[6]);
  Vi[0] = (V[4] * V[8] - V[5] * V[7]) / det;
  Vi[1] = (V[2] * V[7] - V[1] * V[8]) / det;
  Vi[2] = (V[1] * V[5] - V[2] * V[4]) / det;
  Vi[3] = (V[5] * V[6] - V[3] * V[8]) / det;
  Vi[4] = (V[0] * V[8] - V[2] * V[6]) / det;
  Vi[5] = (V[2] * V[3] - V[0] * V[5]) / det;
  Vi[6] = (V[3] * V[7] - V[4] * V[6]) / det;
  Vi[7] = (V[1] * V[6] - V[0] * V[7]) / det;
  Vi[8] = (V[0] * V[4] - V[1] * V[3]) / det;
  double u[3];
  for (int i = 0; i < 3; ++i) u[i] = Vi[3 * i] * t[0] + Vi[3 * i + 1] * t[1] + Vi[3 * i + 2] * t[2];
  // exp(s (w, u))
  const double ws[3] = {s * w[0], s * w[1], s * w[2]};
  const double us[3] = {s * u[0], s * u[1], s * u[2]};
  so3_coeffs(fabs(s) * th, &a, &b, &c);
  hat(ws, W);
  mul3(W, W, W2);
  eye4(out);
  for (int r = 0; r < 3; ++r) {
    double tr = 0.0;
    for (int col = 0; col < 3; ++col) {
      const int i = 3 * r + col;
      const double id = (r == col) ? 1.0 : 0.0;
      out[4 * r + col] = id + a * W[i] + b * W2[i];
      tr += (id + b * W[i] + c * W2[i]) * us[col];
    }
    out[4 * r + 3] = tr;
  }
}

}  // namespace

extern "C" {

int lcr_chain_poses(int32_t n_captures, const double* T, const double* T_close, double* poses) {
  if (n_captures < 1 || !poses || (n_captures > 1 && !T))
    return lcr::set_last_error(LCR_E_ARG, "invalid arguments");
  eye4(poses);
  for (int j = 1; j < n_captures; ++j) mul4(poses + 16 * (j - 1), T + 16 * (j - 1), poses + 16 * j);
  if (!T_close || n_captures < 2) return LCR_OK;
  // E = P_{n-1} T_close: the identity for a consistent loop of n edges
  double E[16];
  mul4(poses + 16 * (n_captures - 1), T_close, E);
  for (int j = 1; j < n_captures; ++j) {
    double C[16];
    se3_pow(E, -(double)j / (double)n_captures, C);
    mul4(C, poses + 16 * j, poses + 16 * j);
  }
  return LCR_OK;
}

int lcr_se3_pow(const double T[16], double s, double out[16]) {
  if (!T || !out) return lcr::set_last_error(LCR_E_ARG, "invalid arguments");
  se3_pow(T, s, out);
  return LCR_OK;
}

int lcr_centroid_align(const float* src, int64_t n_src, const float* tgt, int64_t n_tgt,
                       double roll_deg, double pitch_deg, double yaw_deg, double elevation,
                       double T[16]) {
  if (!T || n_src < 0 || n_tgt < 0 || (n_src > 0 && !src) || (n_tgt > 0 && !tgt))
    return lcr::set_last_error(LCR_E_ARG, "invalid arguments");
  // pcl::compute3DCentroid over the finite points
  auto centroid = [](const float* p, int64_t n, double* c) {
    double s[3] = {0.0, 0.0, 0.0};
    int64_t m = 0;
    for (int64_t i = 0; i < n; ++i) {
      const float* q = p + 3 * i;
      if (!isfinite(q[0]) || !isfinite(q[1]) || !isfinite(q[2])) continue;
      for (int a = 0; a < 3; ++a) s[a] += q[a];
      ++m;
    }
    for (int a = 0; a < 3; ++a) c[a] = m > 0 ? s[a] / (double)m : 0.0;
    return m;
  };
  double cs[3], ct[3];
  if (centroid(src, n_src, cs) == 0 || centroid(tgt, n_tgt, ct) == 0)
    return lcr::set_last_error(LCR_E_ARG, "a cloud has no finite point");
  // Rx(roll) Ry(pitch) Rz(yaw)
  const double d2r = M_PI / 180.0;
  const double a = roll_deg * d2r, b = pitch_deg * d2r, g = yaw_deg * d2r;
  const double Rx[9] = {1, 0, 0, 0, cos(a), -sin(a), 0, sin(a), cos(a)};
  const double Ry[9] = {cos(b), 0, sin(b), 0, 1, 0, -sin(b), 0, cos(b)};
  const double Rz[9] = {cos(g), -sin(g), 0, sin(g), cos(g), 0, 0, 0, 1};
  double R[9];
  mul3(Rx, Ry, R);
  mul3(R, Rz, R);
  // x -> R (x - cs) + ct + (0, elevation, 0)
  eye4(T);
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) T[4 * r + c] = R[3 * r + c];
    T[4 * r + 3] = ct[r] + (r == 1 ? elevation : 0.0) -
                   (R[3 * r] * cs[0] + R[3 * r + 1] * cs[1] + R[3 * r + 2] * cs[2]);
  }
  return LCR_OK;
}

}  // extern "C"
