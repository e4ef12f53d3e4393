// synth.hip -- training pairs for the pose proposal, generated on the device.
//
// Each pair has its own shape: a union of 3 ellipsoids (centre, rotation,
// semi-axes).  Both clouds are area-uniform samples of the shape's boundary
// (independent draws), with N(0, noise^2) added per coordinate; the source is then
// moved by the pair's perturbation P.  The pair is expressed in the frame of the
// target (origin at the target centroid c, scaled by the target's RMS radius s),
// and the label is P^-1 in that frame: omega = log(R), tau = (R c + t - c) / s
// for P^-1 = [R | t].  One block per pair; a counter-based hash (splitmix64) of
// (seed, pair, point, draw) drives every random choice, so the set depends only
// on the seed.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lcr_kernels.hpp"

namespace lcr {

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ float u01(uint64_t h) {  // (0, 1)
  return ((float)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
}

struct Ell {
  float c[3], R[9], a[3];
};

__device__ __forceinline__ bool inside(const Ell& e, float x, float y, float z) {
  const float dx = x - e.c[0], dy = y - e.c[1], dz = z - e.c[2];
  // local = R^T d / a
  const float lx = (e.R[0] * dx + e.R[3] * dy + e.R[6] * dz) / e.a[0];
  const float ly = (e.R[1] * dx + e.R[4] * dy + e.R[7] * dz) / e.a[1];
  const float lz = (e.R[2] * dx + e.R[5] * dy + e.R[8] * dz) / e.a[2];
  return lx * lx + ly * ly + lz * lz < 1.0f;
}

// one area-uniform surface sample of the union of 3 ellipsoids
__device__ void sample_point(const Ell* E, const float* w, uint64_t key, float* out) {
  float p[3] = {0.f, 0.f, 0.f};
  for (int attempt = 0; attempt < 256; ++attempt) {
    const uint64_t h0 = mix64(key ^ (uint64_t)attempt * 0x632be59bd9b4e019ull);
    const uint64_t h1 = mix64(h0), h2 = mix64(h1), h3 = mix64(h2);
    const float r = u01(h0);
    const int e = r < w[0] ? 0 : (r < w[0] + w[1] ? 1 : 2);
    const float zc = 2.f * u01(h1) - 1.f;
    const float ph = 6.28318530718f * u01(h2);
    const float rr = sqrtf(fmaxf(0.f, 1.f - zc * zc));
    const float u[3] = {rr * cosf(ph), rr * sinf(ph), zc};
    const Ell& el = E[e];
    // accept with dA_ellipsoid / dA_sphere over its maximum
    const float f = sqrtf((u[0] / el.a[0]) * (u[0] / el.a[0]) + (u[1] / el.a[1]) * (u[1] / el.a[1]) +
                          (u[2] / el.a[2]) * (u[2] / el.a[2]));
    const float amin = fminf(el.a[0], fminf(el.a[1], el.a[2]));
    const float l[3] = {u[0] * el.a[0], u[1] * el.a[1], u[2] * el.a[2]};
    for (int i = 0; i < 3; ++i)
      p[i] = el.R[3 * i + 0] * l[0] + el.R[3 * i + 1] * l[1] + el.R[3 * i + 2] * l[2] + el.c[i];
    if (u01(h3) >= f * amin) continue;
    bool hidden = false;
    for (int o = 0; o < 3; ++o)
      if (o != e && inside(E[o], p[0], p[1], p[2])) hidden = true;
    if (!hidden) break;
  }
  out[0] = p[0];
  out[1] = p[1];
  out[2] = p[2];
}

__device__ __forceinline__ float gauss(uint64_t h) {
  const float a = u01(h), b = u01(mix64(h));
  return sqrtf(-2.f * logf(a)) * cosf(6.28318530718f * b);
}

// shapes: per pair 3 ellipsoids x (c3, R9, a3) = 45 floats and 3 area weights;
// P: per pair the source perturbation (row-major 4x4, double); Pinv: its inverse.
__global__ void __launch_bounds__(256)
k_synth_pairs(const float* __restrict__ shapes, const float* __restrict__ weights,
              const double* __restrict__ P, const double* __restrict__ Pinv, int npts, float noise,
              uint64_t seed, float4* __restrict__ X, float* __restrict__ labels,
              float* __restrict__ raw, float* __restrict__ frame) {
  const int pair = blockIdx.x, t = threadIdx.x;
  __shared__ Ell E[3];
  __shared__ float w[3];
  __shared__ double red[256][4];
  __shared__ double cs[4];
  if (t < 3) {
    const float* s = shapes + (int64_t)pair * 45 + 15 * t;
    for (int i = 0; i < 3; ++i) E[t].c[i] = s[i];
    for (int i = 0; i < 9; ++i) E[t].R[i] = s[3 + i];
    for (int i = 0; i < 3; ++i) E[t].a[i] = s[12 + i];
    w[t] = weights[(int64_t)pair * 3 + t];
  }
  __syncthreads();
  const double* Pp = P + (int64_t)pair * 16;
  float4* Xs = X + (int64_t)pair * 2 * npts;
  float4* Xt = Xs + npts;
  double acc[4] = {0, 0, 0, 0};
  for (int side = 0; side < 2; ++side) {
    for (int i = t; i < npts; i += 256) {
      const uint64_t key = mix64(seed ^ mix64(((uint64_t)pair << 24) ^ ((uint64_t)side << 23) ^ (uint64_t)i));
      float p[3];
      sample_point(E, w, key, p);
      for (int a = 0; a < 3; ++a) p[a] += noise * gauss(mix64(key + 101 + a));
      if (side == 0) {
        float q[3];
        for (int r = 0; r < 3; ++r)
          q[r] = (float)(Pp[4 * r] * p[0] + Pp[4 * r + 1] * p[1] + Pp[4 * r + 2] * p[2] + Pp[4 * r + 3]);
        Xs[i] = make_float4(q[0], q[1], q[2], 0.f);
      } else {
        Xt[i] = make_float4(p[0], p[1], p[2], 0.f);
        acc[0] += p[0];
        acc[1] += p[1];
        acc[2] += p[2];
      }
    }
  }
  for (int a = 0; a < 3; ++a) red[t][a] = acc[a];
  __syncthreads();
  if (t == 0) {
    double s[3] = {0, 0, 0};
    for (int k = 0; k < 256; ++k)
      for (int a = 0; a < 3; ++a) s[a] += red[k][a];
    for (int a = 0; a < 3; ++a) cs[a] = s[a] / npts;
  }
  __syncthreads();
  double ss = 0.0;
  for (int i = t; i < npts; i += 256) {
    const float4 q = Xt[i];
    const double dx = q.x - cs[0], dy = q.y - cs[1], dz = q.z - cs[2];
    ss += dx * dx + dy * dy + dz * dz;
  }
  red[t][3] = ss;
  __syncthreads();
  if (t == 0) {
    double s = 0;
    for (int k = 0; k < 256; ++k) s += red[k][3];
    cs[3] = sqrt(s / npts) > 1e-12 ? sqrt(s / npts) : 1.0;
  }
  __syncthreads();
  const float c0 = (float)cs[0], c1 = (float)cs[1], c2 = (float)cs[2];
  const float inv = (float)(1.0 / cs[3]);
  for (int side = 0; side < 2; ++side) {
    float4* Xc = side == 0 ? Xs : Xt;
    for (int i = t; i < npts; i += 256) {
      const float4 q = Xc[i];
      if (raw) {
        float* o = raw + (((int64_t)pair * 2 + side) * npts + i) * 3;
        o[0] = q.x;
        o[1] = q.y;
        o[2] = q.z;
      }
      Xc[i] = make_float4((q.x - c0) * inv, (q.y - c1) * inv, (q.z - c2) * inv, 0.f);
    }
  }
  if (t == 0) {
    const double* M = Pinv + (int64_t)pair * 16;
    // omega = log(R): angle from the trace, axis from the skew part
    const double tr = M[0] + M[5] + M[10];
    const double ca = fmin(1.0, fmax(-1.0, 0.5 * (tr - 1.0)));
    const double ang = acos(ca);
    const double sx = M[9] - M[6], sy = M[2] - M[8], sz = M[4] - M[1];
    const double sn = sin(ang);
    const double k = ang < 1e-9 ? 0.5 : 0.5 * ang / sn;
    double om[3] = {k * sx, k * sy, k * sz};
    double tau[3];
    for (int r = 0; r < 3; ++r)
      tau[r] = (M[4 * r] * cs[0] + M[4 * r + 1] * cs[1] + M[4 * r + 2] * cs[2] + M[4 * r + 3] - cs[r]) /
               cs[3];
    float* L = labels + (int64_t)pair * 6;
    for (int a = 0; a < 3; ++a) {
      L[a] = (float)om[a];
      L[3 + a] = (float)tau[a];
    }
    if (frame) {
      for (int a = 0; a < 4; ++a) frame[(int64_t)pair * 4 + a] = (float)cs[a];
    }
  }
}

}  // namespace

void launch_synth_pairs(const float* shapes, const float* weights, const double* P,
                        const double* Pinv, int npairs, int npts, float noise, uint64_t seed,
                        float4* X, float* labels, float* raw, float* frame, hipStream_t s) {
  hipLaunchKernelGGL(k_synth_pairs, dim3(npairs), dim3(256), 0, s, shapes, weights, P, Pinv, npts,
                     noise, seed, X, labels, raw, frame);
}

}  // namespace lcr
