// tile_search.hpp -- radius-limited 1-NN of one wavefront's 64 queries by
// reduced-precision matrix-core distance tiles and an exact FP32 rerank.
//
// Replaces the same KdTreeFLANN::nearestKSearch(k = 1) call as lcr_search.hpp
// (CorrespondenceEstimation, fine_registration.cpp:121) and returns the same
// match: the FLANN-form float distance, ties to the smallest original index.
//
//  1. The wavefront's queries span a box; the grid cells overlapping the box
//     grown by rho = min(r, 0.75 cells) (plus a binning-rounding slack) are
//     staged in LDS: one lane per row of cells finds the runs, a prefix sum
//     places them, then a flat copy.  A query whose nearest staged point is
//     closer than its distance to the grown box (or whose distance to the box
//     exceeds r) is final; the few others continue in the exact scalar search,
//     seeded with the staged answer.
//  2. Coordinates relative to the box centre c are rounded to FP16 or BF16 and
//     d~^2 = |q~|^2 + |p~|^2 - 2 q~.p~ is formed with the dot products on
//     v_mfma_f32_16x16x32_{f16,bf16} (16 queries x 16 candidates per
//     instruction, FP32 accumulation) and the norms in FP32.
//  3. Rerank: with e the rounding bound of one point (|p~ - p| <= e) and E2 the
//     FP32 bound on the computed d~^2, the exact nearest p* satisfies
//        d~^2(p*) <= ((sqrt(m + E2) + 2e) (1 + 2e-6) + 2e)^2 (1 + 2e-6) + E2,
//     where m is the smallest computed d~^2 seen so far.  Every candidate under
//     that bound (a few per query) gets the exact FLANN-form float distance, and
//     the lexicographic minimum of (d2, original index) over those is the
//     answer.  The bound only shrinks as m falls, so one pass suffices.
// Per-point rounding bound: e = (u + 2^-24) R + sqrt(3) 2^-25, u = 2^-11 (FP16)
// or 2^-8 (BF16), R the largest |v - c| in the staged block (the last term is
// FP16's subnormal spacing).  The tile reports failure (the caller falls back to
// the scalar search) when the block holds more than kTileCap points, or, for
// FP16, when R could overflow the format.
#pragma once
#include <hip/hip_runtime.h>

#include "lcr_common.hpp"
#include "lcr_search.hpp"

namespace lcr {

#pragma clang fp contract(off)

constexpr int kTileCap = 512;

// Staging area of one wavefront.
struct TileLds {
  float4 p[kTileCap];  // candidate xyz + bit-cast original index
  int j[kTileCap];     // candidate position in the sorted point array
  float4 q[64];        // queries (w unused)
  float qn[64];        // |q~|^2, NaN for lanes without a query
  float rd[64];        // per-query result: d2, original index, sorted position
  int rid[64];
  int rj[64];
};

typedef _Float16 lcr_f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 lcr_bf16x8 __attribute__((ext_vector_type(8)));
typedef float lcr_f32x4 __attribute__((ext_vector_type(4)));

// Reduced-precision copy of one relative coordinate, and its value back in FP32.
template <int PREC>
__device__ __forceinline__ float round_coord(float v) {
  if constexpr (PREC == 1)
    return (float)(_Float16)v;
  else
    return (float)(__bf16)v;
}

template <int PREC>
__device__ __forceinline__ lcr_f32x4 mfma_dot(float ax, float ay, float az, float bx, float by,
                                              float bz, bool hold) {
  const lcr_f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  if constexpr (PREC == 1) {
    lcr_f16x8 a = {}, b = {};
    if (hold) {
      a[0] = (_Float16)ax;
      a[1] = (_Float16)ay;
      a[2] = (_Float16)az;
      b[0] = (_Float16)bx;
      b[1] = (_Float16)by;
      b[2] = (_Float16)bz;
    }
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, zero, 0, 0, 0);
  } else {
    lcr_bf16x8 a = {}, b = {};
    if (hold) {
      a[0] = (__bf16)ax;
      a[1] = (__bf16)ay;
      a[2] = (__bf16)az;
      b[0] = (__bf16)bx;
      b[1] = (__bf16)by;
      b[2] = (__bf16)bz;
    }
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, zero, 0, 0, 0);
  }
}

__device__ __forceinline__ float wave_minf(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ void tile_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Grid cell range [c0, c1] along one axis covering [lo, hi] (world units).  Points
// are binned with the same (p - o) * inv_h arithmetic; the slack (1e-3 + 1e-6 |f|
// cells, as in cell_pos) covers its rounding.  Returns false if empty.
__device__ __forceinline__ bool axis_cells(float lo, float hi, float o, float inv_h, int n, int& c0,
                                           int& c1) {
  const float f0 = fminf(fmaxf((lo - o) * inv_h, -4.0e6f), 4.0e6f);
  const float f1 = fminf(fmaxf((hi - o) * inv_h, -4.0e6f), 4.0e6f);
  const float s0 = 1.0e-3f + 1.0e-6f * fabsf(f0), s1 = 1.0e-3f + 1.0e-6f * fabsf(f1);
  c0 = max((int)floorf(f0 - s0), 0);
  c1 = min((int)floorf(f1 + s1), n - 1);
  return c0 <= c1;
}

// The staged block of one wavefront's tile.
struct TileBox {
  float lox, hix, loy, hiy, loz, hiz;  // the queries' box
  float rr;                            // growth rho = min(r, margin)
  int x0, x1, y0, y1, z0, z1;          // staged cells
  int M;                               // staged points
};

// Stages the cells overlapping the own queries' box grown by rho = min(r, margin).
// Must be called by all 64 lanes.  Returns 0 when the block cannot be staged
// (more than kTileCap points or rows), 1 when staged (B.M may be 0), 2 when no
// grid cell lies within rho of any query.
__device__ __forceinline__ int tile_stage(const GridDesc& g, const int* __restrict__ cs,
                                          const float4* __restrict__ P, bool own, float qx,
                                          float qy, float qz, float r, float margin, TileLds& L,
                                          TileBox& B) {
  const int lane = threadIdx.x & 63;
  B.lox = wave_minf(own ? qx : INFINITY);
  B.hix = wave_maxf(own ? qx : -INFINITY);
  B.loy = wave_minf(own ? qy : INFINITY);
  B.hiy = wave_maxf(own ? qy : -INFINITY);
  B.loz = wave_minf(own ? qz : INFINITY);
  B.hiz = wave_maxf(own ? qz : -INFINITY);
  B.rr = fminf(r * 1.00001f, margin);
  B.M = 0;
  const float rr = B.rr;
  if (!axis_cells(B.lox - rr, B.hix + rr, g.ox, g.inv_h, g.nx, B.x0, B.x1) ||
      !axis_cells(B.loy - rr, B.hiy + rr, g.oy, g.inv_h, g.ny, B.y0, B.y1) ||
      !axis_cells(B.loz - rr, B.hiz + rr, g.oz, g.inv_h, g.nz, B.z0, B.z1))
    return 2;
  const int x0 = B.x0, x1 = B.x1, y0 = B.y0, y1 = B.y1, z0 = B.z0, z1 = B.z1;
  const int nyb = y1 - y0 + 1, nrows = (z1 - z0 + 1) * nyb;
  if (nrows > kTileCap) return 0;
  tile_lds_sync();
  int M = 0;
  for (int rb = 0; rb < nrows; rb += 64) {
    const int row_i = rb + lane;
    int s = 0, len = 0;
    if (row_i < nrows) {
      const int z = z0 + row_i / nyb, y = y0 + row_i % nyb;
      const int row = (z * g.ny + y) * g.nx;
      s = cs[row + x0];
      len = cs[row + x1 + 1] - s;
    }
    int inc = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(inc, off, 64);
      if (lane >= off) inc += t;
    }
    const int total = __shfl(inc, 63, 64);
    if (M + total > kTileCap) return 0;
    const int excl = inc - len;
    const int nr = min(64, nrows - rb);
    for (int t0 = 0; t0 < total; t0 += 256) {
      int tt[4], jb[4] = {0, 0, 0, 0};
#pragma unroll
      for (int u = 0; u < 4; ++u) tt[u] = min(t0 + lane + 64 * u, total - 1);
      for (int k = 0; k < nr; ++k) {
        const int ex = __builtin_amdgcn_readlane(excl, k);
        const int sr = __builtin_amdgcn_readlane(s, k);
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (ex <= tt[u]) jb[u] = sr - ex;
      }
      float4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = P[jb[u] + tt[u]];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = t0 + lane + 64 * u;
        if (t < total) {
          L.p[M + t] = v[u];
          L.j[M + t] = jb[u] + t;
        }
      }
    }
    M += total;
  }
  B.M = M;
  tile_lds_sync();
  return 1;
}

// Certification of a lane's staged answer: every point within distance D of the
// query lies in the grown box, whose cells were all staged.
__device__ __forceinline__ bool tile_certified(const TileBox& B, float qx, float qy, float qz,
                                               float r, float best) {
  const float rr = B.rr;
  const float D = fminf(fminf(fminf(qx - (B.lox - rr), (B.hix + rr) - qx),
                              fminf(qy - (B.loy - rr), (B.hiy + rr) - qy)),
                        fminf(qz - (B.loz - rr), (B.hiz + rr) - qz));
  return (D >= r * 1.00001f) || (D > 0.0f && best < D * D * 0.99998f);
}

// Exact FP32 variant: each lane scans every staged point with the FLANN-form
// distance (LDS broadcast reads).  Same contract as tile_nn.
__device__ __forceinline__ bool tile_nn_fp32(const GridDesc& g, const int* __restrict__ cs,
                                             const float4* __restrict__ P, bool own, float qx,
                                             float qy, float qz, float r, float margin, TileLds& L,
                                             float& best, int& bid, int& bj, bool& cert) {
  cert = false;
  if (__ballot(own) == 0 || g.npts == 0) return true;
  TileBox B;
  const int st = tile_stage(g, cs, P, own, qx, qy, qz, r, margin, L, B);
  if (st == 0) return false;
  if (st == 2 || B.M == 0) {
    cert = B.rr >= r * 1.00001f;
    return true;
  }
  if (own) {
    float bd = INFINITY;
    int bi = 0x7fffffff, bt = -1;
    int t = 0;
    for (; t + 4 <= B.M; t += 4) {
      const float4 p0 = L.p[t], p1 = L.p[t + 1], p2 = L.p[t + 2], p3 = L.p[t + 3];
      consider(p0, t, qx, qy, qz, bd, bi, bt);
      consider(p1, t + 1, qx, qy, qz, bd, bi, bt);
      consider(p2, t + 2, qx, qy, qz, bd, bi, bt);
      consider(p3, t + 3, qx, qy, qz, bd, bi, bt);
    }
    for (; t < B.M; ++t) consider(L.p[t], t, qx, qy, qz, bd, bi, bt);
    if (bt >= 0 && (bd < best || (bd == best && bi < bid))) {
      best = bd;
      bid = bi;
      bj = L.j[bt];
    }
    cert = tile_certified(B, qx, qy, qz, r, best);
  }
  return true;
}

// Must be called by all 64 lanes.  `own` lanes hold a finite query (qx, qy, qz).
// The staged block covers the queries' box grown by rho = min(r, margin).  On
// success the lane's (best, bid, bj) is the FLANN-form nearest staged point, and
// `cert` says whether that is final: the nearest point overall (its distance is
// below the query's distance to the grown box), or, when that distance exceeds r,
// no point outside the block can be within r.  An uncertified lane's result is a
// valid candidate (an upper bound) for the caller's exact search.  Returns false,
// leaving the outputs untouched, when the tile cannot be staged.
template <int PREC>
__device__ bool tile_nn(const GridDesc& g, const int* __restrict__ cs,
                        const float4* __restrict__ P, bool own, float qx, float qy, float qz,
                        float r, float margin, TileLds& L, float& best, int& bid, int& bj,
                        bool& cert) {
  const int lane = threadIdx.x & 63;
  cert = false;
  if (__ballot(own) == 0 || g.npts == 0) return true;
  TileBox B;
  const int staged = tile_stage(g, cs, P, own, qx, qy, qz, r, margin, L, B);
  if (staged == 0) return false;
  if (staged == 2 || B.M == 0) {
    cert = B.rr >= r * 1.00001f;
    return true;
  }
  const int M = B.M;
  const float lox = B.lox, hix = B.hix, loy = B.loy, hiy = B.hiy, loz = B.loz, hiz = B.hiz;
  const int x0 = B.x0, x1 = B.x1, y0 = B.y0, y1 = B.y1, z0 = B.z0, z1 = B.z1;
  // ---- centre, rounding bounds
  const float cx = 0.5f * (lox + hix), cy = 0.5f * (loy + hiy), cz = 0.5f * (loz + hiz);
  // every staged point and query lies in the cell box [x0, x1+1] x ... (h units)
  const float bx0 = g.ox + (float)x0 * g.h, bx1 = g.ox + (float)(x1 + 1) * g.h;
  const float by0 = g.oy + (float)y0 * g.h, by1 = g.oy + (float)(y1 + 1) * g.h;
  const float bz0 = g.oz + (float)z0 * g.h, bz1 = g.oz + (float)(z1 + 1) * g.h;
  const float ex_ = fmaxf(fmaxf(fabsf(bx0 - cx), fabsf(bx1 - cx)), fmaxf(fabsf(lox - cx), fabsf(hix - cx)));
  const float ey_ = fmaxf(fmaxf(fabsf(by0 - cy), fabsf(by1 - cy)), fmaxf(fabsf(loy - cy), fabsf(hiy - cy)));
  const float ez_ = fmaxf(fmaxf(fabsf(bz0 - cz), fabsf(bz1 - cz)), fmaxf(fabsf(loz - cz), fabsf(hiz - cz)));
  // a cell's extent carries rounding of (o + k h); one more cell of margin
  const float R = sqrtf(ex_ * ex_ + ey_ * ey_ + ez_ * ez_) * 1.001f + 1.8f * g.h;
  if (PREC == 1 && !(R < 1.0e4f)) return false;
  const float u = PREC == 1 ? 4.8828125e-4f : 3.90625e-3f;  // 2^-11, 2^-8
  const float e = (u + 5.9604645e-8f) * R + 5.2e-8f;
  const float E2 = 4.0e-6f * R * R;
  const float e2 = 2.0f * e;
  // ---- queries
  {
    const float ax = round_coord<PREC>(qx - cx), ay = round_coord<PREC>(qy - cy),
                az = round_coord<PREC>(qz - cz);
    L.q[lane] = make_float4(qx, qy, qz, 0.f);
    float n2 = ax * ax;
    n2 = n2 + ay * ay;
    n2 = n2 + az * az;
    L.qn[lane] = own ? n2 : NAN;
  }
  tile_lds_sync();
  const int grp = lane >> 4, col = lane & 15;
  const int ncb = (M + 15) >> 4;
  for (int rb = 0; rb < 4; ++rb) {
    // A: row rb*16 + col of this lane's group-0 copy; other groups hold zero k
    const float4 qa = L.q[rb * 16 + col];
    const float ax = qa.x - cx, ay = qa.y - cy, az = qa.z - cz;
    float qn[4], runmin[4], bd[4];
    int bi[4], bjj[4];
#pragma unroll
    for (int e4 = 0; e4 < 4; ++e4) {
      qn[e4] = L.qn[rb * 16 + 4 * grp + e4];
      runmin[e4] = INFINITY;
      bd[e4] = INFINITY;
      bi[e4] = 0x7fffffff;
      bjj[e4] = -1;
    }
    for (int cb = 0; cb < ncb; ++cb) {
      const int t = cb * 16 + col;
      const bool valid = t < M;
      const float4 p = L.p[valid ? t : 0];
      const float px = p.x - cx, py = p.y - cy, pz = p.z - cz;
      const lcr_f32x4 D = mfma_dot<PREC>(ax, ay, az, px, py, pz, grp == 0);
      const float rx = round_coord<PREC>(px), ry = round_coord<PREC>(py), rz = round_coord<PREC>(pz);
      float pn = rx * rx;
      pn = pn + ry * ry;
      pn = pn + rz * rz;
#pragma unroll
      for (int e4 = 0; e4 < 4; ++e4) {
        const float d2h = valid ? (qn[e4] + pn) - 2.0f * D[e4] : NAN;
        runmin[e4] = fminf(runmin[e4], d2h);  // NaN operands leave the minimum
        const float s1 = (sqrtf(fmaxf(runmin[e4], 0.f) + E2) + e2) * 1.000002f + e2;
        const float thr = s1 * s1 * 1.000002f + E2;
        if (d2h <= thr) {
          const float4 q = L.q[rb * 16 + 4 * grp + e4];
          const float d = flann_d2(q.x, q.y, q.z, p);
          const int id = f2i(p.w);
          if (d < bd[e4] || (d == bd[e4] && id < bi[e4])) {
            bd[e4] = d;
            bi[e4] = id;
            bjj[e4] = L.j[t];
          }
        }
      }
    }
    // lexicographic (d2, id) minimum over the 16 lanes of the group
#pragma unroll
    for (int e4 = 0; e4 < 4; ++e4) {
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        const float od = __shfl_xor(bd[e4], off, 64);
        const int oi = __shfl_xor(bi[e4], off, 64);
        const int oj = __shfl_xor(bjj[e4], off, 64);
        if (od < bd[e4] || (od == bd[e4] && oi < bi[e4])) {
          bd[e4] = od;
          bi[e4] = oi;
          bjj[e4] = oj;
        }
      }
      if (col == 0) {
        const int row = rb * 16 + 4 * grp + e4;
        L.rd[row] = bd[e4];
        L.rid[row] = bi[e4];
        L.rj[row] = bjj[e4];
      }
    }
  }
  tile_lds_sync();
  if (own) {
    const float d = L.rd[lane];
    const int id = L.rid[lane];
    if (d < best || (d == best && id < bid)) {
      best = d;
      bid = id;
      bj = L.rj[lane];
    }
    cert = tile_certified(B, qx, qy, qz, r, best);
  }
  return true;
}

}  // namespace lcr
