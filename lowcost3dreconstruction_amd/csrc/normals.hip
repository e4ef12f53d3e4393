// normals.hip -- K7: k-nearest-neighbour PCA normals.
//
// Replaces pcl::NormalEstimation<PointXYZRGB, Normal>::compute with setKSearch(K)
// (normal_estimation.cpp:84-108): per point, the K nearest neighbours including the
// point itself, their covariance, the eigenvector of the smallest eigenvalue as the
// normal, curvature |l0| / (l0 + l1 + l2), the normal flipped towards the viewpoint
// when (vp - p) . n < 0, then negated when the CLI's flip rule asks
// (normal_estimation.cpp:112-118).  Fewer than 3 neighbours -> NaN normal.
//
// One wavefront per query, 64 consecutive cell-sorted queries per wavefront:
//   1. gather the FLANN-form d2 of every point in rings 0..1 around the query's
//      cell into LDS (coalesced runs of cell-sorted points);
//   2. find the exact K-th smallest d2 by bisection over the float bit pattern
//      (count via wave ballots), resolving equal distances by original index;
//   3. certify: the K-th distance must be below the distance to the searched
//      block's boundary, else grow the block by one ring and repeat;
//   4. accumulate FP64 moments of the selected points relative to the query.
// After its 64 queries each lane solves one FP64 3x3 eigenproblem, so the solve
// runs 64-wide.  Blocks whose candidates overflow LDS take a slower exact path that
// rescans the block from memory for every bisection step.
#include <hip/hip_runtime.h>

#include "lcr_common.hpp"
#include "lcr_kernels.hpp"
#include "lcr_math.hpp"
#include "lcr_search.hpp"

namespace lcr {

#pragma clang fp contract(off)

namespace {

constexpr int kWaves = 4;
constexpr int kCap = 512;
constexpr int kSlots = kCap / 64;

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int popc_ballot(bool p) { return __popcll(__ballot(p)); }

struct Moments {
  double v[9];  // sum dx, dy, dz, dx2, dxdy, dxdz, dy2, dydz, dz2 (d = p - q)
};

__device__ __forceinline__ void add_moment(Moments& m, const float4& q, const float4& p) {
  const double dx = (double)p.x - (double)q.x;
  const double dy = (double)p.y - (double)q.y;
  const double dz = (double)p.z - (double)q.z;
  m.v[0] += dx;
  m.v[1] += dy;
  m.v[2] += dz;
  m.v[3] += dx * dx;
  m.v[4] += dx * dy;
  m.v[5] += dx * dz;
  m.v[6] += dy * dy;
  m.v[7] += dy * dz;
  m.v[8] += dz * dz;
}

// Calls f(j) for this lane's share of the points of rings 0..K.
template <class F>
__device__ __forceinline__ void for_block_points(const GridDesc& g, const int* cs, const CellPos& cp,
                                                 int K, F&& f) {
  const int lane = threadIdx.x & 63;
  for (int ring = 0; ring <= K; ++ring) {
    for_ring_runs(g, cp, ring, [&](int c0, int c1) {
      const int s = cs[c0], e = cs[c1 + 1];
      for (int j = s + lane; j < e; j += 64) f(j);
    });
  }
}

__device__ __forceinline__ int block_count(const GridDesc& g, const int* cs, const CellPos& cp, int K) {
  int total = 0;
  for (int ring = 0; ring <= K; ++ring)
    for_ring_runs(g, cp, ring, [&](int c0, int c1) { total += cs[c1 + 1] - cs[c0]; });
  return total;
}

// Exact path without an LDS buffer: every count rescans the block from memory.
__device__ void select_rescan(const GridDesc& g, const int* cs, const float4* pts, const float4& q,
                              const CellPos& cp, int K, int kk, Moments& acc) {
  unsigned T = 0;
  for (;; ++K) {
    const float Lc = block_gap(g, cp, K);
    const bool covers = (Lc == INFINITY);
    if (block_count(g, cs, cp, K) < kk && !covers) continue;
    unsigned lo = 0u, hi = 0x7f800000u;
    while (lo < hi) {
      const unsigned mid = lo + ((hi - lo) >> 1);
      int c = 0;
      for_block_points(g, cs, cp, K, [&](int j) {
        c += (__float_as_uint(flann_d2(q.x, q.y, q.z, pts[j])) <= mid) ? 1 : 0;
      });
      if (wave_sum_i(c) >= kk)
        hi = mid;
      else
        lo = mid + 1;
    }
    T = lo;
    const float L = (Lc - cp.slack) * g.h;
    if (covers || (L > 0.0f && __uint_as_float(T) < L * L * 0.99998f)) break;
  }
  // equal distances at the threshold: keep the smallest original indices
  int clt = 0, cle = 0;
  for_block_points(g, cs, cp, K, [&](int j) {
    const unsigned d = __float_as_uint(flann_d2(q.x, q.y, q.z, pts[j]));
    clt += d < T ? 1 : 0;
    cle += d <= T ? 1 : 0;
  });
  clt = wave_sum_i(clt);
  cle = wave_sum_i(cle);
  int idcut = 0x7fffffff;
  if (cle > kk) {
    const int need = kk - clt;
    unsigned lo = 0u, hi = 0x7fffffffu;
    while (lo < hi) {
      const unsigned mid = lo + ((hi - lo) >> 1);
      int c = 0;
      for_block_points(g, cs, cp, K, [&](int j) {
        const float4 p = pts[j];
        c += (__float_as_uint(flann_d2(q.x, q.y, q.z, p)) == T &&
              (unsigned)__float_as_int(p.w) <= mid)
                 ? 1
                 : 0;
      });
      if (wave_sum_i(c) >= need)
        hi = mid;
      else
        lo = mid + 1;
    }
    idcut = (int)lo;
  }
  for_block_points(g, cs, cp, K, [&](int j) {
    const float4 p = pts[j];
    const unsigned d = __float_as_uint(flann_d2(q.x, q.y, q.z, p));
    if (d < T || (d == T && __float_as_int(p.w) <= idcut)) add_moment(acc, q, p);
  });
}

__device__ void select_lds(const GridDesc& g, const int* cs, const float4* pts, const float4& q,
                           int kk, unsigned* sd, int* sj, Moments& acc) {
  const int lane = threadIdx.x & 63;
  const CellPos cp = cell_pos(g, q.x, q.y, q.z);
  int cnt = 0;
  bool ovf = false;
  auto gather = [&](int ring) {
    for_ring_runs(g, cp, ring, [&](int c0, int c1) {
      if (ovf) return;
      const int s = cs[c0], e = cs[c1 + 1], len = e - s;
      if (cnt + len > kCap) {
        ovf = true;
        return;
      }
      for (int t = lane; t < len; t += 64) {
        sd[cnt + t] = __float_as_uint(flann_d2(q.x, q.y, q.z, pts[s + t]));
        sj[cnt + t] = s + t;
      }
      cnt += len;
    });
    wave_lds_sync();
  };
  gather(0);
  gather(1);
  int K = 1;
  unsigned slot[kSlots];
  unsigned T = 0;
  for (;;) {
    if (ovf) {
      select_rescan(g, cs, pts, q, cp, K, kk, acc);
      return;
    }
    const float Lc = block_gap(g, cp, K);
    const bool covers = (Lc == INFINITY);
    if (cnt >= kk) {
      const int nslots = (cnt + 63) >> 6;
#pragma unroll
      for (int m = 0; m < kSlots; ++m) {
        const int i = lane + 64 * m;
        slot[m] = (m < nslots && i < cnt) ? sd[i] : 0xffffffffu;
      }
      unsigned lo = 0u, hi = 0x7f800000u;
      while (lo < hi) {
        const unsigned mid = lo + ((hi - lo) >> 1);
        int c = 0;
#pragma unroll
        for (int m = 0; m < kSlots; ++m)
          if (m < nslots) c += popc_ballot(slot[m] <= mid);
        if (c >= kk)
          hi = mid;
        else
          lo = mid + 1;
      }
      T = lo;
      const float L = (Lc - cp.slack) * g.h;
      if (covers || (L > 0.0f && __uint_as_float(T) < L * L * 0.99998f)) break;
    }
    ++K;
    gather(K);
  }
  const int nslots = (cnt + 63) >> 6;
  int clt = 0, cle = 0;
#pragma unroll
  for (int m = 0; m < kSlots; ++m)
    if (m < nslots) {
      clt += popc_ballot(slot[m] < T);
      cle += popc_ballot(slot[m] <= T);
    }
  int idcut = 0x7fffffff;
  if (cle > kk) {
    const int need = kk - clt;
    int ids[kSlots];
#pragma unroll
    for (int m = 0; m < kSlots; ++m) {
      const int i = lane + 64 * m;
      ids[m] = (m < nslots && i < cnt && slot[m] == T) ? __float_as_int(pts[sj[i]].w) : 0x7fffffff;
    }
    unsigned lo = 0u, hi = 0x7fffffffu;
    while (lo < hi) {
      const unsigned mid = lo + ((hi - lo) >> 1);
      int c = 0;
#pragma unroll
      for (int m = 0; m < kSlots; ++m)
        if (m < nslots) c += popc_ballot(slot[m] == T && (unsigned)ids[m] <= mid);
      if (c >= need)
        hi = mid;
      else
        lo = mid + 1;
    }
    idcut = (int)lo;
  }
#pragma unroll
  for (int m = 0; m < kSlots; ++m) {
    const int i = lane + 64 * m;
    if (m < nslots && i < cnt) {
      const unsigned d = slot[m];
      if (d < T || d == T) {
        const float4 p = pts[sj[i]];
        if (d < T || __float_as_int(p.w) <= idcut) add_moment(acc, q, p);
      }
    }
  }
}

__global__ void __launch_bounds__(kWaves * 64)
k_normals(const GridDesc* __restrict__ grids, const int* __restrict__ cell_start,
          const float4* __restrict__ pts, const int* __restrict__ cloud_of_sorted, int64_t total,
          int k, const float* __restrict__ vp, int flip_all, float4* __restrict__ out) {
  __shared__ unsigned s_d[kWaves][kCap];
  __shared__ int s_j[kWaves][kCap];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t base = ((int64_t)blockIdx.x * kWaves + w) * 64;
  if (base >= total) return;
  Moments mine;
#pragma unroll
  for (int i = 0; i < 9; ++i) mine.v[i] = 0.0;
  int mine_n = 0;
  for (int qi = 0; qi < 64; ++qi) {
    const int64_t j = base + qi;
    if (j >= total) break;
    const float4 q = pts[j];
    const GridDesc g = grids[cloud_of_sorted[j]];
    const int kk = min(k, g.npts);
    Moments acc;
#pragma unroll
    for (int i = 0; i < 9; ++i) acc.v[i] = 0.0;
    if (kk >= 3) select_lds(g, cell_start + g.cell_off, pts, q, kk, s_d[w], s_j[w], acc);
#pragma unroll
    for (int i = 0; i < 9; ++i) acc.v[i] = wave_sum(acc.v[i]);
    if (lane == qi) {
      mine = acc;
      mine_n = kk;
    }
    wave_lds_sync();
  }
  const int64_t jq = base + lane;
  if (jq >= total) return;
  const float4 q = pts[jq];
  float4 o;
  if (mine_n < 3) {
    o = make_float4(NAN, NAN, NAN, NAN);
  } else {
    const double n = (double)mine_n;
    const double mx = mine.v[0] / n, my = mine.v[1] / n, mz = mine.v[2] / n;
    double C[3][3];
    C[0][0] = mine.v[3] / n - mx * mx;
    C[0][1] = mine.v[4] / n - mx * my;
    C[0][2] = mine.v[5] / n - mx * mz;
    C[1][1] = mine.v[6] / n - my * my;
    C[1][2] = mine.v[7] / n - my * mz;
    C[2][2] = mine.v[8] / n - mz * mz;
    C[1][0] = C[0][1];
    C[2][0] = C[0][2];
    C[2][1] = C[1][2];
    double wv[3], V[3][3];
    eigh3(C, wv, V);
    double nx = V[0][0], ny = V[1][0], nz = V[2][0];
    const double sum = wv[0] + wv[1] + wv[2];
    const double curv = sum != 0.0 ? fabs(wv[0] / sum) : 0.0;
    const float* v = vp + 3 * cloud_of_sorted[jq];
    const double dot = ((double)v[0] - q.x) * nx + ((double)v[1] - q.y) * ny + ((double)v[2] - q.z) * nz;
    if (dot < 0.0) {
      nx = -nx;
      ny = -ny;
      nz = -nz;
    }
    if (flip_all) {
      nx = -nx;
      ny = -ny;
      nz = -nz;
    }
    o = make_float4((float)nx, (float)ny, (float)nz, (float)curv);
  }
  out[jq] = o;
}

}  // namespace

void launch_normals(const GridDesc* grids, const int* cell_start, const float4* pts,
                    const int* cloud_of_sorted, int64_t npts_total, int k, const float* vp,
                    int flip_all, float4* nrm_sorted, hipStream_t s) {
  if (npts_total <= 0) return;
  const int64_t per_block = (int64_t)kWaves * 64;
  const unsigned blocks = (unsigned)((npts_total + per_block - 1) / per_block);
  hipLaunchKernelGGL(k_normals, dim3(blocks), dim3(kWaves * 64), 0, s, grids, cell_start, pts,
                     cloud_of_sorted, npts_total, k, vp, flip_all, nrm_sorted);
}

}  // namespace lcr
