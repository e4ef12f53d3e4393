// icp.hip -- K2..K6 and K8: one ICP iteration for every active instance of a
// batch in two launches (search, then fold and solve), the fitness pass, and the
// output transform.
//
// Replaces IterativeClosestPoint::computeTransformation's loop body
// (pcl/registration/impl/icp.hpp; driven by icp.align at fine_registration.cpp:121):
//   determineCorrespondences   -> exact grid 1-NN per source point (lcr_search.hpp)
//   estimateRigidTransformation-> FP64 moments per tile, folded per instance in a
//                                 fixed order, Horn (point-to-point) or Cholesky
//                                 (point-to-plane) in FP64 (lcr_math.hpp)
//   final = T * final          -> FP64 composition on the device
//   convergence criteria       -> DefaultConvergenceCriteria on the device
// and Registration::getFitnessScore (fine_registration.cpp:126).
//
// Work decomposition: a tile is `TILE` consecutive cell-sorted source points of
// one instance (one point per thread).  The launch is persistent: a fixed number
// of blocks (a multiple of 8) loops over tiles.  Hardware places block b on XCD
// b % 8, so each XCD is given one contiguous eighth of the tile range: the tiles
// of an instance, and the target grid they search, then stay in that XCD's L2.
// Each block writes its tile's FP64 sums into a per-tile partial; a second
// launch (one block per instance) folds the partials in a fixed order
// (deterministic, no floating-point atomics), solves and updates the instance.
// Converged instances are skipped, so the iteration can be replayed a fixed
// number of times inside a hipGraph.
#include <hip/hip_runtime.h>
#include <float.h>

#include "lcr_common.hpp"
#include "lcr_kernels.hpp"
#include "lcr_math.hpp"
#include "lcr_search.hpp"
#include "tile_search.hpp"

namespace lcr {

#pragma clang fp contract(off)

namespace {

// x' = (float)(((T0*x + T1*y) + T2*z) + T3) in FP64, no FMA: the source is always
// transformed from its original coordinates by the FP64-composed transform.
__device__ __forceinline__ void xform(const double* T, const float4& s, float& qx, float& qy,
                                      float& qz) {
  const double x = s.x, y = s.y, z = s.z;
  double v;
  v = T[0] * x;
  v = v + T[1] * y;
  v = v + T[2] * z;
  v = v + T[3];
  qx = (float)v;
  v = T[4] * x;
  v = v + T[5] * y;
  v = v + T[6] * z;
  v = v + T[7];
  qy = (float)v;
  v = T[8] * x;
  v = v + T[9] * y;
  v = v + T[10] * z;
  v = v + T[11];
  qz = (float)v;
}

// False only when the query lies in a grid cell that the near mask marks as
// holding no point within GridDesc::near_r; queries outside the grid return true
// (the search handles them).
__device__ __forceinline__ bool near_cell(const GridDesc& g, const uint32_t* __restrict__ near,
                                          float qx, float qy, float qz) {
  const CellPos c = cell_pos(g, qx, qy, qz);
  if (c.cx < 0 || c.cy < 0 || c.cz < 0 || c.cx >= g.nx || c.cy >= g.ny || c.cz >= g.nz) return true;
  const uint32_t w =
      near[g.near_off + ((int64_t)c.cz * g.ny + c.cy) * near_row_words(g.nx) + (c.cx >> 5)];
  return ((w >> (c.cx & 31)) & 1u) != 0;
}

struct TileRange {
  int first, end, step;
};

// Persistent mapping (gridDim.x is a multiple of 8).  XCD-contiguous: hardware
// places block b on XCD b % 8, and XCD x takes the x-th eighth of the tile range,
// so an instance's tiles (and its target grid) stay in one L2.  Interleaved:
// plain grid stride, so the tiles of the few instances still active late in a
// batch spread over the whole device.
__device__ __forceinline__ TileRange xcd_tiles(int ntiles, int interleave) {
  if (interleave) return TileRange{(int)blockIdx.x, ntiles, (int)gridDim.x};
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3, per = gridDim.x >> 3;
  const int t0 = (int)(((int64_t)ntiles * xcd) >> 3);
  const int t1 = (int)(((int64_t)ntiles * (xcd + 1)) >> 3);
  return TileRange{t0 + slot, t1, per};
}

// The tile's FP64 sums, written to out[0..NV): the correspondence count, the
// moments of the estimator and the sum of d2.  Point-to-point: q, t and q t^T.
// Point-to-plane (transformation_estimation_point_to_plane_lls.hpp): the upper
// triangle of A^T A and A^T b for rows [s x n ; n], d = n . (dst - src), in
// double; a correspondence whose normal is not finite adds only to the count and
// the d2 sum.  Each value is formed and summed over the wavefront in turn (fixed
// order), so only the per-lane inputs stay live; the waves' sums are added in
// wave order.
//
// tile_moments_emit writes each wavefront's sums to lds[wid * NV + v] (every lane
// of the wavefront calls it); tile_moments_store adds the NW wavefronts' sums in
// wave order into out.
template <bool P2PL>
__device__ __forceinline__ void tile_moments_emit(bool has, float qx, float qy, float qz,
                                                  const float4& t, const float4& n, float d2,
                                                  double* lds) {
  constexpr int NV = P2PL ? kP2LVals : kP2PVals;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int k = 0;
  auto emit = [&](double x) {
    const double sum = wave_sum_dpp(x);
    if (lane == 0) lds[wid * NV + k] = sum;
    ++k;
  };
  emit(has ? 1.0 : 0.0);
  if constexpr (!P2PL) {
    const double q[3] = {has ? (double)qx : 0.0, has ? (double)qy : 0.0, has ? (double)qz : 0.0};
    const double tt[3] = {has ? (double)t.x : 0.0, has ? (double)t.y : 0.0, has ? (double)t.z : 0.0};
#pragma unroll
    for (int a = 0; a < 3; ++a) emit(q[a]);
#pragma unroll
    for (int a = 0; a < 3; ++a) emit(tt[a]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int c = 0; c < 3; ++c) emit(q[a] * tt[c]);
  } else {
    double row[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, d = 0.0;
    if (has && isfinite(n.x) && isfinite(n.y) && isfinite(n.z)) {
      const double sx = qx, sy = qy, sz = qz, nx = n.x, ny = n.y, nz = n.z;
      row[0] = nz * sy - ny * sz;
      row[1] = nx * sz - nz * sx;
      row[2] = ny * sx - nx * sy;
      row[3] = nx;
      row[4] = ny;
      row[5] = nz;
      d = nx * (double)t.x + ny * (double)t.y + nz * (double)t.z - nx * sx - ny * sy - nz * sz;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = i; j < 6; ++j) emit(row[i] * row[j]);
#pragma unroll
    for (int i = 0; i < 6; ++i) emit(row[i] * d);
  }
  emit(has ? (double)d2 : 0.0);
}

template <int NV, int NW>
__device__ __forceinline__ void tile_moments_store(const double* lds, double* out) {
  if (threadIdx.x < NV) {
    double sum = lds[threadIdx.x];
#pragma unroll
    for (int w = 1; w < NW; ++w) sum += lds[w * NV + threadIdx.x];
    out[threadIdx.x] = sum;
  }
}

template <bool P2PL, int TILE>
__device__ __forceinline__ void tile_moments(bool has, float qx, float qy, float qz, const float4& t,
                                             const float4& n, float d2, double* lds, double* out) {
  constexpr int NV = P2PL ? kP2LVals : kP2PVals;
  tile_moments_emit<P2PL>(has, qx, qy, qz, t, n, d2, lds);
  __syncthreads();
  tile_moments_store<NV, TILE / 64>(lds, out);
  __syncthreads();  // lds is reused by the next tile
}

// PREC 0: exact per-lane grid search; 1 / 2: FP16 / BF16 matrix-core distance
// tiles with exact FP32 rerank; 3: staged tiles scanned with exact FP32 distances
// (tile_search.hpp; for 1-3 TILE is 64, one wavefront).
//
// Seven waves per SIMD (72 VGPRs, a few spills) measured 13 % faster over a
// cycle's iterations than the unconstrained five (87 VGPRs), six 9 %, eight
// (64 VGPRs, more spills) 8 %: the search is bound by the latency of its
// dependent loads, and more resident waves hide more of it.
//
// RING: every instance's radius spans at most kRingOnlyCells target cells
// (IcpConst::ring_only), so the bounded search ends inside the ring loop and the
// tree walk is never needed; this variant carries none of its code or registers.
template <bool P2PL, int TILE, int PREC, bool RING = false>
__global__ void __launch_bounds__(TILE, 7) k_icp_iter(const IcpBuffers b, const IcpConst c) {
  constexpr int NV = P2PL ? kP2LVals : kP2PVals;
  static_assert(PREC == 0 || TILE == 64, "distance tiles need one wavefront per tile");
  static_assert(TILE >= NV, "one thread per value writes the tile's sums");
  __shared__ double lds[NV * (TILE / 64)];
  const int nt = b.it_count ? *b.it_count : b.ntiles;
  // order 2: XCD-contiguous while at least half of the tiles are live (each
  // XCD's L2 then holds the targets of its own eighth of the pairs), grid stride
  // once fewer instances are left (they would crowd a few XCDs)
  const int inter = c.interleave == 2 ? (2 * nt < b.ntiles ? 1 : 0) : c.interleave;
  const TileRange tr = xcd_tiles(nt, inter);
  for (int it = tr.first; it < tr.end; it += tr.step) {
    const int2 e = b.it_list ? b.it_list[it] : make_int2(it, b.tile_pair[it]);
    const int tile = e.x, p = e.y;
    const PairState* st = b.state + p;
    if (st->active == 0) continue;  // uniform across the block
    const PairDesc pd = b.pairs[p];
    const GridDesc gs = b.grids[pd.src];
    const GridDesc gt = b.grids[pd.tgt];
    const int rank = (tile - pd.tile_begin) * TILE + (int)threadIdx.x;
    const bool own = rank < gs.npts;
    const float r = (float)sqrt(pd.max_corr_d2);
    const bool bounded = isfinite(r);
    // the previous match of the tile's lane (the final fitness pass takes the
    // same points in the same lanes)
    const int64_t slot = (int64_t)tile * TILE + threadIdx.x;
    float qx = 0.f, qy = 0.f, qz = 0.f;
    NNResult nn{INFINITY, 0x7fffffff, -1, false};
    if (own) xform(st->T, b.morton ? b.tpts[gs.pt_off + rank] : b.pts[gs.pt_off + rank], qx, qy, qz);
    // no target point within r of the query's cell: no correspondence, no search
    const bool empty = own && bounded && b.near && gt.near_r >= r &&
                       !near_cell(gt, b.near, qx, qy, qz);
    bool search = own && !empty;  // lanes that still need the scalar search
    int seed = (search && b.prev) ? b.prev[slot] : -1;
    if constexpr (PREC != 0) {
      __shared__ TileLds tl;
      const bool q_ok = search && isfinite(qx) && isfinite(qy) && isfinite(qz);
      bool cert = false, tiled;
      if constexpr (PREC == 3)
        tiled = bounded && tile_nn_fp32(gt, b.cell_start + gt.cell_off, b.pts, q_ok, qx, qy, qz, r,
                                        0.75f * gt.h, tl, nn.d2, nn.id, nn.j, cert);
      else
        tiled = bounded && tile_nn<PREC>(gt, b.cell_start + gt.cell_off, b.pts, q_ok, qx, qy, qz, r,
                                         0.75f * gt.h, tl, nn.d2, nn.id, nn.j, cert);
      if (tiled) {
        if (q_ok && cert) search = false;
        if (nn.j >= 0) seed = nn.j;
      }
    }
    if constexpr (RING) {
      if (search)
        nn = nn_query(gt, b.cell_start + gt.cell_off, b.pts, Tree{nullptr, nullptr, nullptr}, qx,
                      qy, qz, r, bounded, seed, 2, true);
    } else {
      if (search)
        nn = nn_query(gt, b.cell_start + gt.cell_off, b.pts, b.tree(), qx, qy, qz, r, bounded,
                      seed, 2, true);
      const float rlim = r * 1.00001f;
      // tree walks happen only for radii of more than kRingOnlyCells cells: kept out
      // of the hot path's code layout
      if (__builtin_expect(__ballot(nn.tree) != 0, 0))
        bvh_nn_wave(gt, b.tree(), nn.tree, qx, qy, qz, bounded ? rlim * rlim : INFINITY, nn.d2,
                    nn.id, nn.j);
    }
    bool has = false;
    float4 t = make_float4(0.f, 0.f, 0.f, 0.f), n = t;
    if constexpr (TILE == 64) {
      // the tile has a point without a match: the final fitness pass orders by it
      const uint64_t unmatched = __ballot(own && nn.j < 0);
      if (b.fit_flags && threadIdx.x == 0) b.fit_flags[tile] = unmatched != 0 ? 1 : 0;
    }
    if (own) {
      if (b.prev) b.prev[slot] = nn.j;
      if (nn.j >= 0 && (double)nn.d2 <= pd.max_corr_d2) {
        has = true;
        t = b.pts[nn.j];
        if (P2PL) n = b.nrm[nn.j];
      }
    }
    tile_moments<P2PL, TILE>(has, qx, qy, qz, t, n, nn.d2, lds, b.partials + (int64_t)tile * kMaxVals);
  }
}

// The ring-only search for launches with few live tiles.  Each query gets
// kSplitLanes lanes; each lane scans every kSplitLanes-th (z, y) row of a ring
// (ring_rows), and the lanes take the minimum by (d2, id) after every ring, so the
// result is the one-lane search's.  A block of kSplitLanes wavefronts takes one
// tile, 64 queries per pass (64 / kSplitLanes per wavefront); the TILE results go
// through LDS to the first TILE threads, which form the tile's sums exactly as
// k_icp_iter<TILE> does (query i in lane i % 64 of wavefront i / 64).  With a few
// hundred tiles live the one-lane kernel leaves most SIMDs idle, and each
// wavefront waits on its longest chain of row loads; this one runs kSplitLanes
// times as many wavefronts, each with about 1 / kSplitLanes of the rows.
constexpr int kSplitLanes = 4;

template <int S>
__device__ __forceinline__ void lanes_min(float& best, int& bid, int& bj) {
#pragma unroll
  for (int m = 1; m < S; m <<= 1) {
    const float ob = __shfl_xor(best, m, 64);
    const int oid = __shfl_xor(bid, m, 64);
    const int oj = __shfl_xor(bj, m, 64);
    const bool take = ob < best || (ob == best && oid < bid);
    best = take ? ob : best;
    bid = take ? oid : bid;
    bj = take ? oj : bj;
  }
}

// Bounded ring-only 1-NN over S lanes (sub = this lane's index among them).  The S
// lanes of a query are consecutive and reach every lanes_min together: their
// early returns and ring loop depend only on the query and the shared best.
template <int S>
__device__ __forceinline__ NNResult nn_ring_split(const GridDesc& g, const int* __restrict__ cs,
                                                  const float4* __restrict__ P, float qx, float qy,
                                                  float qz, float r, int seed_j, int sub) {
  NNResult res{INFINITY, 0x7fffffff, -1, false};
  if (g.npts == 0 || !(isfinite(qx) && isfinite(qy) && isfinite(qz))) return res;
  const CellPos c = cell_pos(g, qx, qy, qz);
  const float rlim = r * 1.00001f;
  const int k0 = ring_start(g, c, rlim);
  if (k0 < 0) return res;
  const int kmax = k0 + max(g.nx, max(g.ny, g.nz));
  float best = INFINITY;
  int bid = 0x7fffffff, bj = -1;
  if (seed_j >= 0) scan_points(P, seed_j, seed_j + 1, qx, qy, qz, best, bid, bj);
  const float h2 = g.h * g.h, lim2 = rlim * rlim;
  for (int k = k0; k <= kmax; ++k) {
    ring_rows(g, cs, P, c, k, sub, S, h2, lim2, qx, qy, qz, best, bid, bj);
    lanes_min<S>(best, bid, bj);
    if (ring_stop(g, c, k, best, rlim)) break;
  }
  res.d2 = best;
  res.id = bid;
  res.j = bj;
  return res;
}

template <bool P2PL, int TILE>
__global__ void __launch_bounds__(64 * kSplitLanes) k_icp_iter_split(const IcpBuffers b,
                                                                     const IcpConst c) {
  constexpr int S = kSplitLanes, NV = P2PL ? kP2LVals : kP2PVals;
  static_assert(64 % S == 0 && (S & (S - 1)) == 0, "lanes per query: a power of two dividing 64");
  static_assert(TILE % 64 == 0 && TILE <= 64 * S, "one thread per query forms the sums");
  __shared__ double lds[NV * (TILE / 64)];
  __shared__ float s_d2[TILE];
  __shared__ int s_j[TILE];
  const int nt = b.it_count ? *b.it_count : b.ntiles;
  const int q = (int)threadIdx.x / S, sub = (int)threadIdx.x % S;
  for (int it = blockIdx.x; it < nt; it += gridDim.x) {
    const int2 e = b.it_list ? b.it_list[it] : make_int2(it, b.tile_pair[it]);
    const int tile = e.x, p = e.y;
    const PairState* st = b.state + p;
    if (st->active == 0) continue;  // uniform across the block
    const PairDesc pd = b.pairs[p];
    const GridDesc gs = b.grids[pd.src];
    const GridDesc gt = b.grids[pd.tgt];
    const int base = (tile - pd.tile_begin) * TILE;
    // every radius is finite here (IcpConst::ring_only)
    const float r = (float)sqrt(pd.max_corr_d2);
    // a query's point, in tree or cell order
    auto point_of = [&](int rank) { return b.morton ? b.tpts[gs.pt_off + rank] : b.pts[gs.pt_off + rank]; };
    for (int pass = 0; pass < TILE / 64; ++pass) {
      const int qi = pass * 64 + q;
      const bool own = base + qi < gs.npts;
      float qx = 0.f, qy = 0.f, qz = 0.f;
      NNResult nn{INFINITY, 0x7fffffff, -1, false};
      if (own) xform(st->T, point_of(base + qi), qx, qy, qz);
      const bool empty = own && b.near && gt.near_r >= r && !near_cell(gt, b.near, qx, qy, qz);
      if (own && !empty) {
        const int seed = b.prev ? b.prev[(int64_t)tile * TILE + qi] : -1;
        nn = nn_ring_split<S>(gt, b.cell_start + gt.cell_off, b.pts, qx, qy, qz, r, seed, sub);
      }
      if (sub == 0) {
        s_d2[qi] = nn.d2;
        s_j[qi] = nn.j;
      }
    }
    __syncthreads();
    if (threadIdx.x < TILE) {
      const bool own = base + (int)threadIdx.x < gs.npts;
      float qx = 0.f, qy = 0.f, qz = 0.f;
      if (own) xform(st->T, point_of(base + (int)threadIdx.x), qx, qy, qz);
      const float d2 = s_d2[threadIdx.x];
      const int j = s_j[threadIdx.x];
      bool has = false;
      float4 t = make_float4(0.f, 0.f, 0.f, 0.f), n = t;
      if constexpr (TILE == 64) {
        // as in k_icp_iter (the first wavefront holds the tile's 64 queries)
        const uint64_t unmatched = __ballot(own && j < 0);
        if (b.fit_flags && threadIdx.x == 0) b.fit_flags[tile] = unmatched != 0 ? 1 : 0;
      }
      if (own) {
        if (b.prev) b.prev[(int64_t)tile * TILE + threadIdx.x] = j;
        if (j >= 0 && (double)d2 <= pd.max_corr_d2) {
          has = true;
          t = b.pts[j];
          if (P2PL) n = b.nrm[j];
        }
      }
      tile_moments_emit<P2PL>(has, qx, qy, qz, t, n, d2, lds);
    }
    __syncthreads();
    tile_moments_store<NV, TILE / 64>(lds, b.partials + (int64_t)tile * kMaxVals);
    __syncthreads();  // lds, s_d2 and s_j are reused by the next tile
  }
}

// Second half of an iteration, one block per instance: folds the instance's tile
// partials in a fixed order, then the FP64 solve, the composition final = Tinc *
// final and the convergence step.  Kept out of k_icp_iter so that the search
// kernel carries none of the solver's registers.
//
// The fold reads the partials row by row: lane i of each 32-lane row takes value
// i, row r tiles r, r + kSolveRows, ... (kSolveUnroll loads in flight), so one
// load instruction covers two tiles' contiguous values.  Each value is then the
// sum of the rows' sums in row order.  (One thread per tile reading all of its
// values touched 64 cache lines per instruction: 12 of the 15.5 us of a
// 1611-tile instance went to that fold and a block sum.)  Rows and row sums add
// with compensation (CompSum): a plain sum along ~50 tiles carried several times
// the rounding of the former tree, enough to move a run's final pose by a few
// nanometres between tile sizes.
constexpr int kSolveThreads = 1024;

// Compensated (Kahan) summation: the running sum's rounding error is carried
// and subtracted from the next term, so a row of ~50 tiles and the 32 rows add
// with an error of a few ulps whatever their count and order.  (Every operation
// is rounded as written: the library is built without contraction.)
struct CompSum {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double x) {
    const double y = x - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return s - c; }
};
constexpr int kSolveRows = kSolveThreads / 32;
constexpr int kSolveUnroll = 8;
constexpr int kFoldThreads = 256;

template <bool P2PL>
__global__ void __launch_bounds__(kSolveThreads) k_icp_solve(const IcpBuffers b, const IcpConst c) {
  constexpr int NV = P2PL ? kP2LVals : kP2PVals;
  static_assert(NV <= 32, "one value per lane of a row");
  __shared__ double part[kSolveRows][32];
  __shared__ double tot[32];
  PairState* st = b.state + blockIdx.x;
  if (st->active == 0) return;  // uniform across the block
  const PairDesc pd = b.pairs[blockIdx.x];
  // the state, read now so that its latency hides behind the fold
  PairState ls = *st;
  const int i = threadIdx.x & 31, row = threadIdx.x >> 5;
  CompSum acc;
  if (i < NV) {
    const double* in = b.partials + (int64_t)pd.tile_begin * kMaxVals + i;
    int t = row;
    for (; t + (kSolveUnroll - 1) * kSolveRows < pd.ntiles; t += kSolveUnroll * kSolveRows) {
      double x[kSolveUnroll];
#pragma unroll
      for (int u = 0; u < kSolveUnroll; ++u) x[u] = in[(int64_t)(t + u * kSolveRows) * kMaxVals];
#pragma unroll
      for (int u = 0; u < kSolveUnroll; ++u) acc.add(x[u]);
    }
    for (; t < pd.ntiles; t += kSolveRows) acc.add(in[(int64_t)t * kMaxVals]);
  }
  part[row][i] = acc.value();
  __syncthreads();
  if (threadIdx.x < NV) {
    CompSum sum;
#pragma unroll 8
    for (int r = 0; r < kSolveRows; ++r) sum.add(part[r][threadIdx.x]);
    tot[threadIdx.x] = sum.value();
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = tot[k];
  const double ncorr = v[0];
  ls.ncorr = (int)ncorr;
  if (ncorr < 3.0) {
    // icp.hpp: fewer than min_number_correspondences_ (3) -> not converged
    ls.state = kNoCorrespondences;
    ls.converged = 0;
    ls.active = 0;
    *st = ls;
    return;
  }
  double Tinc[16];
  if constexpr (P2PL) {
    ls.solve_failed = p2pl_increment(v, Tinc) ? 0 : 1;
  } else {
    p2p_increment(v, Tinc);
  }
  double Tn[16];
  mat4_mul(Tinc, ls.T, Tn);
#pragma unroll
  for (int k = 0; k < 16; ++k) ls.T[k] = Tn[k];
  ls.iterations += 1;
  const double mse = v[NV - 1] / ncorr;
  ls.last_mse = mse;
  if (conv_step(&ls, c, Tinc, mse)) {
    ls.converged = 1;
    ls.active = 0;
  }
  *st = ls;
}

// per_instance: each instance's own max_corr_d2 is the range.  A finite range
// bounds the search at sqrt(range): only points within it are counted.
// Six waves per SIMD (80 VGPRs, a few bytes spilled) measured 3 % faster than
// the unconstrained five; eight (64 VGPRs) spilled in the walk and was slower.
//
// With b.fetch and 64-point tiles (one wavefront per block), the tiles are dealt
// to kFetchGroups groups of blocks (group g: blocks g, g + kFetchGroups, ...; its
// tiles g, g + kFetchGroups, ..., a sample of every part of the list), and the
// blocks of a group take its tiles from the group's counter as they finish the
// last, instead of fixed shares.  A tile whose points go to the tree walk costs
// tens of times a seeded one, and with about six tiles per wavefront the fixed
// shares left the mean wavefront idle for half of the pass.  One counter per
// group on its own cache line: atomics on one address serialize, and eight
// counters shared by every block cost more than the balance gained.  Each tile
// writes its own partials, so the order does not change the result.  Larger
// tiles keep fixed shares.
template <int TILE>
__global__ void __launch_bounds__(TILE, 6) k_fitness(const IcpBuffers b, double max_range_arg,
                                                  int per_instance, int interleave) {
  __shared__ double lds[2 * (TILE / 64)];
  // the gate's score (per_instance) counts a fixed subsample of the source's
  // tiles, the same source points for every hypothesis of a pair: the tiles
  // listed in gate_tiles (a pass over every tile that skipped the others spent
  // 320 us of a cycle step on skipping; over the list it takes 190 us)
  // the full pass runs over the tiles of the instances the gate kept (fit_list)
  const bool listed = per_instance && b.gate_tiles;
  const bool kept = !per_instance && b.fit_count;
  const int nt = listed ? b.ngate_tiles : (kept ? *b.fit_count : b.ntiles);
  auto tile_pass = [&](int it) {
    const int tile = listed ? b.gate_tiles[it] : (kept ? b.fit_list[it].x : it);
    const int p = b.tile_pair[tile];
    PairState* st = b.state + p;
    if (st->gated) return;  // uniform across the block
    const PairDesc pd = b.pairs[p];
    const int stride = per_instance ? gate_stride(pd.ntiles) : 1;
    if ((tile - pd.tile_begin) % stride != 0) return;  // uniform across the block
    const GridDesc gs = b.grids[pd.src];
    const GridDesc gt = b.grids[pd.tgt];
    const double max_range = per_instance ? pd.max_corr_d2 : max_range_arg;
    const bool bounded = max_range < DBL_MAX && max_range >= 0.0;
    const float rq = bounded ? (float)sqrt(max_range) : INFINITY;
    const float r = rq * 1.00001f;
    // the point's rank in its instance's source: in cell order for the gate, in
    // the cloud's tree (Morton) order for the final pass (as in the iterations),
    // whose tiles are then compact patches (a tile of cell order can run along a
    // whole row of cells, 20 cm on cycle36, and its far points' tree walks then
    // share little)
    const int rank = (tile - pd.tile_begin) * TILE + (int)threadIdx.x;
    const bool morton = !per_instance && b.morton;
    double v[2] = {0.0, 0.0};
    const bool own = rank < gs.npts;
    float qx = 0.f, qy = 0.f, qz = 0.f;
    NNResult nn{INFINITY, 0x7fffffff, -1, false};
    if (own) {
      xform(st->T, morton ? b.tpts[gs.pt_off + rank] : b.pts[gs.pt_off + rank], qx, qy, qz);
      const bool empty = bounded && b.near && gt.near_r >= rq && !near_cell(gt, b.near, qx, qy, qz);
      // the lane's previous match: the iterations take the same points in the same
      // lanes as the final pass (the gate's pass runs before any match exists)
      const int seed = b.prev ? b.prev[(int64_t)tile * TILE + threadIdx.x] : -1;
      // the final pass over refined instances: a point without a match within r in
      // the last iteration (seed -1) is farther than about r from the target; a
      // ring search would scan two unpruned rings before handing it to the tree,
      // so it goes to the tree walk at once
      if (!bounded && b.prev && seed < 0)
        nn.tree = true;
      else if (!empty)
        nn = nn_query(gt, b.cell_start + gt.cell_off, b.pts, b.tree(), qx, qy, qz, r, bounded, seed,
                      2, true);
    }
    bvh_nn_wave(gt, b.tree(), nn.tree, qx, qy, qz, bounded ? r * r : INFINITY, nn.d2, nn.id,
                nn.j);
    if (own && nn.j >= 0 && (double)nn.d2 <= max_range) {
      v[0] = nn.d2;
      v[1] = 1.0;
    }
    block_sum<2>(v, lds);
    if (threadIdx.x == 0) {
      double* out = b.partials + (int64_t)tile * kMaxVals;
      out[0] = v[0];
      out[1] = v[1];
    }
  };
  if (TILE == 64 && b.fetch) {
    const int ng = min((int)gridDim.x, kFetchGroups);
    const int g = (int)blockIdx.x % ng;
    int* ctr = b.fetch + g * kFetchStride;
    for (;;) {
      // one ticket per tile: lane 0 adds 1 and every other lane 0, and each lane
      // takes lane 0's old value into a scalar register.  No lane-0-only branch
      // and no LDS broadcast: with those the compiler turned this loop's exits
      // into per-lane ones, and lanes 1-63 repeated a tile while lane 0 waited
      // for them to fetch the next.
      const int ticket = atomicAdd(ctr, threadIdx.x == 0 ? 1 : 0);
      const int it = g + ng * __builtin_amdgcn_readfirstlane(ticket);
      if (it >= nt) break;
      tile_pass(it);
    }
  } else {
    const TileRange tr = xcd_tiles(nt, interleave);
    for (int it = tr.first; it < tr.end; it += tr.step) tile_pass(it);
  }
}

// The final pass's list of tiles (fit_list) reordered into out: tiles with a
// point that had no match in its instance's last iteration first (such a point
// goes straight to the tree walk, tens of times the cost of a seeded one), the
// others after, so that the longest tiles start at once.  The search kernels
// flag each tile (fit_flags[tile]) in every iteration: the last one leaves the
// flag the final pass's tile needs (both passes take a tile's 64 points in the
// same order).  Each block takes one range of each part with one atomic on each
// of the two fill counters (ctr, ctr + kFetchStride): heavy tiles fill out from
// the front, the rest from the back.  The order within each part follows the blocks' atomics;
// it changes no result (each tile writes its own partials).
constexpr int kOrderThreads = 256;

__global__ void __launch_bounds__(kOrderThreads) k_fit_order(const IcpBuffers b, int2* __restrict__ out,
                                                             int* __restrict__ ctr) {
  constexpr int kWaves = kOrderThreads / 64;
  __shared__ int s_cnt[2][kWaves];
  __shared__ int s_base[2];
  const int nt = *b.fit_count;
  const int i = (int)(blockIdx.x * kOrderThreads + threadIdx.x);
  const bool valid = i < nt;
  bool heavy = false;
  int2 e = make_int2(0, 0);
  if (valid) {
    e = b.fit_list[i];
    heavy = b.fit_flags[e.x] != 0;
  }
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
  const uint64_t hm = __ballot(valid && heavy), lm = __ballot(valid && !heavy);
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  if (lane == 0) {
    s_cnt[0][w] = __popcll(hm);
    s_cnt[1][w] = __popcll(lm);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int h = 0, l = 0;
    for (int k = 0; k < kWaves; ++k) {
      const int ch = s_cnt[0][k], cl = s_cnt[1][k];
      s_cnt[0][k] = h;
      s_cnt[1][k] = l;
      h += ch;
      l += cl;
    }
    s_base[0] = h ? atomicAdd(ctr, h) : 0;
    s_base[1] = l ? atomicAdd(ctr + kFetchStride, l) : 0;
  }
  __syncthreads();
  if (valid) {
    if (heavy)
      out[s_base[0] + s_cnt[0][w] + __popcll(hm & below)] = e;
    else
      out[nt - 1 - (s_base[1] + s_cnt[1][w] + __popcll(lm & below))] = e;
  }
}

// The fitness pass's second half, one block per instance that took part (every
// instance the gate has not set aside): its tiles' partial sums in a fixed order,
// then the mean.  A separate launch: folding in the pass itself took a
// device-scope fence and an atomic ticket per tile (each fence writes back the
// XCD's L2), which cost cycle36 a third of the pass.
__global__ void __launch_bounds__(kFoldThreads) k_fitness_fold(const IcpBuffers b, int per_instance) {
  __shared__ double lds[2 * (kFoldThreads / 64)];
  // the pass's tile counters, ready for the next pass (the pass has ended: this
  // launch follows it on the stream)
  if (b.fetch && blockIdx.x == 0)
    for (int k = (int)threadIdx.x; k < kFetchLines; k += kFoldThreads) b.fetch[k * kFetchStride] = 0;
  PairState* st = b.state + blockIdx.x;
  if (st->gated) return;  // uniform across the block
  const PairDesc pd = b.pairs[blockIdx.x];
  const int stride = per_instance ? gate_stride(pd.ntiles) : 1;
  const int nparts = (pd.ntiles + stride - 1) / stride;
  double v[2] = {0.0, 0.0};
  for (int t = threadIdx.x; t < nparts; t += kFoldThreads) {
    const double* in = b.partials + (int64_t)(pd.tile_begin + t * stride) * kMaxVals;
    v[0] += in[0];
    v[1] += in[1];
  }
  block_sum<2>(v, lds);
  if (threadIdx.x == 0) {
    st->fit_count = (int)v[1];
    st->fitness = v[1] > 0.0 ? v[0] / v[1] : DBL_MAX;
  }
}

// Per pair: the first hypothesis (the plain seed) is always refined.  Of the
// others, the one with the most source points within its radius (ties: the
// smaller mean d2 of those, then the lower index) is refined as well when it has
// more such points than the seed; every other hypothesis is set aside.  The
// selection after the final fitness pass then picks the lower fitness of the
// (at most two) refined poses, so a hypothesis that wins the inlier count at its
// initial pose but ends in a worse basin never replaces the seed's refinement.
__device__ __forceinline__ bool gate_better(const PairState& a, const PairState& b) {
  return a.fit_count > b.fit_count || (a.fit_count == b.fit_count && a.fitness < b.fitness);
}

__global__ void k_gate(PairState* st, int npairs, int H) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  int alt = -1;
  for (int h = 1; h < H; ++h) {
    const PairState& s = st[p * H + h];
    if (!s.active) continue;
    if (alt < 0 || gate_better(s, st[p * H + alt])) alt = h;
  }
  const PairState& s0 = st[p * H];
  if (alt >= 0 && s0.active && !gate_better(st[p * H + alt], s0)) alt = -1;
  for (int h = 0; h < H; ++h) {
    PairState& s = st[p * H + h];
    s.fitness = DBL_MAX;
    s.fit_count = 0;
    if (h != 0 && h != alt) {
      s.active = 0;
      s.gated = 1;
    }
  }
}

// The (tile, instance) lists of the tiles whose instance takes part: mode 0 --
// active instances (the search), mode 1 -- instances the gate kept (the final
// fitness pass).  With the gate setting most hypotheses aside, a launch over
// every tile left each block a different number of live tiles and waited for the
// fullest; over the list the live tiles are dealt out evenly.  Two launches: one
// block places each instance's run of tiles (a block scan over runs of
// instances per thread, into off[]), then one thread per tile writes its entry.
constexpr int kCompactThreads = 1024;

__device__ __forceinline__ bool takes_part(const PairState& st, int mode) {
  return mode == 0 ? st.active != 0 : st.gated == 0;
}

__global__ void __launch_bounds__(kCompactThreads)
k_compact_offsets(const IcpBuffers b, int mode, int* __restrict__ off, int* __restrict__ count) {
  __shared__ int s_off[kCompactThreads];
  const int t = threadIdx.x;
  const int per = (b.npairs + kCompactThreads - 1) / kCompactThreads;
  const int p0 = min(t * per, b.npairs), p1 = min(p0 + per, b.npairs);
  int n = 0;
  for (int p = p0; p < p1; ++p)
    if (takes_part(b.state[p], mode)) n += b.pairs[p].ntiles;
  s_off[t] = n;
  __syncthreads();
  for (int o = 1; o < kCompactThreads; o <<= 1) {
    const int v = t >= o ? s_off[t - o] : 0;
    __syncthreads();
    s_off[t] += v;
    __syncthreads();
  }
  if (t == kCompactThreads - 1) *count = s_off[t];
  int at = s_off[t] - n;
  for (int p = p0; p < p1; ++p) {
    const bool on = takes_part(b.state[p], mode);
    off[p] = on ? at : -1;
    if (on) at += b.pairs[p].ntiles;
  }
}

__global__ void k_compact_write(const IcpBuffers b, const int* __restrict__ off,
                                int2* __restrict__ list) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= b.ntiles) return;
  const int p = b.tile_pair[t];
  const int o = off[p];
  if (o >= 0) list[o + (t - b.pairs[p].tile_begin)] = make_int2(t, p);
}

__global__ void k_state_init(PairState* st, const double* guesses, const PairDesc* pairs,
                             const GridDesc* grids, int npairs) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  PairState s;
  for (int i = 0; i < 16; ++i)
    s.T[i] = guesses ? guesses[16 * (int64_t)p + i] : ((i % 5 == 0) ? 1.0 : 0.0);
  s.prev_mse = DBL_MAX;
  s.last_mse = 0.0;
  s.fitness = DBL_MAX;
  s.iterations = 0;
  s.state = kNotConverged;
  // Registration::initCompute fails on an empty source or target: nothing runs
  const PairDesc pd = pairs[p];
  s.active = (grids[pd.src].npts > 0 && grids[pd.tgt].npts > 0) ? 1 : 0;
  s.converged = 0;
  s.similar = 0;
  s.ncorr = 0;
  s.solve_failed = 0;
  s.fit_count = 0;
  s.gated = 0;
  st[p] = s;
}

__global__ void k_correspondences(const IcpBuffers b, int pair, int32_t* idx, float* d2) {
  const PairDesc pd = b.pairs[pair];
  const GridDesc gs = b.grids[pd.src];
  const GridDesc gt = b.grids[pd.tgt];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= gs.npts) return;
  const float4 s = b.pts[gs.pt_off + i];
  float qx, qy, qz;
  xform(b.state[pair].T, s, qx, qy, qz);
  const float r = (float)sqrt(pd.max_corr_d2);
  const int o = __float_as_int(s.w);
  if (isfinite(r) && b.near && gt.near_r >= r && !near_cell(gt, b.near, qx, qy, qz)) {
    idx[o] = -1;
    d2[o] = INFINITY;
    return;
  }
  const NNResult nn =
      nn_query(gt, b.cell_start + gt.cell_off, b.pts, b.tree(), qx, qy, qz, r, isfinite(r));
  const bool keep = nn.j >= 0 && (double)nn.d2 <= pd.max_corr_d2;
  idx[o] = keep ? nn.id : -1;
  d2[o] = nn.d2;
}

// raw: the source cloud in original order (x, y, z, *) and its normals (nullable)
__global__ void k_transform_output(const float4* raw, const float4* raw_nrm, int64_t n,
                                   const PairState* st, int pair, float* out_xyz,
                                   float* out_nrm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* T = st[pair].T;
  const float4 p = raw[i];
  if (isfinite(p.x) && isfinite(p.y) && isfinite(p.z)) {
    float x, y, z;
    xform(T, p, x, y, z);
    out_xyz[3 * i] = x;
    out_xyz[3 * i + 1] = y;
    out_xyz[3 * i + 2] = z;
  } else {
    out_xyz[3 * i] = p.x;
    out_xyz[3 * i + 1] = p.y;
    out_xyz[3 * i + 2] = p.z;
  }
  if (out_nrm && raw_nrm) {
    const float4 m = raw_nrm[i];
    if (isfinite(m.x) && isfinite(m.y) && isfinite(m.z)) {
      const double nx = m.x, ny = m.y, nz = m.z;
      out_nrm[3 * i] = (float)((T[0] * nx + T[1] * ny) + T[2] * nz);
      out_nrm[3 * i + 1] = (float)((T[4] * nx + T[5] * ny) + T[6] * nz);
      out_nrm[3 * i + 2] = (float)((T[8] * nx + T[9] * ny) + T[10] * nz);
    } else {
      out_nrm[3 * i] = m.x;
      out_nrm[3 * i + 1] = m.y;
      out_nrm[3 * i + 2] = m.z;
    }
  }
}

// Same arithmetic as k_transform_output for a caller-given transform (pose
// chaining of a cycle's clouds).
struct Mat16 {
  double m[16];
};

__global__ void k_transform_points(const float4* raw, const float4* raw_nrm, int64_t n, Mat16 T,
                                   float* out_xyz, float* out_nrm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = raw[i];
  if (isfinite(p.x) && isfinite(p.y) && isfinite(p.z)) {
    float x, y, z;
    xform(T.m, p, x, y, z);
    out_xyz[3 * i] = x;
    out_xyz[3 * i + 1] = y;
    out_xyz[3 * i + 2] = z;
  } else {
    out_xyz[3 * i] = p.x;
    out_xyz[3 * i + 1] = p.y;
    out_xyz[3 * i + 2] = p.z;
  }
  if (out_nrm && raw_nrm) {
    const float4 m = raw_nrm[i];
    if (isfinite(m.x) && isfinite(m.y) && isfinite(m.z)) {
      const double nx = m.x, ny = m.y, nz = m.z;
      out_nrm[3 * i] = (float)((T.m[0] * nx + T.m[1] * ny) + T.m[2] * nz);
      out_nrm[3 * i + 1] = (float)((T.m[4] * nx + T.m[5] * ny) + T.m[6] * nz);
      out_nrm[3 * i + 2] = (float)((T.m[8] * nx + T.m[9] * ny) + T.m[10] * nz);
    } else {
      out_nrm[3 * i] = m.x;
      out_nrm[3 * i + 1] = m.y;
      out_nrm[3 * i + 2] = m.z;
    }
  }
}

template <bool P2PL, int TILE>
void launch_iter_t(const IcpBuffers& b, const IcpConst& c, int nblocks, hipStream_t s) {
  if (c.ring_only)
    hipLaunchKernelGGL((k_icp_iter<P2PL, TILE, 0, true>), dim3(nblocks), dim3(TILE), 0, s, b, c);
  else
    hipLaunchKernelGGL((k_icp_iter<P2PL, TILE, 0>), dim3(nblocks), dim3(TILE), 0, s, b, c);
}

template <bool P2PL, int TILE, bool RING>
int iter_blocks_per_cu() {
  int nb = 0;
  const hipError_t e =
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<P2PL, TILE, 0, RING>, TILE, 0);
  return (e != hipSuccess || nb < 1) ? 1 : nb;
}

template <int PREC>
void launch_iter_tiles(const IcpBuffers& b, const IcpConst& c, int nblocks, hipStream_t s) {
  if (c.point_to_plane)
    hipLaunchKernelGGL((k_icp_iter<true, 64, PREC>), dim3(nblocks), dim3(64), 0, s, b, c);
  else
    hipLaunchKernelGGL((k_icp_iter<false, 64, PREC>), dim3(nblocks), dim3(64), 0, s, b, c);
}

void launch_iter_scalar(const IcpBuffers& b, const IcpConst& c, int nblocks, hipStream_t s) {
  const bool pl = c.point_to_plane != 0;
  switch (b.tile) {
    case 64:
      pl ? launch_iter_t<true, 64>(b, c, nblocks, s) : launch_iter_t<false, 64>(b, c, nblocks, s);
      break;
    case 128:
      pl ? launch_iter_t<true, 128>(b, c, nblocks, s) : launch_iter_t<false, 128>(b, c, nblocks, s);
      break;
    default:
      pl ? launch_iter_t<true, 256>(b, c, nblocks, s) : launch_iter_t<false, 256>(b, c, nblocks, s);
      break;
  }
}

}  // namespace

void launch_icp_search(const IcpBuffers& b, const IcpConst& c, int nblocks, hipStream_t s) {
  if (c.precision == 1)
    launch_iter_tiles<1>(b, c, nblocks, s);
  else if (c.precision == 2)
    launch_iter_tiles<2>(b, c, nblocks, s);
  else if (c.precision == 3)
    launch_iter_tiles<3>(b, c, nblocks, s);
  else
    launch_iter_scalar(b, c, nblocks, s);
}

template <bool P2PL, int TILE>
void launch_split_t(const IcpBuffers& b, const IcpConst& c, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL((k_icp_iter_split<P2PL, TILE>), dim3(nblocks), dim3(64 * kSplitLanes), 0, s, b,
                     c);
}

void launch_icp_search_split(const IcpBuffers& b, const IcpConst& c, int nblocks, hipStream_t s) {
  const bool pl = c.point_to_plane != 0;
  switch (b.tile) {
    case 64:
      pl ? launch_split_t<true, 64>(b, c, nblocks, s) : launch_split_t<false, 64>(b, c, nblocks, s);
      break;
    case 128:
      pl ? launch_split_t<true, 128>(b, c, nblocks, s) : launch_split_t<false, 128>(b, c, nblocks, s);
      break;
    default:
      pl ? launch_split_t<true, 256>(b, c, nblocks, s) : launch_split_t<false, 256>(b, c, nblocks, s);
      break;
  }
}

template <bool P2PL, int TILE>
int split_blocks_per_cu_t() {
  int nb = 0;
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, k_icp_iter_split<P2PL, TILE>, 64 * kSplitLanes, 0);
  return (e != hipSuccess || nb < 1) ? 1 : nb;
}

int icp_split_blocks_per_cu(int tile, bool p2pl) {
  switch (tile) {
    case 64:
      return p2pl ? split_blocks_per_cu_t<true, 64>() : split_blocks_per_cu_t<false, 64>();
    case 128:
      return p2pl ? split_blocks_per_cu_t<true, 128>() : split_blocks_per_cu_t<false, 128>();
    default:
      return p2pl ? split_blocks_per_cu_t<true, 256>() : split_blocks_per_cu_t<false, 256>();
  }
}

void launch_icp_solve(const IcpBuffers& b, const IcpConst& c, hipStream_t s) {
  if (b.npairs <= 0) return;
  if (c.point_to_plane)
    hipLaunchKernelGGL(k_icp_solve<true>, dim3(b.npairs), dim3(kSolveThreads), 0, s, b, c);
  else
    hipLaunchKernelGGL(k_icp_solve<false>, dim3(b.npairs), dim3(kSolveThreads), 0, s, b, c);
}

void launch_fitness(const IcpBuffers& b, double max_range, int nblocks, hipStream_t s,
                    bool per_instance, int interleave) {
  // (the fold launch below also clears the tile counters, so both run or neither)
  if (b.npairs <= 0) return;
  const int pi = per_instance ? 1 : 0;
  switch (b.tile) {
    case 64:
      if (!per_instance && b.fetch && b.fit_order && b.fit_flags && b.fit_count && b.prev &&
          b.ntiles > 0) {
        // the final pass over fit_list, tree-walk tiles first
        hipLaunchKernelGGL(k_fit_order, dim3((b.ntiles + kOrderThreads - 1) / kOrderThreads),
                           dim3(kOrderThreads), 0, s, b, b.fit_order,
                           b.fetch + kFetchGroups * kFetchStride);
        IcpBuffers bo = b;
        bo.fit_list = b.fit_order;
        hipLaunchKernelGGL(k_fitness<64>, dim3(nblocks), dim3(64), 0, s, bo, max_range, pi,
                           interleave);
      } else {
        hipLaunchKernelGGL(k_fitness<64>, dim3(nblocks), dim3(64), 0, s, b, max_range, pi,
                           interleave);
      }
      break;
    case 128:
      hipLaunchKernelGGL(k_fitness<128>, dim3(nblocks), dim3(128), 0, s, b, max_range, pi,
                         interleave);
      break;
    default:
      hipLaunchKernelGGL(k_fitness<256>, dim3(nblocks), dim3(256), 0, s, b, max_range, pi,
                         interleave);
      break;
  }
  hipLaunchKernelGGL(k_fitness_fold, dim3(b.npairs), dim3(kFoldThreads), 0, s, b, pi);
}

void launch_compact_tiles(const IcpBuffers& b, int* off, int2* it_list, int* it_count,
                          int2* fit_list, int* fit_count, bool fit, hipStream_t s) {
  if (b.npairs <= 0 || b.ntiles <= 0) return;
  const unsigned blocks = (unsigned)((b.ntiles + 255) / 256);
  hipLaunchKernelGGL(k_compact_offsets, dim3(1), dim3(kCompactThreads), 0, s, b, 0, off, it_count);
  hipLaunchKernelGGL(k_compact_write, dim3(blocks), dim3(256), 0, s, b, off, it_list);
  if (!fit) return;
  hipLaunchKernelGGL(k_compact_offsets, dim3(1), dim3(kCompactThreads), 0, s, b, 1, off,
                     fit_count);
  hipLaunchKernelGGL(k_compact_write, dim3(blocks), dim3(256), 0, s, b, off, fit_list);
}

void launch_gate(PairState* st, int npairs, int H, hipStream_t s) {
  if (npairs <= 0) return;
  hipLaunchKernelGGL(k_gate, dim3((npairs + 63) / 64), dim3(64), 0, s, st, npairs, H);
}

void launch_state_init(PairState* st, const double* guesses, const PairDesc* pairs,
                       const GridDesc* grids, int npairs, hipStream_t s) {
  hipLaunchKernelGGL(k_state_init, dim3((npairs + 63) / 64), dim3(64), 0, s, st, guesses, pairs,
                     grids, npairs);
}

void launch_correspondences(const IcpBuffers& b, int pair, int npts_src, int32_t* idx, float* d2,
                            hipStream_t s) {
  if (npts_src <= 0) return;
  hipLaunchKernelGGL(k_correspondences, dim3((npts_src + 255) / 256), dim3(256), 0, s, b, pair,
                     idx, d2);
}

void launch_transform_output(const float4* raw, const float4* raw_nrm, int64_t n,
                             const PairState* st, int pair, float* out_xyz, float* out_nrm,
                             hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_transform_output, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, raw,
                     raw_nrm, n, st, pair, out_xyz, out_nrm);
}

void launch_transform_points(const float4* raw, const float4* raw_nrm, int64_t n, const double* T,
                             float* out_xyz, float* out_nrm, hipStream_t s) {
  if (n <= 0) return;
  Mat16 M;
  for (int i = 0; i < 16; ++i) M.m[i] = T[i];
  hipLaunchKernelGGL(k_transform_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, raw,
                     raw_nrm, n, M, out_xyz, out_nrm);
}

int icp_max_blocks_per_cu(int tile, bool p2pl, int precision, bool ring_only) {
  int nb = 0;
  hipError_t e = hipSuccess;
  if (ring_only && precision == 0) {
    switch (tile) {
      case 64:
        return p2pl ? iter_blocks_per_cu<true, 64, true>() : iter_blocks_per_cu<false, 64, true>();
      case 128:
        return p2pl ? iter_blocks_per_cu<true, 128, true>() : iter_blocks_per_cu<false, 128, true>();
      default:
        return p2pl ? iter_blocks_per_cu<true, 256, true>() : iter_blocks_per_cu<false, 256, true>();
    }
  }
  if (precision >= 1 && precision <= 3) {
    if (precision == 1)
      e = p2pl ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<true, 64, 1>, 64, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<false, 64, 1>, 64, 0);
    else if (precision == 2)
      e = p2pl ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<true, 64, 2>, 64, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<false, 64, 2>, 64, 0);
    else
      e = p2pl ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<true, 64, 3>, 64, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<false, 64, 3>, 64, 0);
    if (e != hipSuccess || nb < 1) nb = 1;
    return nb;
  }
  switch (tile) {
    case 64:
      e = p2pl ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<true, 64, 0>, 64, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<false, 64, 0>, 64, 0);
      break;
    case 128:
      e = p2pl ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<true, 128, 0>, 128, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<false, 128, 0>, 128, 0);
      break;
    default:
      e = p2pl ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<true, 256, 0>, 256, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_icp_iter<false, 256, 0>, 256, 0);
      break;
  }
  if (e != hipSuccess || nb < 1) nb = 1;
  return nb;
}

int fitness_max_blocks_per_cu(int tile) {
  int nb = 0;
  hipError_t e;
  switch (tile) {
    case 64:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fitness<64>, 64, 0);
      break;
    case 128:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fitness<128>, 128, 0);
      break;
    default:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fitness<256>, 256, 0);
      break;
  }
  if (e != hipSuccess || nb < 1) nb = 1;
  return nb;
}

}  // namespace lcr
