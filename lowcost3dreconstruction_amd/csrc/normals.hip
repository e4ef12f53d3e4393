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
//   1. stage every point of the 3x3x3 cells around the query's cell in LDS (one
//      lane per row of cells finds the runs, then a flat coalesced copy); the
//      next query in the same cell reuses the staged block.  A block larger than
//      LDS is staged as an evenly strided subsample, whose K-th smallest d2 bounds
//      the block's from above, and then restaged keeping only the points under
//      that bound;
//   2. find the exact K-th smallest FLANN-form d2: the K-th smallest of the 64
//      per-lane minima bounds it; the points under that bound are compacted to
//      one per lane and a bisection over the float bit pattern (counts by wave
//      ballot) finds it there; equal distances are resolved by original index;
//   3. certify: the K-th distance must be below the distance to the staged
//      block's boundary, else stage the next larger block, keeping only the points
//      within the K-th distance found so far, and repeat;
//   4. list the selected points' original indices (ballot positions, fixed order).
// A second pass, one point per lane, sums the FP64 moments of its listed
// neighbours relative to it and solves the FP64 3x3 eigenproblem (keeping the
// nine cross-lane FP64 sums and the solver's registers out of the search).  Queries the staged
// search gives up on (very many equal distances, K above half the LDS capacity)
// are listed and finished by a rescan pass that reads the block from memory for
// every count.
#include <hip/hip_runtime.h>

#include "lcr_common.hpp"
#include "lcr_kernels.hpp"
#include "lcr_math.hpp"
#include "lcr_search.hpp"

namespace lcr {

#pragma clang fp contract(off)

namespace {

constexpr int kWaves = 4;
// staged points per wave: 4 x (576 x 16 + 64 x 4) B = 37 KiB per block, nine slots
// of 64.  576 keeps the dense regions of Kinect-v2 captures (1.5 mm voxels, 74
// clouds of 25k-88k points) mostly inside LDS -- 20.0 against 21.5 ms per cycle
// step at 496 -- at no cost on Kinect-v1 captures (2 mm voxels, 36 clouds of
// 37k-113k points); 600 and 624 (ten slots) cost those 1 %.
constexpr int kCap = 576;
constexpr int kSlots = (kCap + 63) / 64;

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

// Number of set bits of b in the lanes below this one.
__device__ __forceinline__ int lane_prefix(uint64_t b) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
}

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

// What a selection accumulates: kMoments -- FP64 moments of the neighbours
// relative to the query (normals); kSqrtDist -- the sum of sqrt(d2) over the
// neighbours in FP64 (StatisticalOutlierRemoval's mean distance).
constexpr int kMoments = 0;
constexpr int kSqrtDist = 1;

template <int MODE>
__device__ __forceinline__ void take(Moments& m, const float4& q, const float4& p, float d2) {
  if constexpr (MODE == kMoments)
    add_moment(m, q, p);
  else
    m.v[0] += sqrt((double)d2);
}

// One selected point per lane where sel: kMoments writes its original index to
// nbr[base + rank among the wave's selected lanes] (the moments are summed from
// this list by the solve pass); kSqrtDist adds sqrt(d2).  Every lane of the
// wavefront calls it; returns how many lanes selected.
template <int MODE>
__device__ __forceinline__ int select_out(bool sel, const float4& p, unsigned d, const float4& q,
                                          Moments& acc, int* nbr, int base) {
  if constexpr (MODE == kMoments) {
    const uint64_t b = __ballot(sel);
    if (sel) nbr[base + lane_prefix(b)] = __float_as_int(p.w);
    return __popcll(b);
  } else {
    if (sel) take<MODE>(acc, q, p, __uint_as_float(d));
    return 0;
  }
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
template <int MODE>
__device__ void select_rescan(const GridDesc& g, const int* cs, const float4* pts, const float4& q,
                              const CellPos& cp, int K, int kk, Moments& acc, int* nbr) {
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
  // the selected points, in block order (wave-uniform loop: ballot positions)
  const int lane = threadIdx.x & 63;
  int base = 0;
  for (int ring = 0; ring <= K; ++ring) {
    for_ring_runs(g, cp, ring, [&](int c0, int c1) {
      const int s = cs[c0], e = cs[c1 + 1];
      for (int jb = s; jb < e; jb += 64) {
        const int j = jb + lane;
        bool sel = false;
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        unsigned d = 0u;
        if (j < e) {
          p = pts[j];
          d = __float_as_uint(flann_d2(q.x, q.y, q.z, p));
          sel = d < T || (d == T && __float_as_int(p.w) <= idcut);
        }
        base += select_out<MODE>(sel, p, d, q, acc, nbr, base);
      }
    });
  }
}

// The wave's staged candidate block in LDS.  When the (2K+1)^3 cells around the
// query's cell hold at most kCap points it holds all of them, and the following
// cell-sorted queries in the same cell reuse it.  A larger block is held filtered
// -- the points whose FLANN d2 to the query is <= filt (the first kCap of them
// when more qualify) -- or, without a bound to filter by, as an evenly strided
// subsample; both serve that query only.
struct Staged {
  int64_t cell;   // cell whose complete unfiltered block is staged, else -1
  int K;          // block radius in cells
  int cnt;        // points in LDS
  int total;      // points that qualified (more than kCap: LDS holds a subset)
  unsigned filt;  // d2 filter as a bit pattern; 0xffffffff: none
  unsigned lastT; // the previous query's k-th smallest d2 (bit pattern), 0: none
};

constexpr unsigned kNoFilter = 0xffffffffu;
constexpr unsigned kInfBits = 0x7f800000u;

// Sorted-index start and length of row r (z-major, then y) of block K.
__device__ __forceinline__ void block_row(const GridDesc& g, const int* cs, int x0, int x1, int y0,
                                          int z0, int nyb, int r, int& s, int& len) {
  const int z = z0 + r / nyb, y = y0 + r % nyb;
  const int row = (z * g.ny + y) * g.nx;
  s = cs[row + x0];
  len = cs[row + x1 + 1] - s;
}

// Stages block K around cp into sp (see Staged): the whole block when it fits
// LDS, else the points with d2 <= U when a bound U is given, else an evenly
// strided subsample.  Each lane looks up one x-contiguous row of cells, a wave
// prefix sum places the rows, then the points are loaded four per lane at a time;
// the filtered form keeps the qualifying ones in block order through ballots.
__device__ __forceinline__ void stage_block(const GridDesc& g, const int* cs,
                                            const float4* __restrict__ pts, const CellPos& cp, int K,
                                            const float4& q, unsigned U, float4* sp, Staged& st) {
  const int lane = threadIdx.x & 63;
  wave_lds_sync();  // the previous block's reads come before these writes
  st.K = K;
  st.cnt = 0;
  st.total = 0;
  st.filt = kNoFilter;
  const int z0 = max(cp.cz - K, 0), z1 = min(cp.cz + K, g.nz - 1);
  const int y0 = max(cp.cy - K, 0), y1 = min(cp.cy + K, g.ny - 1);
  const int x0 = max(cp.cx - K, 0), x1 = min(cp.cx + K, g.nx - 1);
  if (z0 > z1 || y0 > y1 || x0 > x1) return;
  const int nyb = y1 - y0 + 1, nrows = (z1 - z0 + 1) * nyb;
  int tot = 0;
  for (int rb = 0; rb < nrows; rb += 64) {
    int s = 0, len = 0;
    if (rb + lane < nrows) block_row(g, cs, x0, x1, y0, z0, nyb, rb + lane, s, len);
    tot += wave_sum_i(len);
  }
  const bool filter = tot > kCap && U != kNoFilter;
  // otherwise block positions i * stride are kept
  const int stride = tot > kCap ? (tot + kCap - 1) / kCap : 1;
  int pos = 0;  // filter: points kept so far; else block position of the group's first point
  for (int rb = 0; rb < nrows; rb += 64) {
    int s = 0, len = 0;
    if (rb + lane < nrows) block_row(g, cs, x0, x1, y0, z0, nyb, rb + lane, s, len);
    int inc = len;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(inc, off, 64);
      if (lane >= off) inc += t;
    }
    const int total = __shfl(inc, 63, 64);
    const int excl = inc - len;
    const int nr = min(64, nrows - rb);
    // filter: group positions; else kept indices i with block position i * stride
    // inside this group
    const int i0 = filter ? 0 : (pos + stride - 1) / stride;
    const int i1 = filter ? total : (pos + total + stride - 1) / stride;
    for (int ib = i0; ib < i1; ib += 256) {
      // positions past the end are clamped to the last one (loaded, not stored)
      int tt[4], jb[4] = {0, 0, 0, 0};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = min(ib + lane + 64 * u, i1 - 1);
        tt[u] = filter ? i : i * stride - pos;
      }
      // row of position t: the last row whose first position is <= t
      for (int rr = 0; rr < nr; ++rr) {
        const int ex = __builtin_amdgcn_readlane(excl, rr);
        const int sr = __builtin_amdgcn_readlane(s, rr);
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (ex <= tt[u]) jb[u] = sr - ex;
      }
      float4 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = pts[jb[u] + tt[u]];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i = ib + lane + 64 * u;
        if (filter) {
          const bool keep = i < i1 && __float_as_uint(flann_d2(q.x, q.y, q.z, v[u])) <= U;
          const uint64_t b = __ballot(keep);
          const int at = pos + lane_prefix(b);
          if (keep && at < kCap) sp[at] = v[u];
          pos += __popcll(b);
        } else if (i < i1) {
          sp[i] = v[u];
        }
      }
    }
    if (!filter) pos += total;
  }
  if (filter) {
    st.filt = U;
    st.total = pos;
    st.cnt = min(pos, kCap);
  } else {
    st.total = tot;
    st.cnt = (tot + stride - 1) / stride;
  }
  wave_lds_sync();
}

// A complete unfiltered block can serve the following queries of its cell.
__device__ __forceinline__ bool reusable(const Staged& st) {
  return st.filt == kNoFilter && st.total <= kCap;
}

// FLANN d2 bit patterns of the staged points, point lane + 64 m in slot m
// (0xffffffff past the end).
__device__ __forceinline__ void fill_slots(const float4* sp, int cnt, const float4& q,
                                           unsigned (&slot)[kSlots]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int m = 0; m < kSlots; ++m) {
    const int i = lane + 64 * m;
    slot[m] = i < cnt ? __float_as_uint(flann_d2(q.x, q.y, q.z, sp[i])) : 0xffffffffu;
  }
}

// Smallest t <= hi0 with at least kk values <= t: the kk-th smallest when at least
// kk of the values are <= hi0.  Bisection over the bit pattern, counts by ballot.
template <int NS>
__device__ __forceinline__ unsigned kth_slots(const unsigned (&d)[NS], int ns, int kk, unsigned hi0) {
  unsigned lo = 0u, hi = hi0;
  while (lo < hi) {
    const unsigned mid = lo + ((hi - lo) >> 1);
    int c = 0;
#pragma unroll
    for (int m = 0; m < NS; ++m)
      if (m < ns) c += popc_ballot(d[m] <= mid);
    if (c >= kk)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// Maximum over the 64 lanes (every lane active): DPP inside each row of 16 lanes,
// then the four row maxima.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
  v = max(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true));   // quad_perm [1,0,3,2]
  v = max(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true));   // quad_perm [2,3,0,1]
  v = max(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x124, 0xF, 0xF, true));  // row_ror:4
  v = max(v, (unsigned)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, true));  // row_ror:8
  return max(max((unsigned)__builtin_amdgcn_readlane((int)v, 0), (unsigned)__builtin_amdgcn_readlane((int)v, 16)),
             max((unsigned)__builtin_amdgcn_readlane((int)v, 32), (unsigned)__builtin_amdgcn_readlane((int)v, 48)));
}

// The kk-th smallest of one value per lane (d: a bit pattern <= 0x7f800000, or
// 0xffffffff for none), given c = count(d <= hi) >= kk.  Bisection over the bit
// pattern keeping count(d <= lo - 1) < kk <= count(d <= hi); it stops as soon as
// exactly kk values lie at or under hi -- the kk-th smallest is then the largest
// of them -- which takes about log2 of the values' spread instead of all 31 bits.
__device__ __forceinline__ unsigned kth_one(unsigned d, int kk, unsigned hi, int c) {
  unsigned lo = 0u;
  while (c > kk && lo < hi) {
    const unsigned mid = lo + ((hi - lo) >> 1);
    const int m = popc_ballot(d <= mid);
    if (m >= kk) {
      hi = mid;
      c = m;
    } else {
      lo = mid + 1;
    }
  }
  if (c == kk) return wave_max_u32(d <= hi ? d : 0u);
  return lo;  // lo == hi: equal values at the kk-th place
}

// The kk-th smallest d2 of the cnt staged points (cnt >= kk).  With more than 64
// points and kk <= 64 the points are filtered by a bound U with at least kk and
// at most 64 staged points under it -- then the kk-th smallest lies under U --
// and compacted to one per lane (cd, ci: d2 bits and staged index, through cb);
// the bisection runs over that single slot.  U is the guess (a nearby query's
// k-th d2, grown by 1.44) when its count qualifies, possibly after one doubling
// or halving; else the kk-th smallest of the 64 per-lane minima, which are kk
// distinct staged points.  comp tells whether the compacted slot was used.
__device__ __forceinline__ unsigned kth_staged(const unsigned (&slot)[kSlots], int cnt, int kk,
                                               unsigned guess, const float4* sp, const float4& q,
                                               int* cb, unsigned& cd, int& ci, bool& comp) {
  const int lane = threadIdx.x & 63;
  const int nslots = (cnt + 63) >> 6;
  comp = false;
  if (nslots > 1 && kk <= 64) {
    unsigned U = kInfBits;
    int c = 65;  // > 64: no qualifying bound yet
    if (guess != 0u) {
      float gf = __uint_as_float(guess) * 1.44f;
      for (int t = 0; t < 2; ++t) {
        const unsigned gb = __float_as_uint(gf);
        int n = 0;
#pragma unroll
        for (int m = 0; m < kSlots; ++m)
          if (m < nslots) n += popc_ballot(slot[m] <= gb);
        if (n >= kk && n <= 64) {
          U = gb;
          c = n;
          break;
        }
        gf = n < kk ? gf * 2.0f : gf * 0.5f;
      }
    }
    if (c > 64) {
      unsigned mn = slot[0];
#pragma unroll
      for (int m = 1; m < kSlots; ++m) mn = min(mn, slot[m]);
      U = kth_one(mn, kk, kInfBits, 64);  // more than 64 staged points: every lane has one
      c = 0;
#pragma unroll
      for (int m = 0; m < kSlots; ++m)
        if (m < nslots) c += popc_ballot(slot[m] <= U);
    }
    if (c <= 64) {
      wave_lds_sync();
      int base = 0;
#pragma unroll
      for (int m = 0; m < kSlots; ++m)
        if (m < nslots) {
          const bool keep = slot[m] <= U;
          const uint64_t b = __ballot(keep);
          if (keep) cb[base + lane_prefix(b)] = lane + 64 * m;
          base += __popcll(b);
        }
      wave_lds_sync();
      ci = lane < c ? cb[lane] : 0;
      cd = lane < c ? __float_as_uint(flann_d2(q.x, q.y, q.z, sp[ci])) : 0xffffffffu;
      comp = true;
      return kth_one(cd, kk, U, c);
    }
  }
  return kth_slots<kSlots>(slot, nslots, kk, kInfBits);
}

// Accumulates the kk points with the smallest (d2, original index) among
// candidates (d bits, staged index ix) that hold every point with d2 <= T, T the
// kk-th smallest d2: all points below T and the smallest indices at T.
template <int MODE, int NS>
__device__ __forceinline__ void take_kth(const unsigned (&d)[NS], const int (&ix)[NS], int ns, int kk,
                                         unsigned T, const float4* sp, const float4& q, Moments& acc,
                                         int* nbr) {
  int clt = 0, cle = 0;
#pragma unroll
  for (int m = 0; m < NS; ++m)
    if (m < ns) {
      clt += popc_ballot(d[m] < T);
      cle += popc_ballot(d[m] <= T);
    }
  int idcut = 0x7fffffff;
  if (cle > kk) {
    const int need = kk - clt;
    int ids[NS];
#pragma unroll
    for (int m = 0; m < NS; ++m) ids[m] = d[m] == T ? __float_as_int(sp[ix[m]].w) : 0x7fffffff;
    unsigned lo = 0u, hi = 0x7fffffffu;
    while (lo < hi) {
      const unsigned mid = lo + ((hi - lo) >> 1);
      int c = 0;
#pragma unroll
      for (int m = 0; m < NS; ++m)
        if (m < ns) c += popc_ballot(d[m] == T && (unsigned)ids[m] <= mid);
      if (c >= need)
        hi = mid;
      else
        lo = mid + 1;
    }
    idcut = (int)lo;
  }
  int base = 0;
#pragma unroll
  for (int m = 0; m < NS; ++m)
    if (m < ns) {
      bool sel = false;
      float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
      if (d[m] <= T) {
        p = sp[ix[m]];
        sel = d[m] < T || __float_as_int(p.w) <= idcut;
      }
      base += select_out<MODE>(sel, p, d[m], q, acc, nbr, base);
    }
}

// Tightens an overflowing block's bound at most this many times before leaving
// the query to the rescan pass.
constexpr int kMaxTighten = 8;

// Exact K-nearest selection for query q from LDS-staged blocks.  Block K grows
// until the kk-th smallest d2 is certified below the distance to the block's
// boundary.  A block larger than LDS is staged filtered by a proven upper bound
// of its kk-th smallest d2 (the previous block's kk-th smallest, or that of a
// staged subset of the block), so it fits LDS unless very many points tie.
// Returns false (acc untouched) when it gives up; the query is then left to the
// rescan pass.
template <int MODE>
__device__ __forceinline__ bool select_staged(const GridDesc& g, const int* cs, const float4* pts,
                                              const float4& q, int kk, float4* sp, int* cb, Staged& st,
                                              Moments& acc, int* nbr) {
  const int lane = threadIdx.x & 63;
  const CellPos cp = cell_pos(g, q.x, q.y, q.z);
  const int64_t cell = (int64_t)g.cell_off + ((int64_t)cp.cz * g.ny + cp.cy) * g.nx + cp.cx;
  if (st.cell != cell) {
    stage_block(g, cs, pts, cp, 1, q, kNoFilter, sp, st);
    st.cell = reusable(st) ? cell : -1;
  }
  int K = st.K;
  unsigned slot[kSlots];
  unsigned T = 0, cd = 0xffffffffu;
  int ci = 0;
  bool comp = false;
  // Every pass grows K (until the block covers the grid) or tightens a bound, at
  // most kMaxTighten times: the loop ends.
  for (int tighten = 0;;) {
    if (st.total > kCap) {
      // LDS holds a subset of the qualifying points of block K: the kk-th smallest
      // d2 among them bounds the block's kk-th smallest from above
      if (st.cnt < kk || ++tighten > kMaxTighten) return false;
      fill_slots(sp, st.cnt, q, slot);
      const unsigned U = kth_slots<kSlots>(slot, (st.cnt + 63) >> 6, kk, kInfBits);
      if (U >= st.filt) return false;  // no progress: very many equal distances
      stage_block(g, cs, pts, cp, K, q, U, sp, st);
      st.cell = -1;
      continue;
    }
    const float Lc = block_gap(g, cp, K);
    const bool covers = (Lc == INFINITY);
    if (st.cnt >= kk) {
      fill_slots(sp, st.cnt, q, slot);
      T = kth_staged(slot, st.cnt, kk, st.lastT, sp, q, cb, cd, ci, comp);
      const float L = (Lc - cp.slack) * g.h;
      if (covers || (L > 0.0f && __uint_as_float(T) < L * L * 0.99998f)) break;
      // the next block's kk-th smallest d2 is at most T
      ++K;
      stage_block(g, cs, pts, cp, K, q, T, sp, st);
      st.cell = reusable(st) ? cell : -1;
      continue;
    }
    // fewer than kk points: only an unfiltered block that does not cover the grid;
    // an isolated point's block grows by half its radius at a time
    if (st.filt != kNoFilter || covers) return false;
    K += max(1, K >> 1);
    stage_block(g, cs, pts, cp, K, q, kNoFilter, sp, st);
    st.cell = reusable(st) ? cell : -1;
  }
  st.lastT = T;
  if (comp) {
    const unsigned d1[1] = {cd};
    const int i1[1] = {ci};
    take_kth<MODE, 1>(d1, i1, 1, kk, T, sp, q, acc, nbr);
  } else {
    int ix[kSlots];
#pragma unroll
    for (int m = 0; m < kSlots; ++m) ix[m] = lane + 64 * m;
    take_kth<MODE, kSlots>(slot, ix, (st.cnt + 63) >> 6, kk, T, sp, q, acc, nbr);
  }
  return true;
}

// ---------------------------------------------------------------- lane path
//
// One query per lane (64 neighbouring cell-sorted queries per wavefront, each
// lane on its own): the candidates are the points of the (2K+1)^3 cells around
// the query's cell, read from memory (the wavefront's queries share most of
// them, so the reads hit the caches).  The kk-th smallest FLANN d2 is certified
// when it lies below D2, the squared distance from the query to the block's
// boundary (minus the rounding slack), so only d2 < D2 counts, and each row of
// the block is trimmed to the cells that can hold such a point (ball_row).  Pass 1 builds a
// per-lane histogram of d2 over [0, D2) in NB equal buckets -- on a surface the
// number of points within a radius grows with its square, so the buckets fill
// about evenly -- and finds the bucket B that holds the kk-th.  Pass 2 lists
// every point of a lower bucket and keeps (d2, original index, position) of
// bucket B's points in LDS; a sorting network orders those by (d2, original
// index) and the first kk - (count below B) complete the list.  Bucket index =
// floor(d2 * scale) is monotone in d2, so the result is exactly the kk smallest
// (d2, original index).  A query is flagged for the next path when its block
// covers the grid, holds fewer than kk points under D2 (not certified), more
// than kLaneMaxCand candidates, or more than NL points in bucket B.
//
// The lane pass runs K = 1 (3x3x3 cells) over every point and appends the
// points it flags (a few percent: near crop edges and silhouettes the 3x3x3
// block holds too few; dense spots overflow bucket B) to a list, which the
// wavefront pass below serves.
constexpr int kLaneMaxCand = 4096;
// The lane pass's histogram buckets (the wavefront pass reads its bucket bounds
// in units of D2 / kLaneBuckets).
constexpr int kLaneBuckets = 32;

constexpr int kLaneThreads = 256;

// Sorted-array start and length of the (2K+1)^2 x-runs (at most 2K+1 cells each)
// of the block of radius K around cp: the cell-table loads are issued together.
template <int K>
struct BlockRuns {
  static constexpr int R = (2 * K + 1) * (2 * K + 1);
  int s[R], n[R];
};

// The cells of row (z, y) of the block of radius K that may hold a point with d2
// below D2: [xa, xb] (empty when xa > xb).  Each cell's lower-bound distance per
// axis is axis_gap (cell units, less the rounding slack), and the bound is
// widened by 2e-5 for the float rounding of d2, as the ring search prunes.
__device__ __forceinline__ void ball_row(const GridDesc& g, const CellPos& cp, int K, int z, int y,
                                         float D2, int& xa, int& xb) {
  const float gz = axis_gap(cp.fz, z, cp.cz, cp.slack), gy = axis_gap(cp.fy, y, cp.cy, cp.slack);
  const float thr = D2 * 1.00002f;
  const float lyz = (gz * gz + gy * gy) * (g.h * g.h);
  const float dxc = sqrtf(fmaxf(thr - lyz, 0.0f)) * g.inv_h * 1.00001f + cp.slack;
  xa = max(max(cp.cx - K, (int)ceilf(cp.fx - dxc - 1.0f)), 0);
  xb = min(min(cp.cx + K, (int)floorf(cp.fx + dxc)), g.nx - 1);
  if (z < 0 || z >= g.nz || y < 0 || y >= g.ny || lyz > thr) xb = xa - 1;
}

// The runs of the block of radius K trimmed to the cells that may hold a point
// with d2 below D2 (ball_row): only those can count towards the certificate.
template <int K>
__device__ __forceinline__ BlockRuns<K> block_runs(const GridDesc& g, const int* cs, const CellPos& cp,
                                                   float D2) {
  BlockRuns<K> b;
#pragma unroll
  for (int r = 0; r < BlockRuns<K>::R; ++r) {
    const int z = cp.cz + r / (2 * K + 1) - K, y = cp.cy + r % (2 * K + 1) - K;
    int xa, xb;
    ball_row(g, cp, K, z, y, D2, xa, xb);
    const bool in = xa <= xb;
    const int row = in ? (z * g.ny + y) * g.nx : 0;
    const int s = in ? cs[row + xa] : 0;
    const int e = in ? cs[row + xb + 1] : 0;
    b.s[r] = s;
    b.n[r] = e - s;
  }
  return b;
}

// The cell-sorted points as a raw buffer (gfx9 resource word 3): a load is one
// 32-bit offset add, and a load past the end returns zeros instead of faulting.
// The caller keeps n * 16 below 2^31 bytes.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t points_rsrc(const float4* pts, int64_t n) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)pts, (short)0, (int)(n * 16), 0x00020000);
}

__device__ __forceinline__ float4 load_point(__amdgpu_buffer_rsrc_t r, int i) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, i * 16, 0, 0));
}

// Calls f(p, i, valid) for the points i of the block's runs, NBATCH loads in
// flight; the loads of a run's last batch read on past its end (the next points,
// or zeros past the buffer) with valid = false.  The runs go column by column:
// the 2K+1 runs of one y offset (z = -K .. K) back to back, each lane moving to
// its next non-empty run as soon as it finishes one.  A loop over one run at a
// time lasts as long as the wavefront's longest lane in every run; per column,
// the lanes' long and short runs even out (about 76 against 101 batches per
// wavefront on the 36-capture cycle's queries).
template <int K, int NBATCH, class F>
__device__ __forceinline__ void for_block(const BlockRuns<K>& b, __amdgpu_buffer_rsrc_t rs,
                                          F&& f) {
  constexpr int S = 2 * K + 1;
#pragma unroll
  for (int dy = 0; dy < S; ++dy) {
    // the column's runs with the empty ones moved to the end, order kept
    int qs[S], qe[S];
#pragma unroll
    for (int t = 0; t < S; ++t) {
      qs[t] = b.s[t * S + dy];
      qe[t] = qs[t] + b.n[t * S + dy];
    }
#pragma unroll
    for (int pass = 0; pass + 1 < S; ++pass)
#pragma unroll
      for (int t = 0; t + 1 < S; ++t) {
        const bool gap = qs[t] >= qe[t];
        qs[t] = gap ? qs[t + 1] : qs[t];
        qe[t] = gap ? qe[t + 1] : qe[t];
        qs[t + 1] = gap ? 0 : qs[t + 1];
        qe[t + 1] = gap ? 0 : qe[t + 1];
      }
    int i = qs[0], e = qe[0];
    while (i < e) {
      float4 p[NBATCH];
#pragma unroll
      for (int u = 0; u < NBATCH; ++u) p[u] = load_point(rs, i + u);
#pragma unroll
      for (int u = 0; u < NBATCH; ++u) f(p[u], i + u, i + u < e);
      i += NBATCH;
      if (i >= e) {
        i = qs[1];
        e = qe[1];
#pragma unroll
        for (int t = 1; t + 1 < S; ++t) {
          qs[t] = qs[t + 1];
          qe[t] = qe[t + 1];
        }
        qs[S - 1] = 0;
        qe[S - 1] = 0;
      }
    }
  }
}

// (key, payload) compare-exchange: smaller key first.
__device__ __forceinline__ void cswap(uint64_t& a, uint64_t& b, int& pa, int& pb) {
  const bool sw = b < a;
  const uint64_t lo = sw ? b : a, hi = sw ? a : b;
  const int plo = sw ? pb : pa, phi = sw ? pa : pb;
  a = lo;
  b = hi;
  pa = plo;
  pb = phi;
}

// Sorts NL (key, payload) pairs: the 19-comparator network for 8, odd-even
// transposition (NL rounds) otherwise.
template <int NL>
__device__ __forceinline__ void sort_keys(uint64_t (&e)[NL], int (&ei)[NL]) {
  if constexpr (NL == 8) {
    cswap(e[0], e[2], ei[0], ei[2]); cswap(e[1], e[3], ei[1], ei[3]);
    cswap(e[4], e[6], ei[4], ei[6]); cswap(e[5], e[7], ei[5], ei[7]);
    cswap(e[0], e[4], ei[0], ei[4]); cswap(e[1], e[5], ei[1], ei[5]);
    cswap(e[2], e[6], ei[2], ei[6]); cswap(e[3], e[7], ei[3], ei[7]);
    cswap(e[0], e[1], ei[0], ei[1]); cswap(e[2], e[3], ei[2], ei[3]);
    cswap(e[4], e[5], ei[4], ei[5]); cswap(e[6], e[7], ei[6], ei[7]);
    cswap(e[2], e[4], ei[2], ei[4]); cswap(e[3], e[5], ei[3], ei[5]);
    cswap(e[1], e[4], ei[1], ei[4]); cswap(e[3], e[6], ei[3], ei[6]);
    cswap(e[1], e[2], ei[1], ei[2]); cswap(e[3], e[4], ei[3], ei[4]); cswap(e[5], e[6], ei[5], ei[6]);
  } else {
#pragma unroll
    for (int round = 0; round < NL; ++round)
#pragma unroll
      for (int t = round & 1; t + 1 < NL; t += 2) cswap(e[t], e[t + 1], ei[t], ei[t + 1]);
  }
}

// Block b of n runs on XCD b % 8 (the hardware's round-robin); this returns the
// logical block such that each XCD takes one contiguous range of them, so that
// neighbouring cell-sorted queries -- and the cells they read -- share an L2.
__device__ __forceinline__ int xcd_block(int b, int n) {
  const int x = b & 7, slot = b >> 3, q = n >> 3, rem = n & 7;
  return x * q + min(x, rem) + slot;
}

template <int NB, int NL>
constexpr int lane_words() {
  return 3 * NL > NB ? 3 * NL : NB;
}

// nsel bit: the query's list holds original indices into the cloud's raw points
// (the rescan path), not positions in the cell-sorted copy.
constexpr int kNselOrig = 1 << 30;

// One query of the lane path. K: block radius in cells; NB: histogram buckets;
// NL: bucket-B capacity; buf: the lane's LDS column (lane_words words).  Writes
// the FP64 moments of the selected neighbours relative to the query (mom, nine
// arrays of n) and nsel, and returns 0; or leaves the query to the next path
// (nothing written) and returns the block radius that path starts from: K + 1
// when this block held too few points or no certificate, K otherwise; when
// bucket B overflowed, B + 1 in bits 8 and up (at least kk points lie below
// (B + 1) D2 / NB, up to the rounding of d2 * scale).
template <int K, int NB, int NL>
__device__ __forceinline__ int lane_query(const GridDesc* __restrict__ grids,
                                          const int* __restrict__ cell_start,
                                          const float4* __restrict__ pts,
                                          __amdgpu_buffer_rsrc_t rs,
                                          const int* __restrict__ cloud_of_sorted, int k, int64_t j,
                                          uint32_t (*buf)[64], int lane, double* __restrict__ mom,
                                          int64_t n, int* __restrict__ nsel) {
  const float4 q = pts[j];
  const GridDesc g = grids[cloud_of_sorted[j]];
  const int kk = min(k, g.npts);
  if (kk < 3) {
    nsel[j] = kk;
    return 0;
  }
  const int* cs = cell_start + g.cell_off;
  const CellPos cp = cell_pos(g, q.x, q.y, q.z);
  const float Lc = block_gap(g, cp, K);
  const float L = (Lc - cp.slack) * g.h;
  const float D2 = L * L * 0.99998f;
  const float scale = (float)NB / D2;
  // blocks that cover the grid (no boundary to certify against) and gaps inside
  // the rounding slack go on; a finite scale keeps every bucket index of a
  // d2 < D2 inside [0, NB]
  if (!(Lc < INFINITY && L > 0.0f && D2 > 1e-30f && scale < INFINITY)) return K;
  const BlockRuns<K> blk = block_runs<K>(g, cs, cp, D2);
  int m = 0;
#pragma unroll
  for (int r = 0; r < BlockRuns<K>::R; ++r) m += blk.n[r];
  if (m < kk) return K + 1;
  if (m > kLaneMaxCand) return K;
#pragma unroll
  for (int c = 0; c < NB; ++c) buf[c][lane] = 0u;
  // pass 1, branch-free: every candidate adds 0 or 1 to its bucket
  for_block<K, 4>(blk, rs, [&](const float4& p, int, bool valid) {
    const float d = flann_d2(q.x, q.y, q.z, p);
    const int b = (int)fminf(d * scale, (float)(NB - 1));
    atomicAdd(&buf[b][lane], (valid && d < D2) ? 1u : 0u);
  });
  int B = -1, below = 0, nB = 0, cum = 0;
  // four counts at a time: unrolled whole, the 32 reads would all be in flight
#pragma unroll 4
  for (int c = 0; c < NB; ++c) {
    const int v = (int)buf[c][lane];
    if (B < 0 && cum + v >= kk) {
      B = c;
      below = cum;
      nB = v;
    }
    cum += v;
  }
  if (B < 0) return K + 1;  // not certified
  if (nB > NL) return K | ((B + 1) << 8);
  // Pass 2 takes points of buckets up to B only: floor(d2 * scale) <= B < NB - 1
  // means d2 < (B + 1) D2 / NB, to within the rounding of d2 * scale (1e-4 covers
  // it).  The rows trimmed to that bound (ball_row, monotone in the bound) keep
  // every such point, and their runs are sub-runs of pass 1's, so the points
  // pass 2 takes and the order it adds them in are unchanged.
  const BlockRuns<K> sel =
      B < NB - 1 ? block_runs<K>(g, cs, cp, fminf(D2, D2 * ((float)(B + 1) / (float)NB) * 1.0001f))
                 : blk;
  // the moments of the lower buckets' points in block order, then of bucket B's
  // picks in (d2, original index) order
  Moments acc;
#pragma unroll
  for (int c = 0; c < 9; ++c) acc.v[c] = 0.0;
  int nl = 0;
  for_block<K, 4>(sel, rs, [&](const float4& p, int i, bool valid) {
    const float d = flann_d2(q.x, q.y, q.z, p);
    if (valid && d < D2) {
      const int b = min((int)(d * scale), NB - 1);
      if (b < B) {
        add_moment(acc, q, p);
      } else if (b == B && nl < NL) {
        buf[3 * nl][lane] = __float_as_uint(d);
        buf[3 * nl + 1][lane] = (uint32_t)__float_as_int(p.w);
        buf[3 * nl + 2][lane] = (uint32_t)i;
        ++nl;
      }
    }
  });
  uint64_t e[NL];
  int ei[NL];
#pragma unroll
  for (int t = 0; t < NL; ++t) {
    e[t] = t < nl ? ((uint64_t)buf[3 * t][lane] << 32) | buf[3 * t + 1][lane] : ~0ull;
    ei[t] = t < nl ? (int)buf[3 * t + 2][lane] : 0;
  }
  sort_keys<NL>(e, ei);
  const int need = kk - below;
#pragma unroll
  for (int t = 0; t < NL; ++t)
    if (t < need) add_moment(acc, q, load_point(rs, ei[t]));
#pragma unroll
  for (int c = 0; c < 9; ++c) mom[c * n + j] = acc.v[c];
  nsel[j] = kk;
  return 0;
}

// Appends the lanes where take holds to list (list[*count ...]: the point and
// the block radius to start from), one atomic per wavefront; every lane of the
// wavefront calls it.  The order of the list depends on scheduling; every
// consumer's result per query does not.
__device__ __forceinline__ void wave_append(bool take, int64_t j, int start_k,
                                            int2* __restrict__ list, int* __restrict__ count) {
  const uint64_t b = __ballot(take);
  if (b == 0) return;
  const int lead = __ffsll((unsigned long long)b) - 1;
  int base = 0;
  if ((int)(threadIdx.x & 63) == lead) base = atomicAdd(count, __popcll(b));
  base = __builtin_amdgcn_readlane(base, lead);
  if (take) list[base + lane_prefix(b)] = make_int2((int)j, start_k);
}

// The lane pass over every point: one lane per point, blocks mapped
// XCD-contiguous; the points it leaves are appended to list.  No barrier: each
// lane uses only its own LDS column.
template <int K, int NB, int NL>
__global__ void __launch_bounds__(kLaneThreads)
k_normal_lane(const GridDesc* __restrict__ grids, const int* __restrict__ cell_start,
              const float4* __restrict__ pts, const int* __restrict__ cloud_of_sorted,
              int64_t total, int k, double* __restrict__ mom, int* __restrict__ nsel,
              int2* __restrict__ list, int* __restrict__ count) {
  static_assert(NB == kLaneBuckets, "the wavefront pass reads bucket bounds of kLaneBuckets");
  __shared__ uint32_t s_buf[kLaneThreads / 64][lane_words<NB, NL>()][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t j = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kLaneThreads + threadIdx.x;
  const __amdgpu_buffer_rsrc_t rs = points_rsrc(pts, total);
  const int start_k = j < total ? lane_query<K, NB, NL>(grids, cell_start, pts, rs, cloud_of_sorted,
                                                        k, j, s_buf[w], lane, mom, total, nsel)
                                : 0;
  wave_append(start_k > 0, j, start_k, list, count);
}

// ---------------------------------------------------------------- wavefront pass
//
// The queries the lane pass leaves, one wavefront per query.  The candidates are
// the (2K+1)^2 x-runs of the block of radius K around the query's cell, each
// trimmed to the cells that can hold a point below D2 (ball_row), taken 64 runs
// at a time; the lanes stride through each chunk's points together (consecutive
// lanes read consecutive points of a run).
//
// Pass 1 counts d2 < D2 (the block's certified range, as in the lane path) in
// kWaveBuckets equal buckets of a histogram the wavefront shares in LDS.  Fewer
// than kk below D2 means the kk-th neighbour may lie outside the block; any kk
// of the block's points bound its distance from above (the kk-th smallest of the
// per-lane minima, distinct points, or the largest d2 when fewer lanes hold one),
// and the next block is the smallest whose certified range holds that bound, so
// it certifies.  That bound (or, for a query whose lane-pass bucket overflowed,
// the bound of that bucket) then narrows the histogram's range: at least kk
// points lie below it, so counting d2 < R2 = min(D2, bound) over the rows
// trimmed to R2 finds the same kk nearest with fewer candidates and finer
// buckets.  A prefix over the buckets gives bucket B of the kk-th; more
// than 64 points in B split it into 256 sub-buckets with one more pass (the
// bucket key floor(256 * d2 * scale) at that level, still monotone in d2).
// Pass 2 adds the points of lower keys to the lane's FP64 moments (each lane its
// share of the block, in block order) and keeps the kk-th's key's entries (d2,
// original index, position) in LDS, where each one's rank among them picks the
// first kk - (count below); a wavefront sum in a fixed order then gives the
// query's moments.  The selected set is exactly the kk smallest (d2, original
// index).  A block that covers the
// grid, a radius above kWaveMaxK, or more than 64 equal keys sends the query to
// the rescan pass.
constexpr int kWaveBuckets = 256;
constexpr int kWaveMaxK = 48;
constexpr int kWaveThreads = 256;

struct WaveLds {
  int hist[kWaveBuckets];
  int pref[65];  // points before run r of the chunk
  int start[64]; // run r's first sorted position
  uint32_t key[64], idx[64];
  int pos[64];
};

// Inclusive prefix sum over the wavefront.
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

__device__ __forceinline__ float wave_max_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// The kk-th smallest of the lanes' values (ties by lane), kk <= 64.
__device__ __forceinline__ float wave_kth_smallest(float v, int kk) {
  const int lane = threadIdx.x & 63;
  int rank = 0;
  for (int u = 0; u < 64; ++u) {
    const float w = __shfl(v, u, 64);
    rank += (w < v || (w == v && u < lane)) ? 1 : 0;
  }
  const int at = __ffsll((unsigned long long)__ballot(rank == kk - 1)) - 1;
  return __shfl(v, at, 64);
}

// Calls f(valid, p, i) for this lane's share of the points of the block of radius
// K that may lie below D2 (i: the sorted position; rows trimmed by ball_row), run
// chunk by run chunk; every lane takes part in every call (valid = false past a
// chunk's end), so f may use ballots.
template <class F>
__device__ __forceinline__ void for_wave_block(WaveLds& L, const GridDesc& g, const int* cs,
                                               const CellPos& cp, int K, float D2,
                                               __amdgpu_buffer_rsrc_t rs, F&& f) {
  const int lane = threadIdx.x & 63;
  const int side = 2 * K + 1, R = side * side;
  for (int r0 = 0; r0 < R; r0 += 64) {
    const int r = r0 + lane;
    int st = 0, len = 0;
    if (r < R) {
      const int z = cp.cz + r / side - K, y = cp.cy + r % side - K;
      int xa, xb;
      ball_row(g, cp, K, z, y, D2, xa, xb);
      if (xa <= xb) {
        const int row = (z * g.ny + y) * g.nx;
        st = cs[row + xa];
        len = cs[row + xb + 1] - st;
      }
    }
    const int incl = wave_incl_scan(len);
    const int total = __shfl(incl, 63, 64);
    if (total == 0) continue;
    wave_lds_sync();  // the previous chunk's reads are done
    L.start[lane] = st;
    L.pref[lane + 1] = incl;
    if (lane == 0) L.pref[0] = 0;
    wave_lds_sync();
    // two points per lane in flight; t past the chunk's end maps to position 0
    // (loaded, not used)
    int rr = 0;
    for (int base = 0; base < total; base += 128) {
      const int t0 = base + lane, t1 = t0 + 64;
      const bool v0 = t0 < total, v1 = t1 < total;
      int i0 = 0, i1 = 0;
      if (v0) {
        while (t0 >= L.pref[rr + 1]) ++rr;  // runs may be empty
        i0 = L.start[rr] + (t0 - L.pref[rr]);
      }
      if (v1) {
        while (t1 >= L.pref[rr + 1]) ++rr;
        i1 = L.start[rr] + (t1 - L.pref[rr]);
      }
      const float4 p0 = load_point(rs, i0), p1 = load_point(rs, i1);
      f(v0, p0, i0);
      f(v1, p1, i1);
    }
  }
}

// Bucket B of the kk-th among the counts in L.hist (lane l holds buckets 4l ..
// 4l + 3) plus `below` points before them: returns false when the counts hold
// fewer than kk; otherwise B, the count before B (below updated) and B's count.
__device__ __forceinline__ bool wave_find_bucket(const WaveLds& L, int kk, int& below, int& B,
                                                 int& nB) {
  const int lane = threadIdx.x & 63;
  int h[4], sum = 0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    h[c] = L.hist[4 * lane + c];
    sum += h[c];
  }
  const int incl = wave_incl_scan(sum) + below;
  if (__shfl(incl, 63, 64) < kk) return false;
  const bool cross = incl >= kk && incl - sum < kk;
  const int cl = __ffsll((unsigned long long)__ballot(cross)) - 1;
  int b = 0, bl = 0, n = 0;
  if (cross) {
    int cum = incl - sum;
    b = -1;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      if (b < 0 && cum + h[c] >= kk) {
        b = 4 * lane + c;
        bl = cum;
        n = h[c];
      }
      cum += h[c];
    }
  }
  B = __builtin_amdgcn_readlane(b, cl);
  below = __builtin_amdgcn_readlane(bl, cl);
  nB = __builtin_amdgcn_readlane(n, cl);
  return true;
}

__global__ void __launch_bounds__(kWaveThreads, 5)
k_normal_wave(const GridDesc* __restrict__ grids, const int* __restrict__ cell_start,
              const float4* __restrict__ pts, const int* __restrict__ cloud_of_sorted, int k,
              const int2* __restrict__ list, const int* __restrict__ count, int64_t total,
              double* __restrict__ mom, int* __restrict__ nsel, int* __restrict__ deferred) {
  __shared__ WaveLds s_w[kWaveThreads / 64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  WaveLds& L = s_w[w];
  const int n = *count;
  const int nwaves = (int)gridDim.x * (kWaveThreads / 64);
  const __amdgpu_buffer_rsrc_t rs = points_rsrc(pts, total);
  for (int qi = (int)blockIdx.x * (kWaveThreads / 64) + w; qi < n; qi += nwaves) {
    const int2 e = list[qi];
    const int64_t j = __builtin_amdgcn_readfirstlane(e.x);
    const float4 q = pts[j];
    const GridDesc g = grids[cloud_of_sorted[j]];
    const int kk = min(k, g.npts);  // at least 3: the lane pass serves the others
    const int* cs = cell_start + g.cell_off;
    const CellPos cp = cell_pos(g, q.x, q.y, q.z);
    bool done = false;
    const int code = __builtin_amdgcn_readfirstlane(e.y);
    int K = code & 0xff;
    // a d2 that at least kk points lie below (INFINITY: none known yet)
    float U = INFINITY;
    for (int attempt = 0; attempt < 8 && !done && K <= kWaveMaxK; ++attempt) {
      const float Lc = block_gap(g, cp, K);
      if (Lc == INFINITY) break;  // covers the grid: the rescan pass
      const float Lm = (Lc - cp.slack) * g.h;
      const float D2 = Lm * Lm * 0.99998f;
      // the lane pass's bucket bound (same block, same D2), widened for rounding
      if (attempt == 0 && (code >> 8) > 0)
        U = D2 * ((float)(code >> 8) / (float)kLaneBuckets) * 1.0001f;
      // The passes count d2 < R2: the certified range, or U below it.  The kk
      // nearest then lie below R2, and the rows trimmed to R2 hold them.
      const float R2 = (U > 1e-30f && U < D2) ? U : D2;
      const float scale = (float)kWaveBuckets / R2;
      if (!(Lm > 0.0f && D2 > 1e-30f && scale < INFINITY)) {
        ++K;
        continue;
      }
      wave_lds_sync();
#pragma unroll
      for (int c = lane; c < kWaveBuckets; c += 64) L.hist[c] = 0;
      // pass 1: the histogram of d2 < R2, and the lane's count, least and largest d2
      int mc = 0;
      float dmin = INFINITY, dmax = 0.0f;
      for_wave_block(L, g, cs, cp, K, R2, rs, [&](bool valid, const float4& p, int) {
        if (!valid) return;
        const float d = flann_d2(q.x, q.y, q.z, p);
        ++mc;
        dmin = fminf(dmin, d);
        dmax = fmaxf(dmax, d);
        if (d < R2) atomicAdd(&L.hist[min((int)(d * scale), kWaveBuckets - 1)], 1);
      });
      wave_lds_sync();
      if (wave_sum_i(mc) < kk) {
        K = 2 * K + 1;
        continue;
      }
      int below = 0, B, nB;
      if (!wave_find_bucket(L, kk, below, B, nB)) {
        // not certified: a block whose certified range holds an upper bound U of
        // the kk-th d2 certifies (its gap is at least its radius in cells)
        const float Ub = (kk <= 64 && popc_ballot(mc > 0) >= kk) ? wave_kth_smallest(dmin, kk)
                                                                  : wave_max_f(dmax);
        const float need = sqrtf(Ub) * g.inv_h * 1.0001f + cp.slack + 1e-3f;
        K = max(K + 1, (int)ceilf(fminf(need, (float)(kWaveMaxK + 1))));
        U = fminf(U, Ub * 1.0001f);  // kk points of this block lie at or below Ub
        continue;
      }
      // The passes below take points of keys up to bucket B only: floor(d2 *
      // scale) <= B < kWaveBuckets - 1 means d2 < (B + 1) R2 / kWaveBuckets, to
      // within the rounding of d2 * scale (1e-4 covers it), so they walk the rows
      // trimmed to that bound (ball_row keeps every cell that may hold such a
      // point).
      const float D2b = B < kWaveBuckets - 1
                            ? fminf(R2, R2 * ((float)(B + 1) / (float)kWaveBuckets) * 1.0001f)
                            : R2;
      // level 1 inside bucket B when it holds more than 64 points
      int SB = 0;
      const bool split = nB > 64;
      if (split) {
        wave_lds_sync();
#pragma unroll
        for (int c = lane; c < kWaveBuckets; c += 64) L.hist[c] = 0;
        wave_lds_sync();
        for_wave_block(L, g, cs, cp, K, D2b, rs, [&](bool valid, const float4& p, int) {
          if (!valid) return;
          const float d = flann_d2(q.x, q.y, q.z, p);
          if (!(d < R2)) return;
          const float x = d * scale;
          if (min((int)x, kWaveBuckets - 1) != B) return;
          atomicAdd(&L.hist[min((int)((x - (float)B) * 256.0f), 255)], 1);
        });
        wave_lds_sync();
        wave_find_bucket(L, kk, below, SB, nB);
        if (nB > 64) break;  // more than 64 equal keys: the rescan pass
      }
      // pass 2: the moments of the lower keys, the kk-th's key kept in LDS; with
      // the split, keys up to (B, SB) lie below (B + (SB + 1) / 256) R2 /
      // kWaveBuckets (x - B and its product by 256 are exact)
      const float D2s =
          split && B < kWaveBuckets - 1
              ? fminf(D2b, R2 * (((float)B + (float)(SB + 1) * (1.0f / 256.0f)) / (float)kWaveBuckets) *
                               1.0001f)
              : D2b;
      Moments acc;
#pragma unroll
      for (int c = 0; c < 9; ++c) acc.v[c] = 0.0;
      int nl = 0;
      for_wave_block(L, g, cs, cp, K, D2s, rs, [&](bool valid, const float4& p, int i) {
        const float d = valid ? flann_d2(q.x, q.y, q.z, p) : INFINITY;
        const float x = d * scale;
        const int b = d < R2 ? min((int)x, kWaveBuckets - 1) : kWaveBuckets;
        const int sb = (split && b == B) ? min((int)((x - (float)B) * 256.0f), 255) : 0;
        const bool lt = b < B || (b == B && sb < SB);
        const bool eq = b == B && sb == SB;
        const uint64_t beq = __ballot(eq);
        if (lt) add_moment(acc, q, p);
        if (eq) {
          const int slot = nl + lane_prefix(beq);
          L.key[slot] = __float_as_uint(d);
          L.idx[slot] = (uint32_t)__float_as_int(p.w);
          L.pos[slot] = i;
        }
        nl += __popcll(beq);
      });
      wave_lds_sync();
      // rank of each kept entry by (d2, original index): unique keys
      if (lane < nl) {
        const uint64_t mine = ((uint64_t)L.key[lane] << 32) | L.idx[lane];
        int rank = 0;
        for (int u = 0; u < nl; ++u) rank += (((uint64_t)L.key[u] << 32) | L.idx[u]) < mine ? 1 : 0;
        if (rank < kk - below) add_moment(acc, q, load_point(rs, L.pos[lane]));
      }
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        const double v = wave_sum_dpp(acc.v[c]);
        if (lane == 0) mom[c * total + j] = v;
      }
      if (lane == 0) nsel[j] = kk;
      done = true;
    }
    if (!done && lane == 0) {
      // deferred[0] counts the list deferred[1..]
      nsel[j] = kk | kNselOrig;
      deferred[1 + atomicAdd(deferred, 1)] = (int)j;
    }
    wave_lds_sync();  // L is reused by the next query
  }
}

// Search pass, per point in cell-sorted order: the neighbour count and either
// the original indices of its K nearest neighbours (kMoments: nbr holds k per
// point, the moments are summed by the solve pass) or the FP64 sum of their
// distances (kSqrtDist: mom holds 1 per point).
template <int MODE>
__global__ void __launch_bounds__(kWaves * 64)
k_normal_moments(const GridDesc* __restrict__ grids, const int* __restrict__ cell_start,
                 const float4* __restrict__ pts, const int* __restrict__ cloud_of_sorted,
                 int64_t total, int k, double* __restrict__ mom, int* __restrict__ nbr,
                 int* __restrict__ nsel, int* __restrict__ deferred) {
  __shared__ float4 s_p[kWaves][kCap];
  __shared__ int s_cb[kWaves][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t base = ((int64_t)xcd_block(blockIdx.x, gridDim.x) * kWaves + w) * 64;
  if (base >= total) return;
  const int64_t jq = base + lane;
  const bool own = jq < total;
  const int nq = (int)min<int64_t>(64, total - base);
  Staged st;
  st.cell = -1;
  st.K = 0;
  st.cnt = 0;
  st.total = 0;
  st.filt = kNoFilter;
  st.lastT = 0u;
  int mine_n = 0;
  for (int qi = 0; qi < nq; ++qi) {
    // wave-uniform addresses: scalar loads, so no vector load issued before the
    // loop stays pending into it (a pending one makes every iteration wait for
    // the previous iteration's stores)
    const float4 q = pts[base + qi];
    const GridDesc g = grids[cloud_of_sorted[base + qi]];
    const int kk = min(k, g.npts);
    Moments acc;
#pragma unroll
    for (int i = 0; i < 9; ++i) acc.v[i] = 0.0;
    int* out = MODE == kMoments ? nbr + (base + qi) * k : nullptr;
    if (kk >= (MODE == kMoments ? 3 : 1) &&
        !select_staged<MODE>(g, cell_start + g.cell_off, pts, q, kk, s_p[w], s_cb[w], st, acc, out)) {
      // deferred[0] counts the list deferred[1..]
      if (lane == 0) deferred[1 + atomicAdd(deferred, 1)] = (int)(base + qi);
      if (lane == qi) mine_n = kk;
      continue;
    }
    if constexpr (MODE == kSqrtDist) {
      const double s = wave_sum_dpp(acc.v[0]);
      if (lane == 0) mom[base + qi] = s;
    }
    if (lane == qi) mine_n = kk;
  }
  if (own) nsel[jq] = mine_n;
}

// Rescan pass for the deferred queries (candidate blocks too large for LDS), one
// wavefront per query, reading the block from memory for every count.
template <int MODE>
__global__ void __launch_bounds__(256)
k_normal_rescan(const GridDesc* __restrict__ grids, const int* __restrict__ cell_start,
                const float4* __restrict__ pts, const int* __restrict__ cloud_of_sorted, int k,
                const int* __restrict__ deferred, double* __restrict__ mom, int* __restrict__ nbr) {
  const int lane = threadIdx.x & 63;
  const int nwaves = (int)(gridDim.x * (blockDim.x >> 6));
  const int count = deferred[0];
  for (int wi = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); wi < count;
       wi += nwaves) {
    const int64_t j = deferred[1 + wi];
    const float4 q = pts[j];
    const GridDesc g = grids[cloud_of_sorted[j]];
    const int kk = min(k, g.npts);
    Moments acc;
#pragma unroll
    for (int i = 0; i < 9; ++i) acc.v[i] = 0.0;
    select_rescan<MODE>(g, cell_start + g.cell_off, pts, q, cell_pos(g, q.x, q.y, q.z), 1, kk, acc,
                        MODE == kMoments ? nbr + j * k : nullptr);
    if constexpr (MODE == kSqrtDist) {
      const double s = wave_sum_dpp(acc.v[0]);
      if (lane == 0) mom[j] = s;
    }
  }
}

// Solve pass, one point per lane: the FP64 moments of its neighbours relative to
// it (from the lane and wavefront passes, or summed here in list order from the
// rescan pass's list of original indices, kNselOrig set), covariance, FP64
// eigen-decomposition, normal (eigenvector of the smallest eigenvalue) oriented
// towards the viewpoint.
__global__ void __launch_bounds__(256)
k_normal_solve(const float4* __restrict__ pts, const int* __restrict__ cloud_of_sorted,
               const float4* __restrict__ raw, const int64_t* __restrict__ cloud_off,
               const int* __restrict__ nbr, int k, const double* __restrict__ mom,
               const int* __restrict__ nsel, int64_t total, const float* __restrict__ vp,
               int flip_all, float4* __restrict__ out) {
  const int64_t jq = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (jq >= total) return;
  const int ns = nsel[jq];
  const int cnt = ns & ~kNselOrig;
  if (cnt < 3) {
    out[jq] = make_float4(NAN, NAN, NAN, NAN);
    return;
  }
  const float4 q = pts[jq];
  Moments acc;
  if (ns & kNselOrig) {
    const float4* R = raw + cloud_off[cloud_of_sorted[jq]];
    const int* list = nbr + jq * k;
#pragma unroll
    for (int i = 0; i < 9; ++i) acc.v[i] = 0.0;
    for (int i = 0; i < cnt; ++i) add_moment(acc, q, R[list[i]]);
  } else {
#pragma unroll
    for (int i = 0; i < 9; ++i) acc.v[i] = mom[i * total + jq];
  }
  const double* m = acc.v;
  const double n = (double)cnt;
  const double mx = m[0] / n, my = m[1] / n, mz = m[2] / n;
  double C[3][3];
  C[0][0] = m[3] / n - mx * mx;
  C[0][1] = m[4] / n - mx * my;
  C[0][2] = m[5] / n - mx * mz;
  C[1][1] = m[6] / n - my * my;
  C[1][2] = m[7] / n - my * mz;
  C[2][2] = m[8] / n - mz * mz;
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
  out[jq] = make_float4((float)nx, (float)ny, (float)nz, (float)curv);
}

// StatisticalOutlierRemoval's value per input point: sum / mean_k in FP64 rounded
// to float at the point's input position; finite flag for the statistics.
__global__ void k_sor_dist(const float4* __restrict__ pts, const int* __restrict__ cloud_of_sorted,
                           const int64_t* __restrict__ cloud_off, const double* __restrict__ sums,
                           int64_t total, int mean_k, float* __restrict__ dist,
                           uint8_t* __restrict__ finite) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= total) return;
  const int64_t i = cloud_off[cloud_of_sorted[j]] + __float_as_int(pts[j].w);
  dist[i] = (float)(sums[j] / (double)mean_k);
  finite[i] = 1;
}

}  // namespace

void launch_sor_distances(const GridDesc* grids, const int* cell_start, const float4* pts,
                          const int* cloud_of_sorted, const int64_t* cloud_off, int64_t npts_total,
                          int mean_k, double* sums, int* nsel, int* deferred, float* dist,
                          uint8_t* finite, hipStream_t s) {
  if (npts_total <= 0) return;
  const int64_t per_block = (int64_t)kWaves * 64;
  const unsigned blocks = (unsigned)((npts_total + per_block - 1) / per_block);
  launch_fill_i32(deferred, 1, 0, s);
  hipLaunchKernelGGL(k_normal_moments<kSqrtDist>, dim3(blocks), dim3(kWaves * 64), 0, s, grids,
                     cell_start, pts, cloud_of_sorted, npts_total, mean_k + 1, sums, (int*)nullptr, nsel,
                     deferred);
  hipLaunchKernelGGL(k_normal_rescan<kSqrtDist>, dim3(1024), dim3(256), 0, s, grids, cell_start,
                     pts, cloud_of_sorted, mean_k + 1, deferred, sums, (int*)nullptr);
  hipLaunchKernelGGL(k_sor_dist, dim3((unsigned)((npts_total + 255) / 256)), dim3(256), 0, s, pts,
                     cloud_of_sorted, cloud_off, sums, npts_total, mean_k, dist, finite);
}

int64_t normals_work_ints(int64_t n) { return 4 * n + 3; }

void launch_normals(const GridDesc* grids, const int* cell_start, const float4* pts,
                    const int* cloud_of_sorted, int64_t npts_total, int k, const float* vp,
                    int flip_all, const float4* raw, const int64_t* cloud_off, int* nbr, double* mom,
                    int* nsel, int* work, float4* nrm_sorted, hipStream_t s) {
  if (npts_total <= 0) return;
  const int64_t n = npts_total;
  // work: [0] the lane pass's list count, [1] the rescan count with its list at
  // [2, 2 + n), the lane pass's list of (point, start radius) at [2 + n, 2 + 3n)
  int* count = work;
  int* deferred = work + 1;
  int2* list = reinterpret_cast<int2*>(work + 2 + n + (n & 1));
  launch_fill_i32(work, 2, 0, s);
  hipLaunchKernelGGL((k_normal_lane<1, kLaneBuckets, 8>), dim3((unsigned)((n + kLaneThreads - 1) / kLaneThreads)),
                     dim3(kLaneThreads), 0, s, grids, cell_start, pts, cloud_of_sorted, n, k, mom,
                     nsel, list, count);
  // persistent over the list (a few percent of the points): at most eight blocks
  // of four wavefronts per CU
  const int64_t wave_blocks = (n + kWaveThreads / 64 - 1) / (kWaveThreads / 64);
  hipLaunchKernelGGL(k_normal_wave, dim3((unsigned)(wave_blocks < 2048 ? wave_blocks : 2048)),
                     dim3(kWaveThreads), 0, s, grids, cell_start, pts, cloud_of_sorted, k,
                     (const int2*)list, (const int*)count, n, mom, nsel, deferred);
  hipLaunchKernelGGL(k_normal_rescan<kMoments>, dim3(1024), dim3(256), 0, s, grids, cell_start, pts,
                     cloud_of_sorted, k, deferred, (double*)nullptr, nbr);
  hipLaunchKernelGGL(k_normal_solve, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pts,
                     cloud_of_sorted, raw, cloud_off, nbr, k, (const double*)mom, nsel, n, vp,
                     flip_all, nrm_sorted);
}

}  // namespace lcr
