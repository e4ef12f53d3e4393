// lcr_search.hpp -- exact nearest-neighbour search over the dense voxel grid.
//
// Replaces KdTreeFLANN::nearestKSearch(k = 1) as used by CorrespondenceEstimation
// (fine_registration.cpp:121) and Registration::getFitnessScore
// (fine_registration.cpp:126).  The search visits Chebyshev rings of cells around
// the query's cell; after ring k every point not yet visited lies outside the
// (2k+1)^3 block, so its distance is at least the query's distance to the block
// boundary on the sides where grid cells remain.  The search stops once the best
// distance is below that bound (the answer is then exact), once the bound exceeds
// the correspondence radius (nothing unvisited can be kept), or once the block
// covers the grid.  Distances use FLANN's float form ((0+dx^2)+dy^2)+dz^2 with
// dx = q - p and no FMA contraction; ties resolve to the smallest original index.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lcr_common.hpp"

namespace lcr {

#pragma clang fp contract(off)

#define LCR_SHD __host__ __device__ __forceinline__

LCR_SHD int f2i(float f) { return __builtin_bit_cast(int, f); }

LCR_SHD float flann_d2(float qx, float qy, float qz, const float4& p) {
  const float dx = qx - p.x;
  const float dy = qy - p.y;
  const float dz = qz - p.z;
  float r = dx * dx;
  r = r + dy * dy;
  r = r + dz * dz;
  return r;
}

// Cell coordinates of a query in cell units (clamped far outside the grid so the
// integer conversion cannot overflow).
struct CellPos {
  float fx, fy, fz;
  int cx, cy, cz;
  float slack;  // cell-unit rounding allowance for the certification bound
};

LCR_SHD CellPos cell_pos(const GridDesc& g, float qx, float qy, float qz) {
  CellPos c;
  c.fx = fminf(fmaxf((qx - g.ox) * g.inv_h, -4.0e6f), 4.0e6f);
  c.fy = fminf(fmaxf((qy - g.oy) * g.inv_h, -4.0e6f), 4.0e6f);
  c.fz = fminf(fmaxf((qz - g.oz) * g.inv_h, -4.0e6f), 4.0e6f);
  c.cx = (int)floorf(c.fx);
  c.cy = (int)floorf(c.fy);
  c.cz = (int)floorf(c.fz);
  // points are binned with the same (p - o) * inv_h arithmetic; cell coordinates
  // of magnitude <= 4096 carry < 5e-4 of rounding, larger ones ~1.2e-7 relative
  c.slack = 1.0e-3f + 1.0e-6f * fmaxf(fabsf(c.fx), fmaxf(fabsf(c.fy), fabsf(c.fz)));
  return c;
}

// Distance in cell units from the query to the nearest face of block k behind
// which unvisited grid cells remain; +inf once the block covers the grid.
LCR_SHD float block_gap(const GridDesc& g, const CellPos& c, int k) {
  float L = INFINITY;
  if (c.cx - k > 0) L = fminf(L, c.fx - (float)(c.cx - k));
  if (c.cx + k + 1 <= g.nx - 1) L = fminf(L, (float)(c.cx + k + 1) - c.fx);
  if (c.cy - k > 0) L = fminf(L, c.fy - (float)(c.cy - k));
  if (c.cy + k + 1 <= g.ny - 1) L = fminf(L, (float)(c.cy + k + 1) - c.fy);
  if (c.cz - k > 0) L = fminf(L, c.fz - (float)(c.cz - k));
  if (c.cz + k + 1 <= g.nz - 1) L = fminf(L, (float)(c.cz + k + 1) - c.fz);
  return L;
}

// Radius (in cells) up to which a bounded query stays in the ring search.
constexpr float kRingOnlyCells = 3.0f;

// First ring that intersects the grid.
LCR_SHD int first_ring(const GridDesc& g, const CellPos& c) {
  const int kx = c.cx < 0 ? -c.cx : (c.cx >= g.nx ? c.cx - g.nx + 1 : 0);
  const int ky = c.cy < 0 ? -c.cy : (c.cy >= g.ny ? c.cy - g.ny + 1 : 0);
  const int kz = c.cz < 0 ? -c.cz : (c.cz >= g.nz ? c.cz - g.nz + 1 : 0);
  return max(kx, max(ky, kz));
}

// Calls f(first_cell, last_cell) for every run of x-contiguous cells of ring k
// (cells at Chebyshev distance exactly k), clipped to the grid.
template <class F>
LCR_SHD void for_ring_runs(const GridDesc& g, const CellPos& c, int k, F&& f) {
  const int z0 = max(c.cz - k, 0), z1 = min(c.cz + k, g.nz - 1);
  const int y0 = max(c.cy - k, 0), y1 = min(c.cy + k, g.ny - 1);
  const int x0 = max(c.cx - k, 0), x1 = min(c.cx + k, g.nx - 1);
  if (z0 > z1 || y0 > y1 || x0 > x1) return;
  for (int z = z0; z <= z1; ++z) {
    const bool zface = (z == c.cz - k) || (z == c.cz + k);
    for (int y = y0; y <= y1; ++y) {
      const int row = (z * g.ny + y) * g.nx;
      if (zface || y == c.cy - k || y == c.cy + k) {
        f(row + x0, row + x1);
      } else {
        if (c.cx - k >= 0) f(row + c.cx - k, row + c.cx - k);
        if (k > 0 && c.cx + k <= g.nx - 1) f(row + c.cx + k, row + c.cx + k);
      }
    }
  }
}

struct NNResult {
  float d2;   // FLANN-form squared distance (+inf if none)
  int id;     // original index of the match (INT_MAX if none)
  int j;      // position of the match in the sorted point array (-1 if none)
  bool tree;  // deferred: finish with a tree traversal from (d2, id, j)
};

// Lower bound, in cell units, of the distance from the query coordinate f (cell
// units) to cell c along one axis, less the rounding slack.
LCR_SHD float axis_gap(float f, int c, int qc, float slack) {
  const float d = (c < qc) ? f - (float)(c + 1) : ((c > qc) ? (float)c - f : 0.0f);
  return fmaxf(d - slack, 0.0f);
}

LCR_SHD void consider(const float4& p, int j, float qx, float qy, float qz, float& best, int& bid,
                      int& bj) {
  const float d = flann_d2(qx, qy, qz, p);
  const int id = f2i(p.w);
  if (d < best || (d == best && id < bid)) {
    best = d;
    bid = id;
    bj = j;
  }
}

// Scans points [s, e) of the sorted array, keeping the lexicographic minimum of
// (d2, original index).  Four loads are issued before any is used, so a run costs
// one memory latency per four points rather than one per point.
LCR_SHD void scan_points(const float4* __restrict__ P, int s, int e, float qx, float qy, float qz,
                         float& best, int& bid, int& bj) {
  int j = s;
  for (; j + 4 <= e; j += 4) {
    const float4 p0 = P[j], p1 = P[j + 1], p2 = P[j + 2], p3 = P[j + 3];
    consider(p0, j, qx, qy, qz, best, bid, bj);
    consider(p1, j + 1, qx, qy, qz, best, bid, bj);
    consider(p2, j + 2, qx, qy, qz, best, bid, bj);
    consider(p3, j + 3, qx, qy, qz, best, bid, bj);
  }
  if (j < e) {
    // the tail: up to three more, loaded together (indices clamped inside the run)
    const int n = e - j;
    const float4 p0 = P[j], p1 = P[j + (n > 1 ? 1 : 0)], p2 = P[j + (n > 2 ? 2 : 0)];
    consider(p0, j, qx, qy, qz, best, bid, bj);
    if (n > 1) consider(p1, j + 1, qx, qy, qz, best, bid, bj);
    if (n > 2) consider(p2, j + 2, qx, qy, qz, best, bid, bj);
  }
}

// Squared distance from q to the box [lo, hi] (0 inside), float.
LCR_SHD float box_d2(const float4& lo, const float4& hi, float qx, float qy, float qz) {
  const float dx = fmaxf(fmaxf(lo.x - qx, qx - hi.x), 0.0f);
  const float dy = fmaxf(fmaxf(lo.y - qy, qy - hi.y), 0.0f);
  const float dz = fmaxf(fmaxf(lo.z - qz, qz - hi.z), 0.0f);
  return (dx * dx + dy * dy) + dz * dz;
}

// 10 bits spread to every third bit.
LCR_SHD uint32_t spread3(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

// Morton code (kTreeKeyBits bits) of a point on a 256^3 lattice over the grid's
// box (the same scale on every axis).  Any order is exact for the search; this
// one makes the tree's boxes compact.  Points of one lattice cell keep their
// cell order (the sort is stable); at a 2 mm point spacing on a 1 m cloud a
// lattice cell holds a few points, and a 1024^3 lattice (30 bits) cost the radix
// sort a fifth pass.
constexpr int kTreeKeyBits = 24;
LCR_SHD uint32_t tree_key(const GridDesc& g, float x, float y, float z) {
  const int n = max(g.nx, max(g.ny, g.nz));
  const float s = 256.0f / (g.h * (float)n);
  const int ix = min(max((int)((x - g.ox) * s), 0), 255);
  const int iy = min(max((int)((y - g.oy) * s), 0), 255);
  const int iz = min(max((int)((z - g.oz) * s), 0), 255);
  return spread3((uint32_t)ix) | (spread3((uint32_t)iy) << 1) | (spread3((uint32_t)iz) << 2);
}

// Exact 1-NN by traversal of the cloud's tree (see Tree).  Depth-first, nearer child
// first, no stack: `order` keeps one bit per depth saying whether the right child
// was entered first.  A subtree is skipped only when its box is farther than the
// bound by a relative margin, so equal distances are still visited.  `limit2`
// (may be +inf) caps the bound for radius-limited searches.
LCR_SHD void bvh_nn(const GridDesc& g, const Tree& tr, float qx, float qy, float qz, float limit2,
                    float& best, int& bid, int& bj) {
  const float4* N = tr.nodes + 2 * g.node_off;
  int tj = -1;  // tree-order index of the best point once a leaf improves it
  const unsigned first_leaf = 1u << g.depth;
  const int end = (int)g.pt_off + g.npts;
  unsigned i = 1, order = 0;
  int depth = 0;
  for (;;) {
    const float bound = fminf(best, limit2) * 1.00002f;
    bool descend = box_d2(N[2 * i], N[2 * i + 1], qx, qy, qz) <= bound;
    if (descend && i >= first_leaf) {
      const int s = (int)g.pt_off + (int)(i - first_leaf) * kLeaf;
      scan_points(tr.P, s, min(s + kLeaf, end), qx, qy, qz, best, bid, tj);
      descend = false;
    }
    if (descend) {
      const unsigned l = 2 * i;
      const float dl = box_d2(N[2 * l], N[2 * l + 1], qx, qy, qz);
      const float dr = box_d2(N[2 * l + 2], N[2 * l + 3], qx, qy, qz);
      const bool right_first = dr < dl;
      ++depth;
      order = right_first ? (order | (1u << depth)) : (order & ~(1u << depth));
      i = right_first ? l + 1 : l;
      continue;
    }
    // next subtree: the sibling if i was entered first, else climb
    for (;;) {
      if (depth == 0) {
        if (tj >= 0) bj = tr.pos[tj];
        return;
      }
      const unsigned right_first = (order >> depth) & 1u;
      if ((i & 1u) == right_first) {
        i ^= 1u;
        break;
      }
      i >>= 1;
      --depth;
    }
  }
}

// Row (z, y) of ring k around the query's cell c: the whole x range of the ring on
// its z and y faces, else the ring's two x ends.  The row is skipped when its
// lower-bound distance exceeds the bound min(best, lim2) (with a relative margin,
// so equal distances are still visited), and its x range is trimmed to the cells
// whose x gap keeps the bound under it.  gz, zface: z's gap and whether z is on a
// z face of the ring.
LCR_SHD void scan_ring_row(const GridDesc& g, const int* __restrict__ cs,
                           const float4* __restrict__ P, const CellPos& c, int k, int z, int y,
                           float gz, bool zface, int x0, int x1, float h2, float lim2, float qx,
                           float qy, float qz, float& best, int& bid, int& bj) {
  const float gy = axis_gap(c.fy, y, c.cy, c.slack);
  // bound on d2 for any point of this row (cell units^2 -> m^2)
  const float thr = fminf(best, lim2) * 1.00002f;
  const float lyz = (gz * gz + gy * gy) * h2;
  if (lyz > thr) return;
  int xa, xb;
  bool two = false;
  if (zface || y == c.cy - k || y == c.cy + k) {
    xa = x0;
    xb = x1;
  } else {
    xa = c.cx - k;
    xb = c.cx + k;
    two = true;
  }
  int xlo = xa, xhi = xb;
  if (thr < INFINITY) {
    const float rem = fmaxf(thr - lyz, 0.0f);
    const float dxc = sqrtf(rem) * g.inv_h * 1.00001f + c.slack;
    xlo = max(xa, (int)ceilf(c.fx - dxc - 1.0f));
    xhi = min(xb, (int)floorf(c.fx + dxc));
  }
  const int row = (z * g.ny + y) * g.nx;
  if (!two) {
    if (xlo <= xhi) scan_points(P, cs[row + xlo], cs[row + xhi + 1], qx, qy, qz, best, bid, bj);
  } else {
    if (xa >= 0 && xa >= xlo && xa <= xhi)
      scan_points(P, cs[row + xa], cs[row + xa + 1], qx, qy, qz, best, bid, bj);
    if (k > 0 && xb <= g.nx - 1 && xb >= xlo && xb <= xhi)
      scan_points(P, cs[row + xb], cs[row + xb + 1], qx, qy, qz, best, bid, bj);
  }
}

// The rows of ring k dealt to `nsub` lanes: this one (`sub`) scans rows sub,
// sub + nsub, ... of the ring's (z, y) rows in row-major order, each pruned by
// its own best.  Taking the minimum (by d2, then id) over the lanes afterwards
// gives what one lane scanning every row finds: a lane's best is never below the
// shared one, so it prunes no more than a single lane would.
LCR_SHD void ring_rows(const GridDesc& g, const int* __restrict__ cs, const float4* __restrict__ P,
                       const CellPos& c, int k, int sub, int nsub, float h2, float lim2, float qx,
                       float qy, float qz, float& best, int& bid, int& bj) {
  const int z0 = max(c.cz - k, 0), z1 = min(c.cz + k, g.nz - 1);
  const int y0 = max(c.cy - k, 0), y1 = min(c.cy + k, g.ny - 1);
  const int x0 = max(c.cx - k, 0), x1 = min(c.cx + k, g.nx - 1);
  if (!(z0 <= z1 && y0 <= y1 && x0 <= x1)) return;
  const int ny = y1 - y0 + 1;
  int z = z0, y = y0 + sub;
  while (y > y1) {
    y -= ny;
    ++z;
  }
  while (z <= z1) {
    const float gz = axis_gap(c.fz, z, c.cz, c.slack);
    const bool zface = (z == c.cz - k) || (z == c.cz + k);
    scan_ring_row(g, cs, P, c, k, z, y, gz, zface, x0, x1, h2, lim2, qx, qy, qz, best, bid, bj);
    y += nsub;
    while (y > y1) {
      y -= ny;
      ++z;
    }
  }
}

// Ball radius (in cells) up to which a seeded query scans its ball's rows instead
// of the rings.
constexpr float kSeedBallCells = 3.0f;

// Whether a bounded search (limit rlim) may stop after ring k with `best`: the
// grid is exhausted, no unvisited cell can hold a point closer than best, or
// none is within rlim.
LCR_SHD bool ring_stop(const GridDesc& g, const CellPos& c, int k, float best, float rlim) {
  const float Lc = block_gap(g, c, k);
  if (Lc == INFINITY) return true;
  const float L = (Lc - c.slack) * g.h;
  if (L > 0.0f && best < L * L * 0.99998f) return true;
  return L > rlim;
}

// Where a bounded ring-only search starts: k0 = the first ring that can hold a
// point, -1 if no cell within rlim of q can (the result is then "none").
LCR_SHD int ring_start(const GridDesc& g, const CellPos& c, float rlim) {
  const int k0 = first_ring(g, c);
  if (k0 > 0 && (block_gap(g, c, k0 - 1) - c.slack) * g.h > rlim) return -1;
  return k0;
}

// Exact 1-NN of q among the grid's points.  With `bounded`, points farther than
// `r` may be skipped (the result is then either the true nearest or a point the
// caller rejects as d2 > r^2).  `seed_j` (or -1) is a known candidate, e.g. the
// previous iteration's match: it only tightens the pruning bound, the answer is
// the same.  Inside each ring, rows and cells whose lower-bound distance exceeds
// the current best are skipped (with a relative margin, so equal distances are
// still visited and the smallest index wins).
//
// After `ring_limit` rings without certification the rest of the search runs as a
// tree traversal (bvh_nn) seeded with the best point found so far, so queries far
// from the cloud cost O(log N) instead of O(distance^3) cells.
//
// With `defer_tree`, the tree part is left to the caller (NNResult::tree), which
// runs it for a whole wavefront at once (bvh_nn_wave); a query that found nothing
// in its first ring is deferred straight away.  A bounded query whose radius is at
// most kRingOnlyCells cells never uses the tree: rows and cells farther than r
// are pruned, and the rings pass r within a few steps.
LCR_SHD NNResult nn_query(const GridDesc& g, const int* __restrict__ cs,
                          const float4* __restrict__ P, const Tree& tr, float qx, float qy,
                          float qz, float r, bool bounded, int seed_j = -1, int ring_limit = 2,
                          bool defer_tree = false) {
  const float4* nodes = tr.nodes;
  NNResult res{INFINITY, 0x7fffffff, -1, false};
  if (g.npts == 0 || !(isfinite(qx) && isfinite(qy) && isfinite(qz))) return res;
  const CellPos c = cell_pos(g, qx, qy, qz);
  const int k0 = first_ring(g, c);
  const int kmax = k0 + max(g.nx, max(g.ny, g.nz));
  const float rlim = r * 1.00001f;
  if (bounded && k0 > 0) {
    const float L = (block_gap(g, c, k0 - 1) - c.slack) * g.h;
    if (L > rlim) return res;
  }
  float best = INFINITY;
  int bid = 0x7fffffff, bj = -1;
  if (seed_j >= 0) scan_points(P, seed_j, seed_j + 1, qx, qy, qz, best, bid, bj);
  const float h2 = g.h * g.h;
  // a radius of at most kRingOnlyCells cells: rows and cells beyond it are pruned
  // and the rings reach it within a few steps, so the tree is never needed
  const bool ring_only = bounded && rlim * g.inv_h <= kRingOnlyCells;
  const float lim2 = bounded ? rlim * rlim : INFINITY;
  // A seed within kSeedBallCells cells, or a limit that close: every point at
  // least as close (within the limit) lies in a cell that the ball of that radius
  // around q touches, so the rows of the ball's cell box are scanned instead of
  // the rings.  The rows outside the box's z and y range are those a ring walk
  // prunes (their gap exceeds the ball's radius, slack included), and so are the
  // cells outside its x range; inside a row scan_ring_row prunes and trims by the
  // current best as the rings do.  The result is the same point, in fewer row
  // iterations per wavefront (at most 2 x 2 rows for a radius under one cell,
  // against ring 0 and the nine rows of ring 1; about 16 for a query with no
  // target point within a limit of 1.6 cells, against rings 0-2; the unbounded
  // fitness pass's seeds up to three cells away as well).
  if (seed_j >= 0) {
    const float thr = fminf(best, lim2) * 1.00002f;
    const float rho = sqrtf(thr) * g.inv_h * 1.00001f + c.slack;
    if (rho < kSeedBallCells) {
      const int za = max((int)floorf(c.fz - rho), 0), zb = min((int)floorf(c.fz + rho), g.nz - 1);
      const int ya = max((int)floorf(c.fy - rho), 0), yb = min((int)floorf(c.fy + rho), g.ny - 1);
      const int x0 = max((int)floorf(c.fx - rho), 0), x1 = min((int)floorf(c.fx + rho), g.nx - 1);
      if (x0 <= x1) {
        for (int z = za; z <= zb; ++z) {
          const float gz = axis_gap(c.fz, z, c.cz, c.slack);
          for (int y = ya; y <= yb; ++y)
            scan_ring_row(g, cs, P, c, 1, z, y, gz, true, x0, x1, h2, lim2, qx, qy, qz, best, bid,
                          bj);
        }
      }
      res.d2 = best;
      res.id = bid;
      res.j = bj;
      return res;
    }
  }
  if (nodes && !ring_only && k0 > ring_limit) {  // far outside the grid: straight to the tree
    if (defer_tree)
      res.tree = true;
    else
      bvh_nn(g, tr, qx, qy, qz, bounded ? rlim * rlim : INFINITY, best, bid, bj);
    res.d2 = best;
    res.id = bid;
    res.j = bj;
    return res;
  }
  for (int k = k0; k <= kmax; ++k) {
    const int z0 = max(c.cz - k, 0), z1 = min(c.cz + k, g.nz - 1);
    const int y0 = max(c.cy - k, 0), y1 = min(c.cy + k, g.ny - 1);
    const int x0 = max(c.cx - k, 0), x1 = min(c.cx + k, g.nx - 1);
    if (z0 <= z1 && y0 <= y1 && x0 <= x1) {
      for (int z = z0; z <= z1; ++z) {
        const float gz = axis_gap(c.fz, z, c.cz, c.slack);
        const bool zface = (z == c.cz - k) || (z == c.cz + k);
        for (int y = y0; y <= y1; ++y)
          scan_ring_row(g, cs, P, c, k, z, y, gz, zface, x0, x1, h2, lim2, qx, qy, qz, best, bid, bj);
      }
    }
    const float Lc = block_gap(g, c, k);
    if (Lc == INFINITY) break;
    const float L = (Lc - c.slack) * g.h;
    if (L > 0.0f && best < L * L * 0.99998f) break;
    if (bounded && L > rlim) break;
    if (nodes && !ring_only && (k - k0 >= ring_limit || (defer_tree && best == INFINITY))) {
      if (defer_tree)
        res.tree = true;
      else
        bvh_nn(g, tr, qx, qy, qz, bounded ? rlim * rlim : INFINITY, best, bid, bj);
      break;
    }
  }
  res.d2 = best;
  res.id = bid;
  res.j = bj;
  return res;
}

// A read-only array seen through the constant address space: a load from it
// whose address is the same in every lane of the wavefront is served by the
// scalar cache, one request per wavefront instead of one per lane.
// (The host pass only parses the device functions that use it.)
#if defined(__HIP_DEVICE_COMPILE__)
typedef const float4 __attribute__((address_space(4)))* UniformF4;
#else
typedef const float4* UniformF4;
#endif

// Tree levels one step of the wavefront walk descends: the 2^e descendants of a
// node e levels down are contiguous in the implicit tree, so one load brings all
// their boxes (4 boxes, 128 bytes).  Three levels (8 boxes) measured slower.
constexpr int kWalkLevels = 2;
constexpr int kWalkFan = 1 << kWalkLevels;
static_assert((kWalkFan - 1) * (27 / kWalkLevels) + kWalkFan <= 64, "walk stack fits a register");

// The set bit of `mask` whose key is smallest (nearest) or largest, lowest bit on
// ties.  Keys and mask are wave-uniform.
__device__ __forceinline__ int walk_pick(uint32_t mask, const float (&key)[kWalkFan], bool nearest) {
  int sel = -1;
  float kb = 0.0f;
#pragma unroll
  for (int c = 0; c < kWalkFan; ++c) {
    const bool better = nearest ? key[c] < kb : key[c] > kb;
    if (((mask >> c) & 1u) && (sel < 0 || better)) {
      sel = c;
      kb = key[c];
    }
  }
  return sel;
}

// Points of leaf node `leaf` against every active lane's query.  A full leaf has
// all of its loads in flight at once, one latency per leaf; selects rather than a
// branch on `active` keep the compiler from sinking the loads into it.
__device__ __forceinline__ void walk_leaf(const GridDesc& g, const Tree& tr, UniformF4 TP,
                                          unsigned leaf, bool active, float qx, float qy,
                                          float qz, float& best, int& bid, int& tj) {
  const int end = (int)g.pt_off + g.npts;
  const int s = (int)g.pt_off + (int)(leaf - (1u << g.depth)) * kLeaf;
  if (s + kLeaf <= end) {
    float4 p[kLeaf];
#pragma unroll
    for (int j = 0; j < kLeaf; ++j) p[j] = TP[s + j];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < kLeaf; ++j) {
      const float d = flann_d2(qx, qy, qz, p[j]);
      const int id = f2i(p[j].w);
      const bool take = active && (d < best || (d == best && id < bid));
      best = take ? d : best;
      bid = take ? id : bid;
      tj = take ? s + j : tj;
    }
  } else if (active && s < end) {
    scan_points(tr.P, s, end, qx, qy, qz, best, bid, tj);
  }
}

// bvh_nn for the lanes of a wavefront at once (`active` lanes; every lane of the
// wavefront must call it).  The traversal is shared and the node index is
// wave-uniform, so node boxes and leaf points are read once per wavefront through
// the scalar cache.  Queries of one wavefront are neighbouring cell-sorted source
// points, so their paths mostly coincide.
//
// Each step expands a node kWalkLevels levels at once (the root by depth mod
// kWalkLevels, so that every later step ends exactly on the leaves) and keeps the
// descendants whose box some active lane may find a closer point in (bound with
// the same relative margin as bvh_nn).  Leaves are then scanned nearest first,
// each tested again against the tightened bounds from registers; inner nodes are
// pushed farthest first on a stack that holds one entry per lane of a register,
// the nearest is entered at once, and a popped node is tested again before it is
// expanded.  "Nearest" is by the first active lane's box distance.  A step costs
// one load latency where a binary walk spends one per level and another per
// sibling.  The stack never holds more than kWalkFan - 1 entries per inner step
// along one path: at most 3 * 13 + 1 < 64 for the 27 levels of a tree over 2^31
// points.
__device__ __forceinline__ void bvh_nn_wave(const GridDesc& g, const Tree& tr, bool active,
                                            float qx, float qy, float qz, float limit2,
                                            float& best, int& bid, int& bj) {
  const uint64_t live = __ballot(active);
  if (live == 0) return;
  const int leader = __ffsll((unsigned long long)live) - 1;
  const UniformF4 N = (UniformF4)(tr.nodes + 2 * g.node_off);
  const UniformF4 TP = (UniformF4)tr.P;
  const int D = g.depth;
  int tj = -1;  // tree-order index of the best point once a leaf improves it
  const int lane = (int)__lane_id();
  int stack = 0, sp = 0;
  unsigned i = 1;
  bool fresh = false;  // i passed its test under the current bounds
  for (;;) {
    if (!fresh) {
      const float bound = fminf(best, limit2) * 1.00002f;
      if (__ballot(active && box_d2(N[2 * i], N[2 * i + 1], qx, qy, qz) <= bound) == 0) {
        if (sp == 0) break;
        i = (unsigned)__builtin_amdgcn_readlane(stack, --sp);
        continue;
      }
    }
    const int di = 31 - __clz((int)i);
    if (di == D) {  // the root of a one-leaf tree
      walk_leaf(g, tr, TP, i, active, qx, qy, qz, best, bid, tj);
      break;
    }
    const int e = (di == 0 && D % kWalkLevels) ? D % kWalkLevels : kWalkLevels;
    const unsigned c0 = i << e;
    const int nc = 1 << e;
    float dc[kWalkFan], key[kWalkFan];
    if (e == kWalkLevels) {
      float4 bx[2 * kWalkFan];
#pragma unroll
      for (int c = 0; c < 2 * kWalkFan; ++c) bx[c] = N[2 * c0 + c];
#pragma unroll
      for (int c = 0; c < kWalkFan; ++c) dc[c] = box_d2(bx[2 * c], bx[2 * c + 1], qx, qy, qz);
    } else {
#pragma unroll
      for (int c = 0; c < kWalkFan; ++c)
        dc[c] = c < nc ? box_d2(N[2 * (c0 + c)], N[2 * (c0 + c) + 1], qx, qy, qz) : INFINITY;
    }
    const float bound = fminf(best, limit2) * 1.00002f;
    uint32_t pass = 0;
#pragma unroll
    for (int c = 0; c < kWalkFan; ++c) {
      key[c] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(dc[c]), leader));
      if (c < nc && __ballot(active && dc[c] <= bound) != 0) pass |= 1u << c;
    }
    if (di + e == D) {
      while (pass) {
        const int c = walk_pick(pass, key, true);
        pass &= ~(1u << c);
        float d = dc[0];
#pragma unroll
        for (int k = 1; k < kWalkFan; ++k) d = (c == k) ? dc[k] : d;
        const float bnd = fminf(best, limit2) * 1.00002f;
        if (__ballot(active && d <= bnd) != 0)
          walk_leaf(g, tr, TP, c0 + (unsigned)c, active, qx, qy, qz, best, bid, tj);
      }
    } else if (pass) {
      const int cn = walk_pick(pass, key, true);
      pass &= ~(1u << cn);
      while (pass) {
        const int c = walk_pick(pass, key, false);
        pass &= ~(1u << c);
        stack = (lane == sp) ? (int)(c0 + (unsigned)c) : stack;
        ++sp;
      }
      i = c0 + (unsigned)cn;
      fresh = true;
      continue;
    }
    if (sp == 0) break;
    i = (unsigned)__builtin_amdgcn_readlane(stack, --sp);
    fresh = false;
  }
  if (tj >= 0) bj = tr.pos[tj];
}

// ------------------------------------------------------------------ reductions

// v + (v moved by the DPP pattern `ctrl` inside each row of 16 lanes), in FP64.
template <int ctrl>
__device__ __forceinline__ double dpp_add(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xF, 0xF, true);
  return v + __hiloint2double(hi, lo);
}

// Sum over the 64 lanes, returned in every lane: inside each row of 16 lanes by
// DPP (quad swaps, then rotations by 4 and 8), then the four row totals through
// readlane.  Same value in every lane; the order of the additions is fixed.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v = dpp_add<0xB1>(v);   // quad_perm [1,0,3,2]
  v = dpp_add<0x4E>(v);   // quad_perm [2,3,0,1]
  v = dpp_add<0x124>(v);  // row_ror:4
  v = dpp_add<0x128>(v);  // row_ror:8
  double r[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    r[k] = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 16 * k),
                            __builtin_amdgcn_readlane(__double2loint(v), 16 * k));
  return (r[0] + r[1]) + (r[2] + r[3]);
}

// Block-wide sum of N doubles per thread (every thread of the block calls it);
// the totals are valid in thread 0.  `lds` must hold N * (blockDim.x / 64)
// doubles.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double* lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int i = 0; i < N; ++i) v[i] = wave_sum_dpp(v[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < N; ++i) lds[wid * N + i] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) {
#pragma unroll
      for (int i = 0; i < N; ++i) v[i] += lds[w * N + i];
    }
  }
  __syncthreads();
}

}  // namespace lcr
