// grid.hip -- K1: dense voxel grid build for a batch of clouds.
//
// Replaces KdTreeFLANN::setInputCloud (run by Registration::initCompute when the
// target changes; fine_registration.cpp:109,121) and the search-tree build of
// NormalEstimation (normal_estimation.cpp:89-90).  Per cloud: bounding box ->
// per-cell counts (with occupancy) -> exclusive scan into global point offsets ->
// scatter into a cell-sorted float4 array -> sort each cell by original index, so
// the layout does not depend on atomic ordering.  All kernels are HBM-bound
// integer work: one pass over the points per step, coalesced.
#include <hip/hip_runtime.h>

#include "lcr_common.hpp"
#include "lcr_kernels.hpp"

namespace lcr {

namespace {

constexpr int kBBoxThreads = 256;

__device__ __forceinline__ bool finite3(const float4& p) {
  return isfinite(p.x) && isfinite(p.y) && isfinite(p.z);
}

// One block per (cloud, chunk): partial min/max of finite points.
__global__ void __launch_bounds__(kBBoxThreads)
k_bbox_partial(const float4* __restrict__ raw, const int64_t* __restrict__ cloud_off, int nclouds,
               int chunks_per_cloud, float* __restrict__ part) {
  const int cloud = blockIdx.x / chunks_per_cloud;
  const int chunk = blockIdx.x % chunks_per_cloud;
  const int64_t b = cloud_off[cloud], e = cloud_off[cloud + 1];
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  float cnt = 0.0f;
  for (int64_t i = b + (int64_t)chunk * blockDim.x + threadIdx.x; i < e;
       i += (int64_t)chunks_per_cloud * blockDim.x) {
    const float4 p = raw[i];
    if (!finite3(p)) continue;
    mn[0] = fminf(mn[0], p.x);
    mn[1] = fminf(mn[1], p.y);
    mn[2] = fminf(mn[2], p.z);
    mx[0] = fmaxf(mx[0], p.x);
    mx[1] = fmaxf(mx[1], p.y);
    mx[2] = fmaxf(mx[2], p.z);
    cnt += 1.0f;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      mn[a] = fminf(mn[a], __shfl_xor(mn[a], off, 64));
      mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off, 64));
    }
    cnt += __shfl_xor(cnt, off, 64);
  }
  __shared__ float s[kBBoxThreads / 64][7];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    s[w][0] = mn[0];
    s[w][1] = mn[1];
    s[w][2] = mn[2];
    s[w][3] = mx[0];
    s[w][4] = mx[1];
    s[w][5] = mx[2];
    s[w][6] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
      for (int a = 0; a < 3; ++a) {
        s[0][a] = fminf(s[0][a], s[k][a]);
        s[0][3 + a] = fmaxf(s[0][3 + a], s[k][3 + a]);
      }
      s[0][6] += s[k][6];
    }
    float* o = part + (int64_t)blockIdx.x * 7;
    for (int a = 0; a < 7; ++a) o[a] = s[0][a];
  }
}

// One thread per cloud: fold the chunk partials.  Output per cloud: min xyz,
// max xyz, finite count.
__global__ void k_bbox_final(const float* __restrict__ part, int nclouds, int chunks_per_cloud,
                             float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nclouds) return;
  float r[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0f};
  for (int k = 0; k < chunks_per_cloud; ++k) {
    const float* p = part + ((int64_t)c * chunks_per_cloud + k) * 7;
    for (int a = 0; a < 3; ++a) {
      r[a] = fminf(r[a], p[a]);
      r[3 + a] = fmaxf(r[3 + a], p[3 + a]);
    }
    r[6] += p[6];
  }
  for (int a = 0; a < 7; ++a) out[c * 7 + a] = r[a];
}

__device__ __forceinline__ int cell_of(const GridDesc& g, const float4& p) {
  int x = (int)floorf((p.x - g.ox) * g.inv_h);
  int y = (int)floorf((p.y - g.oy) * g.inv_h);
  int z = (int)floorf((p.z - g.oz) * g.inv_h);
  x = min(max(x, 0), g.nx - 1);
  y = min(max(y, 0), g.ny - 1);
  z = min(max(z, 0), g.nz - 1);
  return (z * g.ny + y) * g.nx + x;
}

// Per point: global cell index (or -1 for non-finite points) and per-cell counts.
// counter[key] += 1 for every calling lane, one atomic per distinct key in the
// wavefront (raw points in image order share cells, so a wavefront holds a few
// keys); returns the counter's value before the lane's own increment.  It
// ballots: every lane of the wavefront that is still running must call it.
__device__ __forceinline__ int wave_counter_add(int* __restrict__ counter, int key) {
  const int lane = (int)(threadIdx.x & 63);
  uint64_t todo = __ballot(true);
  int result = 0;
  while (todo) {
    const int lead = __ffsll((unsigned long long)todo) - 1;
    const int k = __shfl(key, lead, 64);
    const uint64_t same = __ballot(key == k) & todo;
    int base = 0;
    if (lane == lead) base = atomicAdd(&counter[k], __popcll(same));
    base = __shfl(base, lead, 64);
    if ((same >> lane) & 1ull)
      result = base + __popcll(same & ((1ull << lane) - 1ull));
    todo &= ~same;
  }
  return result;
}

__global__ void k_cell_count(const float4* __restrict__ raw, const int* __restrict__ point_cloud,
                             const GridDesc* __restrict__ grids, int64_t n,
                             int* __restrict__ cid, int* __restrict__ counts,
                             int* __restrict__ occupied) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = raw[i];
  if (!finite3(p)) {
    cid[i] = -1;
    return;
  }
  const int c = point_cloud[i];
  const GridDesc g = grids[c];
  const int gc = (int)g.cell_off + cell_of(g, p);
  cid[i] = gc;
  wave_counter_add(counts, gc);
}

// occupied[cloud] = number of non-empty cells of the cloud's grid (one atomic per
// block after a block reduction)
__global__ void k_occupancy(const GridDesc* __restrict__ grids, const int* __restrict__ counts,
                            int* __restrict__ occupied) {
  __shared__ int s[256];
  const GridDesc g = grids[blockIdx.y];
  const int64_t ncells = (int64_t)g.nx * g.ny * g.nz;
  int a = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < ncells; i += (int64_t)gridDim.x * 256)
    a += counts[g.cell_off + i] > 0 ? 1 : 0;
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0 && s[0] > 0) atomicAdd(&occupied[blockIdx.y], s[0]);
}

constexpr int kScanThreads = 1024;
constexpr int kScanPerThread = 4;
constexpr int kScanTile = kScanThreads * kScanPerThread;

__device__ __forceinline__ int block_excl_scan(int v, int* lds, int& total) {
  // inclusive wave scan with shuffles, then a scan over wave totals
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  if (w == 0) {
    const int nw = blockDim.x >> 6;
    int t = lane < nw ? lds[lane] : 0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(t, off, 64);
      if (lane >= off) t += y;
    }
    if (lane < nw) lds[lane] = t;
  }
  __syncthreads();
  const int before = w > 0 ? lds[w - 1] : 0;
  total = lds[(blockDim.x >> 6) - 1];
  __syncthreads();
  return before + x - v;
}

// CONSUME: the input is set to zero as it is read.
template <bool CONSUME>
__global__ void __launch_bounds__(kScanThreads)
k_scan_tiles(int* __restrict__ in, int* __restrict__ out, int64_t n, int* __restrict__ sums) {
  __shared__ int lds[kScanThreads / 64];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanPerThread;
  int v[kScanPerThread];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  if constexpr (CONSUME) {
#pragma unroll
    for (int k = 0; k < kScanPerThread; ++k)
      if (base + k < n) in[base + k] = 0;
  }
  int total;
  int run = block_excl_scan(s, lds, total);
#pragma unroll
  for (int k = 0; k < kScanPerThread; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Single block: exclusive scan of the tile sums, in chunks with a carry.
__global__ void __launch_bounds__(kScanThreads) k_scan_sums(int* __restrict__ sums, int n) {
  __shared__ int lds[kScanThreads / 64];
  int carry = 0;
  for (int b = 0; b < n; b += kScanThreads) {
    const int i = b + threadIdx.x;
    const int v = i < n ? sums[i] : 0;
    int total;
    const int ex = block_excl_scan(v, lds, total);
    if (i < n) sums[i] = carry + ex;
    carry += total;
  }
}

__global__ void k_scan_add(int* __restrict__ out, int64_t n, const int* __restrict__ sums) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] += sums[i / kScanTile];
}

// Cell-sorted placement, step 1: each finite point takes a slot of its cell in
// arbitrary (atomic) order and leaves its raw index and cell there.
__global__ void k_scatter(const int* __restrict__ cid, int64_t n, const int* __restrict__ cell_start,
                          int* __restrict__ fill, int* __restrict__ slot_raw,
                          int* __restrict__ slot_cell) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cid[i];
  if (c < 0) return;
  const int pos = cell_start[c] + wave_counter_add(fill, c);
  slot_raw[pos] = (int)i;
  slot_cell[pos] = c;
}

// Step 2, one thread per slot: the point's final position is its cell's start plus
// the number of the cell's points with a smaller raw index (raw order within a
// cloud is original-index order), so points of a cell end up sorted by original
// index whatever order step 1 produced.  The thread of a cell's first slot sets
// the cell's step-1 counter back to zero (empty cells were never counted), so the
// counters are all zero again for the next build.  keep_w: the input points
// already carry their original index in w (an earlier build's sorted points).
__global__ void k_place(const float4* __restrict__ raw, const float4* __restrict__ raw_nrm,
                        const int* __restrict__ slot_raw, const int* __restrict__ slot_cell,
                        int64_t n_sorted, const int* __restrict__ cell_start,
                        const int64_t* __restrict__ cloud_off, const int* __restrict__ point_cloud,
                        float4* __restrict__ pts, float4* __restrict__ nrm, int* __restrict__ fill,
                        int keep_w) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_sorted) return;
  const int c = slot_cell[j], key = slot_raw[j];
  const int b = cell_start[c], e = cell_start[c + 1];
  if (j == b) fill[c] = 0;
  int rank = 0;
  for (int k = b; k < e; ++k) rank += slot_raw[k] < key ? 1 : 0;
  float4 p = raw[key];
  if (!keep_w) p.w = __int_as_float((int)(key - cloud_off[point_cloud[key]]));
  pts[b + rank] = p;
  if (raw_nrm) nrm[b + rank] = raw_nrm[key];
}

// Bounding-volume tree over each cloud's cell-sorted points: leaf i of cloud c
// covers sorted points [16 i, 16 i + 16); padding leaves are empty boxes.  Level
// l holds nodes 2^l .. 2^(l+1) - 1 of a 1-based heap, the leaves at l = depth.
//
// One launch builds kBvhLevels levels: block x of cloud y takes 256 nodes of the
// input level depth - kBvhLevels * pass (the leaves, boxed from the points, when
// pass is 0) and combines them in LDS level by level up to its subtree's root,
// writing every node on the way.  A tree of depth d takes ceil(d / kBvhLevels)
// launches (at least one), where one launch per level took d + 1.
constexpr int kBvhThreads = 256;
constexpr int kBvhLevels = 8;
static_assert(kBvhThreads == 1 << kBvhLevels, "a block's subtree spans kBvhLevels levels");

__global__ void __launch_bounds__(kBvhThreads)
k_bvh_levels(const GridDesc* __restrict__ grids, int pass, const float4* __restrict__ pts,
             float4* __restrict__ nodes) {
  __shared__ float4 slo[kBvhThreads], shi[kBvhThreads];
  const GridDesc g = grids[blockIdx.y];
  const int in_level = g.depth - kBvhLevels * pass;
  if (in_level < 0) return;  // uniform across the block
  const int64_t first = (int64_t)blockIdx.x * kBvhThreads;
  const int64_t width = (int64_t)1 << in_level;
  if (first >= width) return;  // uniform across the block
  const int count = (int)min<int64_t>(kBvhThreads, width);
  float4* N = nodes + 2 * g.node_off;
  const int t = (int)threadIdx.x;
  float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
  float4 hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
  if (t < count) {
    const int64_t node = width + first + t;
    if (pass == 0) {
      const int64_t s = (first + t) * kLeaf, e = min<int64_t>(s + kLeaf, g.npts);
      for (int64_t j = s; j < e; ++j) {
        const float4 p = pts[g.pt_off + j];
        lo.x = fminf(lo.x, p.x);
        lo.y = fminf(lo.y, p.y);
        lo.z = fminf(lo.z, p.z);
        hi.x = fmaxf(hi.x, p.x);
        hi.y = fmaxf(hi.y, p.y);
        hi.z = fmaxf(hi.z, p.z);
      }
      N[2 * node] = lo;
      N[2 * node + 1] = hi;
    } else {
      lo = N[2 * node];
      hi = N[2 * node + 1];
    }
  }
  slo[t] = lo;
  shi[t] = hi;
  __syncthreads();
  int level = in_level, n = count;
  int64_t base = first;  // index of the block's first node within its level
  while (n > 1 && level > 0) {
    n >>= 1;
    --level;
    base >>= 1;
    float4 a, b;
    if (t < n) {
      const float4 l0 = slo[2 * t], l1 = shi[2 * t], r0 = slo[2 * t + 1], r1 = shi[2 * t + 1];
      a = make_float4(fminf(l0.x, r0.x), fminf(l0.y, r0.y), fminf(l0.z, r0.z), 0.f);
      b = make_float4(fmaxf(l1.x, r1.x), fmaxf(l1.y, r1.y), fmaxf(l1.z, r1.z), 0.f);
      const int64_t node = ((int64_t)1 << level) + base + t;
      N[2 * node] = a;
      N[2 * node + 1] = b;
    }
    __syncthreads();
    if (t < n) {
      slo[t] = a;
      shi[t] = b;
    }
    __syncthreads();
  }
}

// grid owning a global cell_start entry (grids in cell_off order)
__device__ __forceinline__ int grid_of_cell(const GridDesc* __restrict__ g, int n, int64_t c) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (g[mid].cell_off <= c)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// The near mask, one bit per cell, each x-row of a grid padded to whole 32-bit
// words (near_row_words); one row of blocks per grid (blockIdx.y), grid-stride
// over its words.  Cells within kd cells of an occupied cell along every axis are
// marked, kd = ceil(r / h) + 1 (at most 24, so an x-dilation reaches only the
// neighbouring words).

// Occupancy: bit b of word w of row (y, z) is set when cell x = 32 w + b holds a
// point (padding bits past nx stay clear).  One wavefront per 64 cells of a row:
// a lane per cell reads its two cell-table entries (coalesced), a ballot forms
// two words.
__global__ void k_near_occ(const GridDesc* __restrict__ grids, const int* __restrict__ cs,
                           uint32_t* __restrict__ out) {
  const GridDesc& gd = grids[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int W = near_row_words(gd.nx);
  const int segs = (gd.nx + 63) >> 6;
  const int64_t items = (int64_t)gd.ny * gd.nz * segs;
  const int* c = cs + gd.cell_off;
  uint32_t* o = out + gd.near_off;
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
  for (int64_t it = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); it < items;
       it += nwaves) {
    const int64_t row = it / segs;
    const int x = (int)(it - row * segs) * 64 + lane;
    bool occ = false;
    if (x < gd.nx) {
      const int64_t cell = row * gd.nx + x;
      occ = c[cell + 1] > c[cell];
    }
    const uint64_t m = __ballot(occ);
    const int wx = (x - lane) / 32 + lane;  // lanes 0 and 1 write the segment's two words
    if (lane < 2 && wx < W) o[row * W + wx] = (uint32_t)(m >> (32 * lane));
  }
}

// Dilation along x: out bit x = OR of in bits x - kd .. x + kd of the row.
__global__ void k_near_x(const GridDesc* __restrict__ grids, const uint32_t* __restrict__ in,
                         uint32_t* __restrict__ out, float r) {
  const GridDesc& gd = grids[blockIdx.y];
  const int W = near_row_words(gd.nx);
  const int64_t nwords = (int64_t)W * gd.ny * gd.nz;
  const int kd = (int)ceilf(r * 1.0001f * gd.inv_h) + 1;
  const uint32_t* i0 = in + gd.near_off;
  uint32_t* o = out + gd.near_off;
  for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
       w += (int64_t)gridDim.x * blockDim.x) {
    const int wx = (int)(w % W);
    const uint32_t m = i0[w];
    const uint32_t L = wx > 0 ? i0[w - 1] : 0u;
    const uint32_t R = wx + 1 < W ? i0[w + 1] : 0u;
    uint32_t acc = m;
    for (int d = 1; d <= kd; ++d) acc |= (m << d) | (L >> (32 - d)) | (m >> d) | (R << (32 - d));
    o[w] = acc;
  }
}

// Dilation along y (axis 1) or z (axis 2): whole words of the rows within kd.
__global__ void k_near_yz(const GridDesc* __restrict__ grids, const uint32_t* __restrict__ in,
                          uint32_t* __restrict__ out, int axis, float r) {
  const GridDesc& gd = grids[blockIdx.y];
  const int W = near_row_words(gd.nx);
  const int64_t nwords = (int64_t)W * gd.ny * gd.nz;
  const int kd = (int)ceilf(r * 1.0001f * gd.inv_h) + 1;
  const int dim = axis == 1 ? gd.ny : gd.nz;
  const int64_t stride = axis == 1 ? (int64_t)W : (int64_t)W * gd.ny;
  const uint32_t* i0 = in + gd.near_off;
  uint32_t* o = out + gd.near_off;
  for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
       w += (int64_t)gridDim.x * blockDim.x) {
    const int pos = axis == 1 ? (int)((w / W) % gd.ny) : (int)(w / ((int64_t)W * gd.ny));
    const int lo = max(pos - kd, 0), hi = min(pos + kd, dim - 1);
    uint32_t acc = 0u;
    for (int p = lo; p <= hi; ++p) acc |= i0[w + (int64_t)(p - pos) * stride];
    o[w] = acc;
  }
}

// Every grid's near_r and the first word of its mask (the host computes the same).
__global__ void k_set_near(GridDesc* grids, int n, float r) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int64_t off = 0;
  for (int i = 0; i < n; ++i) {
    grids[i].near_r = r;
    grids[i].near_off = off;
    off += (int64_t)near_row_words(grids[i].nx) * grids[i].ny * grids[i].nz;
  }
}

// sorted-order normals back to input order: raw[cloud_off[c] + id] = nrm[j]
__global__ void k_unsort_normals(const float4* __restrict__ pts, const float4* __restrict__ nrm,
                                 const int* __restrict__ cloud_of_sorted,
                                 const int64_t* __restrict__ cloud_off, int64_t n,
                                 float4* __restrict__ raw_nrm) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int id = __float_as_int(pts[j].w);
  raw_nrm[cloud_off[cloud_of_sorted[j]] + id] = nrm[j];
}

__global__ void k_fill_f4(float4* __restrict__ p, int64_t n, float4 v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_i32(int* __restrict__ p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void launch_bbox(const float4* raw, const int64_t* cloud_off, int nclouds, float* part,
                 float* out, hipStream_t s) {
  constexpr int chunks = 64;
  hipLaunchKernelGGL(k_bbox_partial, dim3(nclouds * chunks), dim3(kBBoxThreads), 0, s, raw,
                     cloud_off, nclouds, chunks, part);
  hipLaunchKernelGGL(k_bbox_final, dim3(blocks_for(nclouds, 64)), dim3(64), 0, s, part, nclouds,
                     chunks, out);
}

void launch_bvh_build(const GridDesc* grids, int nclouds, int max_depth, const float4* pts,
                      float4* nodes, hipStream_t s) {
  if (nclouds <= 0) return;
  for (int pass = 0; pass == 0 || kBvhLevels * pass < max_depth; ++pass) {
    const int64_t width = (int64_t)1 << (max_depth - kBvhLevels * pass);
    hipLaunchKernelGGL(k_bvh_levels, dim3(blocks_for(width, kBvhThreads), nclouds),
                       dim3(kBvhThreads), 0, s, grids, pass, pts, nodes);
  }
}

void launch_unsort_normals(const float4* pts, const float4* nrm, const int* cloud_of_sorted,
                           const int64_t* cloud_off, int64_t n_sorted, float4* raw_nrm,
                           int64_t n_raw, hipStream_t s) {
  if (n_raw > 0)
    hipLaunchKernelGGL(k_fill_f4, dim3(blocks_for(n_raw, 256)), dim3(256), 0, s, raw_nrm, n_raw,
                       make_float4(NAN, NAN, NAN, NAN));
  if (n_sorted > 0)
    hipLaunchKernelGGL(k_unsort_normals, dim3(blocks_for(n_sorted, 256)), dim3(256), 0, s, pts, nrm,
                       cloud_of_sorted, cloud_off, n_sorted, raw_nrm);
}

void launch_near_mask(const GridDesc* grids, int ngrids, const int* cell_start, int64_t total_words,
                      float r, uint32_t* out, uint32_t* tmp, hipStream_t s) {
  if (total_words <= 0 || ngrids <= 0) return;
  // per grid a row of blocks, grid-stride over its words
  const unsigned bx = (unsigned)std::min<int64_t>(blocks_for(total_words / ngrids + 1, 256), 256);
  const dim3 grid(bx, (unsigned)ngrids);
  // occupancy: one wavefront item per 64 cells (two words), each a round trip to
  // the cell table; enough blocks that a wavefront takes a few items (with the
  // words' grid each took ~32 in turn: 74 us per cycle36 build)
  const unsigned bo = (unsigned)std::min<int64_t>(blocks_for(total_words / ngrids / 2 + 1, 4), 1024);
  hipLaunchKernelGGL(k_near_occ, dim3(bo, (unsigned)ngrids), dim3(256), 0, s, grids, cell_start, tmp);
  hipLaunchKernelGGL(k_near_x, grid, dim3(256), 0, s, grids, (const uint32_t*)tmp, out, r);
  hipLaunchKernelGGL(k_near_yz, grid, dim3(256), 0, s, grids, (const uint32_t*)out, tmp, 1, r);
  hipLaunchKernelGGL(k_near_yz, grid, dim3(256), 0, s, grids, (const uint32_t*)tmp, out, 2, r);
}

void launch_set_near(GridDesc* grids, int ngrids, float r, hipStream_t s) {
  if (ngrids <= 0) return;
  hipLaunchKernelGGL(k_set_near, dim3(1), dim3(64), 0, s, grids, ngrids, r);
}

void launch_fill_i32(int* p, int64_t n, int v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
}

void launch_cell_count(const float4* raw, const int* point_cloud, const GridDesc* grids,
                       int64_t n, int* cid, int* counts, int* occupied, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cell_count, dim3(blocks_for(n, 256)), dim3(256), 0, s, raw, point_cloud,
                     grids, n, cid, counts, occupied);
}

void launch_occupancy(const GridDesc* grids, int nclouds, const int* counts, int* occupied,
                      hipStream_t s) {
  if (nclouds <= 0) return;
  hipLaunchKernelGGL(k_occupancy, dim3(64, nclouds), dim3(256), 0, s, grids, counts, occupied);
}

int64_t scan_tmp_ints(int64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

void launch_exclusive_scan(const int* in, int* out, int64_t n, int* tmp, hipStream_t s) {
  if (n <= 0) return;
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_tiles<false>, dim3((unsigned)tiles), dim3(kScanThreads), 0, s,
                     const_cast<int*>(in), out, n, tmp);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, tmp, (int)tiles);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(n, 256)), dim3(256), 0, s, out, n, tmp);
}

void launch_exclusive_scan_consume(int* in, int* out, int64_t n, int* tmp, hipStream_t s) {
  if (n <= 0) return;
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_tiles<true>, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, in, out,
                     n, tmp);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanThreads), 0, s, tmp, (int)tiles);
  hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(n, 256)), dim3(256), 0, s, out, n, tmp);
}

void launch_place_points(const float4* raw, const float4* raw_nrm, const int* cid, int64_t n_raw,
                         int64_t n_sorted, const int* cell_start, int* fill, int* slot_raw,
                         int* slot_cell, const int64_t* cloud_off, const int* point_cloud,
                         float4* pts, float4* nrm, bool keep_w, hipStream_t s) {
  if (n_raw <= 0) return;
  hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n_raw, 256)), dim3(256), 0, s, cid, n_raw,
                     cell_start, fill, slot_raw, slot_cell);
  if (n_sorted <= 0) return;
  hipLaunchKernelGGL(k_place, dim3(blocks_for(n_sorted, 256)), dim3(256), 0, s, raw, raw_nrm,
                     slot_raw, slot_cell, n_sorted, cell_start, cloud_off, point_cloud, pts, nrm,
                     fill, keep_w ? 1 : 0);
}

}  // namespace lcr
