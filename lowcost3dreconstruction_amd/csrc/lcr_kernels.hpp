// lcr_kernels.hpp -- host-side launchers of the HIP kernels (all asynchronous on
// the given stream).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lcr_common.hpp"

namespace lcr {

// ---- K1 grid build (grid.hip)
void launch_bbox(const float4* raw, const int64_t* cloud_off, int nclouds, float* part,
                 float* out, hipStream_t s);
constexpr int kBBoxChunks = 64;  // partial boxes per cloud (part holds 7 floats each)
void launch_fill_i32(int* p, int64_t n, int v, hipStream_t s);
void launch_cell_count(const float4* raw, const int* point_cloud, const GridDesc* grids,
                       int64_t n, int* cid, int* counts, int* occupied, hipStream_t s);
void launch_occupancy(const GridDesc* grids, int nclouds, const int* counts, int* occupied,
                      hipStream_t s);
int64_t scan_tmp_ints(int64_t n);
void launch_exclusive_scan(const int* in, int* out, int64_t n, int* tmp, hipStream_t s);
// The same, setting the input to zero as it is read.
void launch_exclusive_scan_consume(int* in, int* out, int64_t n, int* tmp, hipStream_t s);
// Places every finite input point at its cell-sorted position (cells in order,
// points of a cell by input index); fill must be zero on entry and is zero again
// on exit, slot_* hold n_sorted ints.  The input is the raw upload (w set to the
// point's index in its cloud) or, with keep_w, an earlier build's sorted points
// (w already holds it).
void launch_place_points(const float4* raw, const float4* raw_nrm, const int* cid, int64_t n_raw,
                         int64_t n_sorted, const int* cell_start, int* fill, int* slot_raw,
                         int* slot_cell, const int64_t* cloud_off, const int* point_cloud,
                         float4* pts, float4* nrm, bool keep_w, hipStream_t s);
// Dilated occupancy per cell, one bit (set: some point within r of the cell):
// occupancy, then three separable passes; tmp and out hold total_words words
// (the grids' near_off must be set first, launch_set_near).
void launch_near_mask(const GridDesc* grids, int ngrids, const int* cell_start, int64_t total_words,
                      float r, uint32_t* out, uint32_t* tmp, hipStream_t s);
// Sets every grid's near_r and its mask's first word (near_off).
void launch_set_near(GridDesc* grids, int ngrids, float r, hipStream_t s);
void launch_unsort_normals(const float4* pts, const float4* nrm, const int* cloud_of_sorted,
                           const int64_t* cloud_off, int64_t n_sorted, float4* raw_nrm,
                           int64_t n_raw, hipStream_t s);
// bounding-volume trees (GridDesc::node_off / depth set by the host)
// Tree order (Tree): Morton keys of the cell-sorted points, a radix sort of
// (cloud, key) carrying each point's cell-sorted position into tpos, then the
// points gathered into tpts.  keys, keys_sorted, vals: npts each; tmp: at least
// tree_order_tmp_bytes(npts, nclouds).
size_t tree_order_tmp_bytes(int64_t npts, int nclouds);
hipError_t launch_tree_order(const GridDesc* grids, int nclouds, const float4* pts,
                             const int* cloud_of_sorted, int64_t npts, uint64_t* keys,
                             uint64_t* keys_sorted, int* vals, int* tpos, void* tmp, size_t tmp_bytes,
                             float4* tpts, hipStream_t s);
void launch_bvh_build(const GridDesc* grids, int nclouds, int max_depth, const float4* pts,
                      float4* nodes, hipStream_t s);

// ---- K2-K6 ICP iteration, fitness, output (icp.hip)
// Tile counters of the fitness passes: kFetchGroups groups of blocks, then the
// two fill counters of the final pass's reordered list, each counter on its own
// 128-byte line (kFetchStride ints apart).
constexpr int kFetchGroups = 256;
constexpr int kFetchStride = 32;
constexpr int kFetchLines = kFetchGroups + 2;
constexpr int kFetchCounters = kFetchLines * kFetchStride;

struct IcpBuffers {
  const GridDesc* grids;
  const int* cell_start;
  const float4* pts;      // cell-sorted points of all clouds
  const float4* nrm;      // cell-sorted normals (nullable for point-to-point)
  const float4* nodes;    // bounding-volume trees (2 float4 per node)
  const float4* tpts;     // points in tree order (Tree::P)
  const int* tpos;        // cell-sorted position of each tree-order point
  const PairDesc* pairs;
  const int* tile_pair;   // tile -> instance
  PairState* state;
  double* partials;       // ntiles * kMaxVals
  int* prev;              // ntiles * tile: last match (sorted target position) per source
                          // point, -1 = none; seeds the next search (nullable)
  const uint32_t* near;   // per target cell, one bit: some point within GridDesc::near_r
                          // (nullable; GridDesc::near_off, near_row_words)
  const int* gate_tiles;  // the tiles the gate's pass scores (every gate_stride-th of an
                          // instance), in tile order (nullable: all, skipping the others)
  int ngate_tiles;
  // (tile, instance) lists the search and the final fitness pass run over, in tile
  // order, and their lengths (device memory: rebuilt by launch_compact_tiles as
  // instances stop or are set aside; nullable: every tile)
  const int2* it_list;
  const int* it_count;
  const int2* fit_list;
  const int* fit_count;
  // (kFetchCounters ints) the fitness passes' tile counters (nullable: fixed
  // shares); zero before each pass, and zeroed again by the fold launch that
  // follows it
  int* fetch;
  // room for one list like fit_list: the final pass's tiles reordered with the
  // tree-walk tiles first (nullable: fit_list's order), and one flag per entry
  int2* fit_order;
  int* fit_flags;
  int ntiles;
  int npairs;
  int tile;               // source points per tile (= threads per block)
  int morton;             // 1: the iterations and the final fitness pass take a tile's
                          // points in the source's tree order (tpts), 0: in cell order
  __host__ __device__ Tree tree() const { return Tree{nodes, tpts, tpos}; }
};

// One ICP iteration for every active instance: the correspondence search with
// per-tile moment partials, then the per-instance fold, solve and update.
void launch_icp_search(const IcpBuffers& b, const IcpConst& c, int nblocks, hipStream_t s);
// The same search with four lanes per query, for ring-only instances
// (c.ring_only): the launches with few live tiles.
void launch_icp_search_split(const IcpBuffers& b, const IcpConst& c, int nblocks, hipStream_t s);
int icp_split_blocks_per_cu(int tile, bool p2pl);
void launch_icp_solve(const IcpBuffers& b, const IcpConst& c, hipStream_t s);
// getFitnessScore(max_range) of every instance that is not gated; with
// per_instance, each instance's own max_corr_d2 replaces max_range.
void launch_fitness(const IcpBuffers& b, double max_range, int nblocks, hipStream_t s,
                    bool per_instance = false, int interleave = 0);
// Hypothesis gate over n pairs of H consecutive instances, after a fitness pass
// with each instance's radius: the best keeps running, the others are set aside.
void launch_gate(PairState* st, int npairs, int H, hipStream_t s);
// Rebuilds IcpBuffers::it_list (tiles of active instances) and, with `fit`,
// fit_list (tiles of instances the gate kept) from the instance states; off holds
// npairs ints of scratch.
void launch_compact_tiles(const IcpBuffers& b, int* off, int2* it_list, int* it_count,
                          int2* fit_list, int* fit_count, bool fit, hipStream_t s);
// Initialise instance states from row-major 4x4 guesses (nullable = identity);
// instances with an empty source or target start inactive.
void launch_state_init(PairState* st, const double* guesses, const PairDesc* pairs,
                       const GridDesc* grids, int npairs, hipStream_t s);
// Correspondences of each source point of instance `pair` under its current
// transform: original target index (-1 if d2 > r^2) and d2, in original order.
void launch_correspondences(const IcpBuffers& b, int pair, int npts_src, int32_t* idx, float* d2,
                            hipStream_t s);
// K8: raw source points (original order) transformed by the instance's final
// transform (xyz, and normals when given); non-finite points are copied.
void launch_transform_output(const float4* raw, const float4* raw_nrm, int64_t n,
                             const PairState* st, int pair, float* out_xyz, float* out_nrm,
                             hipStream_t s);
// x' = T x (xyz, FP64, rounded once) and n' = R n for a host-given row-major T.
void launch_transform_points(const float4* raw, const float4* raw_nrm, int64_t n, const double* T,
                             float* out_xyz, float* out_nrm, hipStream_t s);
int icp_max_blocks_per_cu(int tile, bool p2pl, int precision, bool ring_only = false);
// Resident blocks per CU of the fitness pass at `tile` points per block.
int fitness_max_blocks_per_cu(int tile);

// ---- K7 normals (normals.hip)
// Ints of the normals pass's workspace for n points (its two query lists).
int64_t normals_work_ints(int64_t n);
void launch_normals(const GridDesc* grids, const int* cell_start, const float4* pts,
                    const int* cloud_of_sorted, int64_t npts_total, int k, const float* vp,
                    int flip_all, const float4* raw, const int64_t* cloud_off, int* nbr, double* mom,
                    int* nsel, int* work, float4* nrm_sorted, hipStream_t s);

// ---- K11-K13 filters before ICP (filters.hip)
// VoxelGrid lattice: voxel of p = floor(p * inv) - minb (FP32, per axis); index
// x + y * div[0] + z * div[0] * div[1].
struct VoxLattice {
  float inv;
  int minb[3];
  int64_t div[3];
};
// device scratch of one VoxelGrid pass (sizes per field)
struct VoxScratch {
  uint32_t* skey;          // n: voxel per point (nvox: non-finite)
  uint32_t* skey_sorted;   // n
  int* idx;                // n: input index per point
  int* sorted_raw;         // n: input indices in (voxel, input index) order
  int* counts;             // nvox
  int* start;              // nvox
  int* flags;              // nvox
  int* outidx;             // nvox
  int* scan_tmp;           // scan_tmp_ints(nvox)
  void* sort_tmp;          // voxel_sort_tmp_bytes(n, nvox)
  size_t sort_tmp_bytes;
};
size_t voxel_sort_tmp_bytes(int64_t n, int64_t nvox);
hipError_t launch_voxel_grid(const float4* raw, const float4* raw_nrm, const float* curv,
                             const uint8_t* rgb, int64_t n, const VoxLattice& L, int64_t nvox,
                             VoxScratch w, float* out_xyz, float* out_nrm, float* out_curv,
                             uint8_t* out_rgb, hipStream_t s);
// StatisticalOutlierRemoval statistics and keep flags from per-point mean
// distances (finite[i] = 1 for points that count); part holds 3 doubles per block
// of sor_stat_blocks(n), stats 3 doubles (mean, stddev, threshold).
int sor_stat_blocks(int64_t n);
void launch_sor_stats(const float* dist, const uint8_t* finite, int64_t n, double mul, double* part,
                      double* stats, uint8_t* keep, hipStream_t s);
// depth back-projection (16-bit millimetre depth, row-major w x h)
struct DepthParams {
  double fx, fy, cx, cy;
  int depth_min, depth_max, top, bottom, left, right;
};
void launch_depth_flags(const uint16_t* depth, int w, int h, const DepthParams& P, int* flags,
                        hipStream_t s);
void launch_depth_points(const uint16_t* depth, int w, int h, const uint8_t* color,
                         const DepthParams& P, const int* pos, float* out_xyz, float* out_nrm,
                         uint8_t* out_rgb, hipStream_t s);
// *out_n = pos[n - 1] + flags[n - 1]: the number of points (n = w * h > 0)
void launch_depth_count(const int* pos, const int* flags, int64_t n, int64_t* out_n, hipStream_t s);

// StatisticalOutlierRemoval distances (normals.hip's selection with mean_k + 1
// neighbours): dist / finite per input point (the caller zeroes both first);
// sums holds one double per sorted point, deferred npts_total + 1 ints.
void launch_sor_distances(const GridDesc* grids, const int* cell_start, const float4* pts,
                          const int* cloud_of_sorted, const int64_t* cloud_off, int64_t npts_total,
                          int mean_k, double* sums, int* nsel, int* deferred, float* dist,
                          uint8_t* finite, hipStream_t s);

// ---- K9 pose-proposal network (encoder.hip)
struct EncBuffers {
  const float4* X;       // 2 * npairs clouds of npts normalised points (src, tgt, ...)
  float* W;              // FP32 master weights (kEncNParam)
  __bf16* W2t;           // bf16 W2 transposed [128][64]
  __bf16* W3t;           // bf16 W3 transposed [256][128]
  float* part_max;       // [cloud][chunk][256]
  int* part_arg;
  float* F;              // pooled features [cloud][256]
  int* ARG;              // argmax point per feature
  float* Z;              // head hidden [pair][256]
  float* OUT;            // [pair][6]
  float* DOUT;           // [pair][6]
  float* loss_part;      // [pair][6]
  float* DZ;             // [pair][256]
  float* DG;             // [pair][512]
  float* XA;             // rows [cloud*256 + c]: x (3), h1 (64), da1 (64), h2 (128), da2 (128)
  float* H1A;
  float* DA1;
  float* H2A;
  float* DA2;
  float* DF;
  float* part_grad;
  float* G;              // gradient (kEncNParam)
};
void launch_enc_bf16_weights(const float* W, __bf16* W2t, __bf16* W3t, hipStream_t s);
// Forward blocks (and max-pool partials per channel) per cloud of npts points.
int enc_forward_chunks(int npts);
void launch_enc_forward(const EncBuffers& e, int npairs, int npts, const float* labels,
                        hipStream_t s);
void launch_enc_backward(const EncBuffers& e, int npairs, int npts, hipStream_t s);
void launch_adam(float* W, const float* G, float* M, float* V, int n, float lr, int step,
                 hipStream_t s);
int64_t enc_part_grad_floats(int npairs);
// training pairs generated on the device (synth.hip)
void launch_synth_pairs(const float* shapes, const float* weights, const double* P,
                        const double* Pinv, int npairs, int npts, float noise, uint64_t seed,
                        float4* X, float* labels, float* raw, float* frame, hipStream_t s);

}  // namespace lcr
