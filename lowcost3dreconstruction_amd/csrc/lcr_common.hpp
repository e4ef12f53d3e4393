// lcr_common.hpp -- device data layout shared by the HIP kernels and the host
// runtime of liblcr.  Everything lives in HBM as:
//   * per cloud: float4 points sorted by voxel cell (x, y, z, bit-cast original
//     index), float4 normals in the same order (nx, ny, nz, curvature), and a
//     dense cell-start table over the cloud's bounding box (ncells + 1 int32);
//   * per pair / hypothesis: a PairDesc (which clouds, tile range, radius) and a
//     PairState (FP64 transform, convergence bookkeeping, fitness);
//   * per source tile: FP64 partial sums written by the tile's block.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lcr {

// pcl::registration::DefaultConvergenceCriteria::ConvergenceState
enum ConvState : int {
  kNotConverged = 0,
  kIterations = 1,
  kTransform = 2,
  kAbsMse = 3,
  kRelMse = 4,
  kNoCorrespondences = 5,
  kFailureAfterMaxIterations = 6,
};

// Uniform dense voxel grid over one cloud's bounding box.  Cell (x, y, z) has id
// (z * ny + y) * nx + x; its points are pts[cell_start[id] .. cell_start[id+1]).
// cell_start holds global offsets into the concatenated point array.
struct GridDesc {
  float ox, oy, oz;   // min corner
  float h, inv_h;     // cell edge and its reciprocal
  int nx, ny, nz;
  int64_t cell_off;   // first entry of this grid's cell_start (ncells + 1 entries)
  int64_t pt_off;     // first point of the cloud in the sorted arrays
  int npts;           // finite points in the grid
  int nraw;           // points handed in (including non-finite ones)
  int64_t node_off;   // first node of this cloud's bounding-volume tree (2 float4 each)
  int depth;          // tree depth: leaves are nodes [2^depth, 2^(depth+1)), 1-based
  float near_r;       // radius of the cell mask "some point within near_r" (0 = no mask)
  int64_t near_off;   // first 32-bit word of this grid's cell mask (near_row_words per x-row)
};

// 32-bit words per x-row of a grid's cell mask: bit x & 31 of word x >> 5.
__host__ __device__ inline int near_row_words(int nx) { return (nx + 31) >> 5; }

// Points per leaf of the bounding-volume tree over a cloud's points.
constexpr int kLeaf = 16;

// The cloud's bounding-volume tree: an implicit complete binary tree (node i has
// children 2i and 2i+1; empty padding leaves have lo > hi) over runs of kLeaf
// points in tree order -- the cloud's points sorted by the Morton code of their
// position (tree_key), so that every node bounds a compact region.  P holds the
// points in that order (pt_off-based like the cell-sorted array) and pos[i] the
// cell-sorted position of P[i]; nodes == nullptr: no tree.
struct Tree {
  const float4* nodes;
  const float4* P;
  const int* pos;
};

// One ICP instance (a pair, or one pose hypothesis of a pair).
struct PairDesc {
  int src;            // cloud index of the source
  int tgt;            // cloud index of the target
  int tile_begin;     // first work tile of this instance
  int ntiles;         // tiles = ceil(source points / tile size)
  double max_corr_d2; // setMaxCorrespondenceDistance^2
};

// Per-instance state, updated on the device by each iteration's solve launch.
struct PairState {
  double T[16];        // composed transform (row-major), FP64
  double prev_mse;     // DefaultConvergenceCriteria::correspondences_prev_mse_
  double last_mse;     // MSE of the last iteration's correspondences
  double fitness;      // getFitnessScore result
  int iterations;      // nr_iterations_
  int state;           // ConvState
  int active;          // 1 while the instance still iterates
  int converged;       // Registration::converged_
  int similar;         // iterations_similar_transforms_
  int ncorr;           // correspondences of the last iteration
  int solve_failed;    // 6x6 system not solvable in the last iteration
  int fit_count;       // points counted by the fitness pass
  int gated;           // 1: set aside by the hypothesis gate (never refined or chosen)
};

struct IcpConst {
  int max_iterations;
  int point_to_plane;
  double rotation_threshold;     // 1 - transformation_epsilon
  double translation_threshold;  // transformation_epsilon
  double mse_rel;                // euclidean_fitness_epsilon
  double mse_abs;                // 1e-12
  int precision;                 // correspondence search: 0 FP32, 1 FP16 / 2 BF16 tiles
  int interleave;                // tile order: 0 XCD-contiguous, 1 grid stride, 2 by live tiles
  int ring_only;                 // 1: every instance's radius spans at most kRingOnlyCells
                                 // target cells, so the search never needs the tree
  int morton;                    // 1: a tile's points in the source's tree (Morton) order
                                 // (compact patches), 0: in cell order
};

// The hypothesis gate scores each pose on every s-th work tile of the source,
// s = clamp(ntiles / kGateMinTiles, 1, kGateStride): at least 1/32 of the tiles,
// and at least kGateMinTiles tiles where the source has them.  32 tiles (2048
// points) halve the gate's pass against 64 on 74 pairs x 15 hypotheses.
constexpr int kGateStride = 32;
constexpr int kGateMinTiles = 32;
__host__ __device__ inline int gate_stride(int ntiles) {
  const int s = ntiles / kGateMinTiles;
  return s < 1 ? 1 : (s > kGateStride ? kGateStride : s);
}

// Flat FP32 weight layout of the pose-proposal network, row-major [in][out]:
// W1 3x64, b1, W2 64x128, b2, W3 128x256, b3, Hw1 512x256, hb1, Hw2 256x6, hb2.
constexpr int kEncW1 = 0;
constexpr int kEncB1 = kEncW1 + 3 * 64;
constexpr int kEncW2 = kEncB1 + 64;
constexpr int kEncB2 = kEncW2 + 64 * 128;
constexpr int kEncW3 = kEncB2 + 128;
constexpr int kEncB3 = kEncW3 + 128 * 256;
constexpr int kEncHw1 = kEncB3 + 256;
constexpr int kEncHb1 = kEncHw1 + 512 * 256;
constexpr int kEncHw2 = kEncHb1 + 256;
constexpr int kEncHb2 = kEncHw2 + 256 * 6;
constexpr int kEncNParam = kEncHb2 + 6;
static_assert(kEncNParam == 174470, "parameter count");

// FP64 partial sums per tile.
//   point-to-point: [0] n, [1..3] sum s, [4..6] sum t, [7..15] sum s_a t_b (a-major),
//                   [16] sum d2
//   point-to-plane: [0] n, [1..21] upper triangle of A^T A (row-major),
//                   [22..27] A^T b, [28] sum d2
constexpr int kP2PVals = 17;
constexpr int kP2LVals = 29;
constexpr int kMaxVals = 29;

}  // namespace lcr
