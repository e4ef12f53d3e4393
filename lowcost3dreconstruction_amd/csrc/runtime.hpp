// runtime.hpp -- host runtime of liblcr: device buffers, the cloud store (upload +
// batched grid build + normals), the batched ICP engine (one persistent launch
// per iteration, captured once into a hipGraph), and the kernel profiler.
#pragma once
#include <hip/hip_runtime.h>

#include <string.h>

#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "lcr_common.hpp"
#include "lcr_kernels.hpp"

struct lcr_proposal;

namespace lcr {

struct Error {
  int code;
  std::string msg;
};

#define LCR_HIP(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::lcr::Error{-2, std::string(#expr) + " failed: " + hipGetErrorString(e_)};  \
  } while (0)

// Device allocation that only grows.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void reserve(size_t n) {
    if (n == 0) n = 256;
    if (n <= bytes) return;
    if (p) LCR_HIP(hipFree(p));
    p = nullptr;
    bytes = 0;
    LCR_HIP(hipMalloc(&p, n));
    bytes = n;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Pinned host allocation that only grows.
struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  HostBuf() = default;
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
  void reserve(size_t n) {
    if (n == 0) n = 256;
    if (n <= bytes) return;
    if (p) LCR_HIP(hipHostFree(p));
    p = nullptr;
    bytes = 0;
    LCR_HIP(hipHostMalloc(&p, n, hipHostMallocDefault));
    bytes = n;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Per-kernel GPU time from HIP events recorded around each launch on its stream.
class Profiler {
 public:
  struct Span {
    std::string name;
    hipEvent_t a, b;
  };
  static Profiler& get();
  bool on() const { return on_; }
  void enable(bool v) { on_ = v; }
  void reset();
  // Records a begin event; returns the span index (or -1 when off).
  int begin(const char* name, hipStream_t s, std::vector<Span>* into);
  void end(int span, hipStream_t s, std::vector<Span>* into);
  // Adds the elapsed times of completed spans to the totals (events stay owned
  // by `spans` when `keep` is set, e.g. spans recorded inside a graph).
  void collect(std::vector<Span>& spans, bool keep);
  bool read(const std::string& name, int64_t* launches, double* ms);

 private:
  hipEvent_t take();
  bool on_ = false;
  std::mutex mu_;
  std::map<std::string, std::pair<int64_t, double>> totals_;
  std::vector<hipEvent_t> pool_;
};

// A set of clouds resident on one device: raw points (input order) and the
// cell-sorted layout with one dense grid per cloud.
struct CloudSet {
  int n = 0;
  std::vector<int64_t> nraw, raw_off;  // raw_off: n + 1 prefix
  int64_t total_raw = 0;
  bool has_nrm = false;
  std::vector<GridDesc> grids;
  int64_t total_pts = 0, total_cells = 0, total_nodes = 0;
  // per cloud: bounding box + finite count (7 floats) and occupied probe cells,
  // measured by the first build after an upload
  std::vector<float> geom_box;
  std::vector<int> geom_occ;
  bool geom_valid = false;
  // recorded after the last copy out of the pinned build region / viewpoint
  // buffer: the next build or normals pass waits on it before rewriting them
  // (instead of synchronising the stream at the end of each)
  hipEvent_t ev_build = nullptr, ev_vp = nullptr;
  // radix-sort scratch size for (tree order points, clouds)
  int64_t ttmp_pts = -1;
  int ttmp_n = -1;
  size_t ttmp_bytes = 0;
  CloudSet() = default;
  CloudSet(const CloudSet&) = delete;
  CloudSet& operator=(const CloudSet&) = delete;
  ~CloudSet();
  int max_depth = 0;

  DevBuf d_raw, d_raw_nrm, d_point_cloud, d_cloud_off;
  DevBuf d_grids, d_counts, d_cell_start, d_cid, d_scan_tmp, d_bbox_part, d_bbox, d_occ;
  // d_counts is zero over its whole allocation once the stream has run the work
  // issued so far (each build leaves it so; the geometry probe does not)
  bool counts_zero = false;
  DevBuf d_pts, d_nrm, d_cloud_of_sorted, d_vp, d_poff, d_nodes, d_nbr, d_nsel;
  DevBuf d_npts, d_nnrm;  // the normals stage's sorted points and normals
  DevBuf d_tpts, d_tpos, d_tkeys, d_tkeys_sorted, d_tvals, d_ttmp;  // tree order
  DevBuf d_slot_raw, d_slot_cell, d_near, d_near_tmp;
  DevBuf d_nwork;  // normals: the query lists of the wavefront and rescan passes
  DevBuf d_nmom;   // normals: FP64 neighbour moments, nine arrays of one per point
  HostBuf h_stage, h_stage_nrm, h_small, h_build, h_vp;

  // Packs and uploads clouds (x, y, z at the start of each `stride` bytes record;
  // normals at the same stride from their own pointer).
  void upload(hipStream_t s, const std::vector<const float*>& xyz,
              const std::vector<const float*>& nrm, const std::vector<int64_t>& counts,
              int64_t stride_bytes);
  // Builds every cloud's grid; cell edge = factor * estimated point spacing.
  // with_tree: also the tree order and the bounding-volume tree (the 1-NN searches
  // need them; the k-NN selections of normals and outlier removal do not)
  // stage: kBuildRaw sorts the uploaded points into d_pts (d_nrm); kBuildNormals
  // sorts them into d_npts for compute_normals(..., true); kBuildFromNormals sorts
  // d_npts and the normals computed there (d_nnrm) into d_pts / d_nrm, so the
  // second grid of a run starts from spatially ordered points and the normals
  // need no trip back to input order.
  static constexpr int kBuildRaw = 0, kBuildNormals = 1, kBuildFromNormals = 2;
  void build(hipStream_t s, double factor, std::vector<Profiler::Span>* spans, bool with_tree = true,
             int stage = kBuildRaw);
  // K7 normals for every cloud into d_nrm (cell-sorted order); normals_stage: over
  // d_npts into d_nnrm (after a kBuildNormals build).
  void compute_normals(hipStream_t s, int k, const float* vp_host, int flip_all,
                       std::vector<Profiler::Span>* spans, bool normals_stage = false);
  // Copies the cell-sorted normals back to input order (d_raw_nrm), so the next
  // build() carries them into its own order.
  void normals_to_raw(hipStream_t s);
  // Per grid cell: 1 if some point of the cloud may lie within r of a position in
  // the cell (d_near; GridDesc::near_r = r).  Skipped (near_r = 0) for an
  // infinite r or one spanning more than kMaxNearCells cells.
  void build_near_mask(hipStream_t s, double r);
  static constexpr int kMaxNearCells = 24;
};

// A batch of ICP instances over a CloudSet.
struct Engine {
  int npairs = 0, ntiles = 0, tile = 256, nblocks = 8;
  int nblocks_ring = 8;  // persistent blocks of the ring-only search variant
  // The four-lanes-per-query search (launch_icp_search_split) runs the chunks of
  // ring-only instances that start with at most split_units live tiles, counted
  // in 64-point units, on nblocks_split blocks.
  int split_units = 0, nblocks_split = 8;
  bool use_split() const { return c.ring_only && split_units > 0; }
  bool split_fits(int64_t live_tiles) const {
    return use_split() && live_tiles * (tile / 64) <= split_units;
  }
  // IcpConst::ring_only for the instances over these grids: the FP32 scan, and
  // every radius finite and spanning at most kRingOnlyCells target cells
  bool ring_only(const CloudSet& cs) const;
  int nblocks_fit = 8;  // persistent blocks of the fitness passes (their own occupancy)
  IcpConst c{};
  std::vector<PairDesc> hp;
  DevBuf d_pairs, d_tile_pair, d_state, d_partials, d_guess, d_prev, d_gate_tiles;
  DevBuf d_it_list, d_fit_list, d_list_count, d_list_off;  // live-tile lists (IcpBuffers::it_list)
  DevBuf d_fetch;      // IcpBuffers::fetch
  DevBuf d_fit_order;  // IcpBuffers::fit_order
  DevBuf d_fit_flags;  // IcpBuffers::fit_flags
  int ngate_tiles = 0;
  HostBuf h_state, h_guess;
  static constexpr int kChunk = 4;  // iterations per captured chunk
  DevBuf d_active;
  HostBuf h_active;
  hipGraph_t graph = nullptr, tail_graph = nullptr, split_graph = nullptr;
  hipGraphExec_t exec = nullptr, tail_exec = nullptr, split_exec = nullptr;
  struct GraphKey {
    IcpBuffers b{};
    IcpConst c{};
    int nblocks = 0, nblocks_fit = 0, nblocks_split = 0;
    const void* guess = nullptr;
    const void* hstate = nullptr;
    const void* active = nullptr;
    const void* hactive = nullptr;
    bool operator==(const GraphKey& o) const { return memcmp(this, &o, sizeof *this) == 0; }
  };
  GraphKey graph_key_{};
  GraphKey graph_key(const IcpBuffers& b) const;
  static void capture(hipStream_t s, const std::function<void()>& f, hipGraph_t* g,
                      hipGraphExec_t* e);

  ~Engine();
  void reset_graph();
  // (src, tgt) cloud indices per instance; squared radius per instance.
  void setup(const CloudSet& cs, const std::vector<int>& src, const std::vector<int>& tgt,
             const std::vector<double>& r2, int tile_pts);
  int setup_precision = -1;  // c.precision the launch shape was sized for
  // > 1: instances come in groups of gate_h hypotheses of one pair; before the
  // iterations, each group keeps only its best hypothesis by inliers within r.
  int gate_h = 0;
  IcpBuffers buffers(const CloudSet& cs) const;
  // Runs state init (from guesses, 16 doubles per instance or null) and up to
  // max_iterations iterations plus the fitness pass; results land in h_state.
  // `pre` (optional) runs on stream s after the guesses reach the device and
  // before the first iteration (it may rewrite d_guess).
  void run(hipStream_t s, const CloudSet& cs, const double* guesses, bool use_graph,
           const std::function<void()>& pre = {});
  void run_fitness(hipStream_t s, const CloudSet& cs, double max_range);
  const PairState* states() const { return h_state.as<PairState>(); }
};

int device_cu_count();

// Records the message returned by lcr_last_error() on this thread; returns code.
int set_last_error(int code, const std::string& msg);

// Takes the address of the fixed scratch array of lcr_depth_to_cloud_noalloc on
// `device`, which must be current (noalloc.hip).
void noalloc_init(int device);

// The pose proposal behind an lcr_proposal handle, and its per-cycle inference:
// writes each pair's learned guess into slot `slot` of d_guess (pair-major,
// `stride` slots per pair, slot 0 = the plain seed), on stream s.
struct ProposalImpl;
ProposalImpl* proposal_impl(lcr_proposal* p);
void proposal_guesses(ProposalImpl& P, hipStream_t s, const CloudSet& cs, const int* d_src,
                      const int* d_tgt, int npairs, int npts, double* d_guess, int stride, int slot);

}  // namespace lcr
