// runtime.cpp -- host runtime of liblcr (see runtime.hpp).
#include "runtime.hpp"

#include "lcr_search.hpp"

#include <atomic>
#include <thread>

#include <float.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>

namespace lcr {

// ------------------------------------------------------------------ profiler

Profiler& Profiler::get() {
  static Profiler p;
  return p;
}

hipEvent_t Profiler::take() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  LCR_HIP(hipEventCreate(&e));
  return e;
}

void Profiler::reset() {
  std::lock_guard<std::mutex> g(mu_);
  totals_.clear();
}

int Profiler::begin(const char* name, hipStream_t s, std::vector<Span>* into) {
  if (!on_ || !into) return -1;
  std::lock_guard<std::mutex> g(mu_);
  Span sp{name, take(), take()};
  LCR_HIP(hipEventRecord(sp.a, s));
  into->push_back(sp);
  return (int)into->size() - 1;
}

void Profiler::end(int span, hipStream_t s, std::vector<Span>* into) {
  if (span < 0 || !into) return;
  LCR_HIP(hipEventRecord((*into)[span].b, s));
}

void Profiler::collect(std::vector<Span>& spans, bool keep) {
  std::lock_guard<std::mutex> g(mu_);
  for (auto& sp : spans) {
    LCR_HIP(hipEventSynchronize(sp.b));
    float ms = 0.f;
    LCR_HIP(hipEventElapsedTime(&ms, sp.a, sp.b));
    auto& t = totals_[sp.name];
    t.first += 1;
    t.second += ms;
    if (!keep) {
      pool_.push_back(sp.a);
      pool_.push_back(sp.b);
    }
  }
  if (!keep) spans.clear();
}

bool Profiler::read(const std::string& name, int64_t* launches, double* ms) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = totals_.find(name);
  if (it == totals_.end()) {
    *launches = 0;
    *ms = 0.0;
    return false;
  }
  *launches = it->second.first;
  *ms = it->second.second;
  return true;
}

int device_cu_count() {
  int dev = 0;
  LCR_HIP(hipGetDevice(&dev));
  int cus = 0;
  LCR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return cus > 0 ? cus : 1;
}

// ------------------------------------------------------------------ clouds

namespace {

__global__ void k_segments(const int64_t* __restrict__ off, int n, int64_t total,
                           int* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  int lo = 0, hi = n - 1;
  while (lo < hi) {  // last segment whose start is <= i
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  out[i] = lo;
}

void launch_segments(const int64_t* off, int n, int64_t total, int* out, hipStream_t s) {
  if (total <= 0) return;
  hipLaunchKernelGGL(k_segments, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, off, n,
                     total, out);
}

constexpr int64_t kMaxDim = 4096;

}  // namespace

// Host threads staging an upload, and the points below which one thread does it.
constexpr int kStageThreads = 8;
constexpr int64_t kStageMinPoints = 1 << 18;

void CloudSet::upload(hipStream_t s, const std::vector<const float*>& xyz,
                      const std::vector<const float*>& nrm, const std::vector<int64_t>& counts,
                      int64_t stride_bytes) {
  n = (int)xyz.size();
  geom_valid = false;
  nraw = counts;
  raw_off.assign(n + 1, 0);
  for (int c = 0; c < n; ++c) raw_off[c + 1] = raw_off[c] + counts[c];
  total_raw = raw_off[n];
  has_nrm = true;
  for (int c = 0; c < n; ++c)
    if (!nrm[c]) has_nrm = false;
  const size_t np = (size_t)std::max<int64_t>(total_raw, 1);
  h_stage.reserve(np * sizeof(float4));
  float4* st = h_stage.as<float4>();
  int* pc = nullptr;
  h_stage_nrm.reserve(np * (sizeof(float4) + sizeof(int)));
  float4* sn = h_stage_nrm.as<float4>();
  pc = reinterpret_cast<int*>(sn + np);
  // host staging into pinned memory, clouds spread over a few threads
  auto stage = [&](int c) {
    const char* base = reinterpret_cast<const char*>(xyz[c]);
    const char* nb = reinterpret_cast<const char*>(nrm[c]);
    for (int64_t i = 0; i < counts[c]; ++i) {
      const float* p = reinterpret_cast<const float*>(base + i * stride_bytes);
      st[raw_off[c] + i] = make_float4(p[0], p[1], p[2], 0.0f);
      pc[raw_off[c] + i] = c;
      if (has_nrm) {
        const float* m = reinterpret_cast<const float*>(nb + i * stride_bytes);
        sn[raw_off[c] + i] = make_float4(m[0], m[1], m[2], 0.0f);
      }
    }
  };
  const int nthreads = (int)std::min<int64_t>(
      {(int64_t)kStageThreads, (int64_t)n, std::max<int64_t>(total_raw / kStageMinPoints, 1)});
  if (nthreads <= 1) {
    for (int c = 0; c < n; ++c) stage(c);
  } else {
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
      pool.emplace_back([&] {
        for (int c = next++; c < n; c = next++) stage(c);
      });
    for (std::thread& th : pool) th.join();
  }
  d_raw.reserve(np * sizeof(float4));
  d_point_cloud.reserve(np * sizeof(int));
  d_cloud_off.reserve((n + 1) * sizeof(int64_t));
  LCR_HIP(hipMemcpyAsync(d_raw.p, st, total_raw * sizeof(float4), hipMemcpyHostToDevice, s));
  LCR_HIP(hipMemcpyAsync(d_point_cloud.p, pc, total_raw * sizeof(int), hipMemcpyHostToDevice, s));
  h_small.reserve((n + 1) * sizeof(int64_t) + 4096);
  memcpy(h_small.p, raw_off.data(), (n + 1) * sizeof(int64_t));
  LCR_HIP(hipMemcpyAsync(d_cloud_off.p, h_small.p, (n + 1) * sizeof(int64_t),
                         hipMemcpyHostToDevice, s));
  if (has_nrm) {
    d_raw_nrm.reserve(np * sizeof(float4));
    LCR_HIP(hipMemcpyAsync(d_raw_nrm.p, sn, total_raw * sizeof(float4), hipMemcpyHostToDevice, s));
  }
  LCR_HIP(hipStreamSynchronize(s));  // staging buffers are reused by the next upload
}

namespace {
// Disjoint regions of the pinned scratch used by one grid build (async copies of
// different regions may be in flight together).
struct BuildRegions {
  size_t bbox, occ, grids, poff, total;
  explicit BuildRegions(int n) {
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    bbox = 0;
    occ = al(bbox + (size_t)n * 7 * sizeof(float));
    grids = al(occ + (size_t)n * sizeof(int));
    poff = al(grids + (size_t)n * sizeof(GridDesc));
    total = al(poff + (size_t)(n + 1) * sizeof(int64_t));
  }
};
}  // namespace

CloudSet::~CloudSet() {
  if (ev_build) (void)hipEventDestroy(ev_build);
  if (ev_vp) (void)hipEventDestroy(ev_vp);
}

void CloudSet::build(hipStream_t s, double factor, std::vector<Profiler::Span>* spans,
                     bool with_tree, int stage) {
  const bool from_normals = stage == kBuildFromNormals;
  // the sorted input's layout (finite points per cloud) comes from the same upload
  if (from_normals && !geom_valid) throw Error{-5, "no normals-stage build to start from"};
  // the previous build's copies out of the pinned region must be done
  if (ev_build)
    LCR_HIP(hipEventSynchronize(ev_build));
  else
    LCR_HIP(hipEventCreateWithFlags(&ev_build, hipEventDisableTiming));
  Profiler& prof = Profiler::get();
  if (factor <= 0.0) factor = 2.5;
  const BuildRegions R(n);
  h_build.reserve(R.total);
  char* hb = h_build.as<char>();
  float* bb = reinterpret_cast<float*>(hb + R.bbox);
  int* occ = reinterpret_cast<int*>(hb + R.occ);
  GridDesc* hg = reinterpret_cast<GridDesc*>(hb + R.grids);
  int64_t* hpoff = reinterpret_cast<int64_t*>(hb + R.poff);
  const int span = prof.begin("grid_build", s, spans);
  // 1. bounding boxes and finite counts (and 2.'s probe occupancy) depend only on
  //    the uploaded points: measured once per upload
  if (!geom_valid) {
    d_bbox_part.reserve((size_t)n * kBBoxChunks * 7 * sizeof(float));
    d_bbox.reserve((size_t)n * 7 * sizeof(float));
    launch_bbox(d_raw.as<float4>(), d_cloud_off.as<int64_t>(), n, d_bbox_part.as<float>(),
                d_bbox.as<float>(), s);
    LCR_HIP(hipMemcpyAsync(bb, d_bbox.p, (size_t)n * 7 * sizeof(float), hipMemcpyDeviceToHost, s));
    LCR_HIP(hipStreamSynchronize(s));
    geom_box.assign(bb, bb + (size_t)n * 7);
  }
  const std::vector<float>& box = geom_box;

  // 2. probe grid (<= 65 cells on the longest axis) to measure occupancy
  grids.assign(n, GridDesc{});
  int64_t cells = 0;
  for (int c = 0; c < n; ++c) {
    GridDesc& g = grids[c];
    const float* b = &box[(size_t)c * 7];
    g.npts = (int)b[6];
    g.nraw = (int)nraw[c];
    if (g.npts == 0) {
      g.ox = g.oy = g.oz = 0.f;
      g.h = 1.f;
      g.inv_h = 1.f;
      g.nx = g.ny = g.nz = 1;
    } else {
      const double ext = std::max({(double)b[3] - b[0], (double)b[4] - b[1], (double)b[5] - b[2]});
      const double h0 = std::max(ext / 64.0, 1e-6);
      g.ox = b[0];
      g.oy = b[1];
      g.oz = b[2];
      g.h = (float)h0;
      g.inv_h = (float)(1.0 / h0);
      g.nx = (int)std::min<double>(floor(((double)b[3] - b[0]) / h0) + 1, 65);
      g.ny = (int)std::min<double>(floor(((double)b[4] - b[1]) / h0) + 1, 65);
      g.nz = (int)std::min<double>(floor(((double)b[5] - b[2]) / h0) + 1, 65);
    }
    g.cell_off = cells;
    g.near_r = 0.0f;
    g.near_off = 0;
    cells += (int64_t)g.nx * g.ny * g.nz + 1;
  }
  d_grids.reserve((size_t)n * sizeof(GridDesc));
  d_cid.reserve((size_t)std::max<int64_t>(total_raw, 1) * sizeof(int));
  if (!geom_valid) {
    d_counts.reserve((size_t)cells * sizeof(int));
    counts_zero = false;  // the probe counts below stay in the array
    d_occ.reserve((size_t)n * sizeof(int));
    memcpy(hg, grids.data(), (size_t)n * sizeof(GridDesc));
    LCR_HIP(hipMemcpyAsync(d_grids.p, hg, (size_t)n * sizeof(GridDesc), hipMemcpyHostToDevice, s));
    launch_fill_i32(d_counts.as<int>(), cells, 0, s);
    launch_fill_i32(d_occ.as<int>(), n, 0, s);
    launch_cell_count(d_raw.as<float4>(), d_point_cloud.as<int>(), d_grids.as<GridDesc>(),
                      total_raw, d_cid.as<int>(), d_counts.as<int>(), nullptr, s);
    launch_occupancy(d_grids.as<GridDesc>(), n, d_counts.as<int>(), d_occ.as<int>(), s);
    LCR_HIP(hipMemcpyAsync(occ, d_occ.p, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s));
    LCR_HIP(hipStreamSynchronize(s));
    geom_occ.assign(occ, occ + n);
    geom_valid = true;
  } else {
    memcpy(occ, geom_occ.data(), (size_t)n * sizeof(int));
  }

  // 3. final cell edge: factor * spacing, spacing ~ h0 / sqrt(points per occupied
  //    probe cell) for surface-sampled clouds; capped cell count and extent
  cells = 0;
  total_pts = 0;
  for (int c = 0; c < n; ++c) {
    GridDesc& g = grids[c];
    const float* b = &box[(size_t)c * 7];
    g.pt_off = total_pts;
    total_pts += g.npts;
    if (g.npts > 0) {
      const double h0 = g.h;
      const double per_cell = (double)g.npts / std::max(occ[c], 1);
      const double spacing = h0 / sqrt(std::max(per_cell, 1.0));
      double h = std::max(factor * spacing, 1e-6);
      const double ex = (double)b[3] - b[0], ey = (double)b[4] - b[1], ez = (double)b[5] - b[2];
      const double max_cells = std::max<double>(16.0 * g.npts, 4096.0);
      for (int it = 0; it < 64; ++it) {
        const double nx = floor(ex / h) + 1, ny = floor(ey / h) + 1, nz = floor(ez / h) + 1;
        if (nx <= kMaxDim && ny <= kMaxDim && nz <= kMaxDim && nx * ny * nz <= max_cells) break;
        h *= 1.25;
      }
      g.h = (float)h;
      g.inv_h = (float)(1.0 / (double)g.h);
      g.nx = (int)(floor(ex * (double)g.inv_h) + 1);
      g.ny = (int)(floor(ey * (double)g.inv_h) + 1);
      g.nz = (int)(floor(ez * (double)g.inv_h) + 1);
    }
    g.cell_off = cells;
    g.near_r = 0.0f;
    g.near_off = 0;
    cells += (int64_t)g.nx * g.ny * g.nz + 1;
  }
  if (cells > (int64_t)1 << 30) throw Error{-5, "grid too large"};
  total_cells = cells;
  // bounding-volume trees: leaves of kLeaf sorted points, complete binary tree
  int64_t nodes = 0;
  max_depth = 0;
  for (int c = 0; c < n; ++c) {
    GridDesc& g = grids[c];
    const int64_t leaves = (g.npts + kLeaf - 1) / kLeaf;
    int d = 0;
    while (((int64_t)1 << d) < leaves) ++d;
    g.depth = d;
    g.node_off = nodes;
    nodes += (int64_t)2 << d;  // 1-based heap: indices 1 .. 2^(d+1) - 1
    max_depth = std::max(max_depth, d);
  }
  total_nodes = nodes;
  memcpy(hg, grids.data(), (size_t)n * sizeof(GridDesc));
  LCR_HIP(hipMemcpyAsync(d_grids.p, hg, (size_t)n * sizeof(GridDesc), hipMemcpyHostToDevice, s));

  // 4. count, scan, scatter, per-cell sort.  The counters are zero over the whole
  // allocation between builds (the scan consumes the counts, the placement
  // clears its slot counters), so they are filled only after a reallocation or
  // the geometry probe.
  const void* counts_before = d_counts.p;
  d_counts.reserve((size_t)cells * sizeof(int));
  if (d_counts.p != counts_before) counts_zero = false;
  d_cell_start.reserve((size_t)cells * sizeof(int));
  d_scan_tmp.reserve((size_t)scan_tmp_ints(cells) * sizeof(int));
  const size_t npts = (size_t)std::max<int64_t>(total_pts, 1);
  DevBuf& out_pts = stage == kBuildNormals ? d_npts : d_pts;
  DevBuf& out_nrm = stage == kBuildNormals ? d_nnrm : d_nrm;
  out_pts.reserve(npts * sizeof(float4));
  out_nrm.reserve(npts * sizeof(float4));
  d_cloud_of_sorted.reserve(npts * sizeof(int));
  // input: the upload, or the normals stage's sorted points (finite only, the
  // clouds in the same order) with the normals computed on them
  const float4* in_pts = from_normals ? d_npts.as<float4>() : d_raw.as<float4>();
  const float4* in_nrm = from_normals ? d_nnrm.as<float4>()
                                      : (has_nrm ? d_raw_nrm.as<float4>() : nullptr);
  const int* in_cloud = from_normals ? d_cloud_of_sorted.as<int>() : d_point_cloud.as<int>();
  const int64_t n_in = from_normals ? total_pts : total_raw;
  if (!counts_zero) launch_fill_i32(d_counts.as<int>(), (int64_t)(d_counts.bytes / sizeof(int)), 0, s);
  counts_zero = false;
  launch_cell_count(in_pts, in_cloud, d_grids.as<GridDesc>(), n_in, d_cid.as<int>(),
                    d_counts.as<int>(), nullptr, s);
  launch_exclusive_scan_consume(d_counts.as<int>(), d_cell_start.as<int>(), cells,
                                d_scan_tmp.as<int>(), s);
  d_slot_raw.reserve(npts * sizeof(int));
  d_slot_cell.reserve(npts * sizeof(int));
  launch_place_points(in_pts, in_nrm, d_cid.as<int>(), n_in, total_pts, d_cell_start.as<int>(),
                      d_counts.as<int>(), d_slot_raw.as<int>(), d_slot_cell.as<int>(),
                      d_cloud_off.as<int64_t>(), d_point_cloud.as<int>(), out_pts.as<float4>(),
                      out_nrm.as<float4>(), from_normals, s);
  if (from_normals) has_nrm = true;
  counts_zero = true;
  // sorted point -> cloud
  for (int c = 0; c < n; ++c) hpoff[c] = grids[c].pt_off;
  hpoff[n] = total_pts;
  d_poff.reserve((size_t)(n + 1) * sizeof(int64_t));
  LCR_HIP(hipMemcpyAsync(d_poff.p, hpoff, (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
  launch_segments(d_poff.as<int64_t>(), n, total_pts, d_cloud_of_sorted.as<int>(), s);
  // tree order, then the bounding-volume tree over it
  if (!with_tree) {
    prof.end(span, s, spans);
    LCR_HIP(hipEventRecord(ev_build, s));  // the pinned regions are reused by the next build
    return;
  }
  d_tpts.reserve(npts * sizeof(float4));
  d_tpos.reserve(npts * sizeof(int));
  d_tkeys.reserve(npts * sizeof(uint64_t));
  d_tkeys_sorted.reserve(npts * sizeof(uint64_t));
  d_tvals.reserve(npts * sizeof(int));
  if (ttmp_pts != total_pts || ttmp_n != n) {  // the size query is slow: once per shape
    ttmp_bytes = tree_order_tmp_bytes(total_pts, n);
    ttmp_pts = total_pts;
    ttmp_n = n;
  }
  const size_t tmp_bytes = ttmp_bytes;
  d_ttmp.reserve(std::max<size_t>(tmp_bytes, 16));
  LCR_HIP(launch_tree_order(d_grids.as<GridDesc>(), n, d_pts.as<float4>(), d_cloud_of_sorted.as<int>(),
                            total_pts, d_tkeys.as<uint64_t>(), d_tkeys_sorted.as<uint64_t>(),
                            d_tvals.as<int>(), d_tpos.as<int>(), d_ttmp.p, tmp_bytes,
                            d_tpts.as<float4>(), s));
  d_nodes.reserve((size_t)std::max<int64_t>(total_nodes, 1) * 2 * sizeof(float4));
  launch_bvh_build(d_grids.as<GridDesc>(), n, max_depth, d_tpts.as<float4>(), d_nodes.as<float4>(), s);
  prof.end(span, s, spans);
  LCR_HIP(hipEventRecord(ev_build, s));  // the pinned regions are reused by the next build
}

void CloudSet::compute_normals(hipStream_t s, int k, const float* vp_host, int flip_all,
                               std::vector<Profiler::Span>* spans, bool normals_stage) {
  Profiler& prof = Profiler::get();
  // the previous pass's copy out of h_vp must be done before it is rewritten
  if (ev_vp)
    LCR_HIP(hipEventSynchronize(ev_vp));
  else
    LCR_HIP(hipEventCreateWithFlags(&ev_vp, hipEventDisableTiming));
  d_vp.reserve((size_t)n * 3 * sizeof(float));
  h_vp.reserve((size_t)n * 3 * sizeof(float));
  float* v = h_vp.as<float>();
  for (int c = 0; c < n; ++c)
    for (int a = 0; a < 3; ++a) v[3 * c + a] = vp_host ? vp_host[3 * c + a] : 0.0f;
  LCR_HIP(hipMemcpyAsync(d_vp.p, v, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice, s));
  LCR_HIP(hipEventRecord(ev_vp, s));
  // the selection reads the sorted points through a raw buffer of 32-bit byte offsets
  if (total_pts > ((int64_t)1 << 27) - 1) throw Error{-5, "more than 2^27 points in one normals batch"};
  const int span = prof.begin("normals", s, spans);
  const size_t np = (size_t)std::max<int64_t>(total_pts, 1);
  d_nbr.reserve(np * (size_t)k * sizeof(int));
  d_nsel.reserve(np * sizeof(int));
  d_nmom.reserve(np * 9 * sizeof(double));
  d_nwork.reserve((size_t)normals_work_ints((int64_t)np) * sizeof(int));
  DevBuf& pts = normals_stage ? d_npts : d_pts;
  DevBuf& nrm = normals_stage ? d_nnrm : d_nrm;
  nrm.reserve(np * sizeof(float4));
  launch_normals(d_grids.as<GridDesc>(), d_cell_start.as<int>(), pts.as<float4>(),
                 d_cloud_of_sorted.as<int>(), total_pts, k, d_vp.as<float>(), flip_all,
                 d_raw.as<float4>(), d_cloud_off.as<int64_t>(), d_nbr.as<int>(), d_nmom.as<double>(),
                 d_nsel.as<int>(), d_nwork.as<int>(), nrm.as<float4>(), s);
  prof.end(span, s, spans);
  has_nrm = true;
}

void CloudSet::build_near_mask(hipStream_t s, double r) {
  // dilation in cells: a point within r of a query lies at most ceil(r / h) + 1
  // cells away along each axis (the +1 covers positions inside the cells and
  // binning rounding)
  int kd_max = 0;
  for (const GridDesc& g : grids)
    if (g.npts > 0) kd_max = std::max(kd_max, (int)ceil(r * 1.0001 * (double)g.inv_h) + 1);
  const bool use = std::isfinite(r) && r > 0.0 && kd_max <= kMaxNearCells && total_cells > 0;
  // the same word offsets launch_set_near gives the device copy
  int64_t words = 0;
  for (GridDesc& g : grids) {
    g.near_r = use ? (float)r : 0.0f;
    g.near_off = words;
    words += (int64_t)near_row_words(g.nx) * g.ny * g.nz;
  }
  launch_set_near(d_grids.as<GridDesc>(), n, use ? (float)r : 0.0f, s);
  if (use) {
    d_near.reserve((size_t)std::max<int64_t>(words, 1) * sizeof(uint32_t));
    d_near_tmp.reserve((size_t)std::max<int64_t>(words, 1) * sizeof(uint32_t));
    launch_near_mask(d_grids.as<GridDesc>(), n, d_cell_start.as<int>(), words, (float)r,
                     d_near.as<uint32_t>(), d_near_tmp.as<uint32_t>(), s);
  }
}

void CloudSet::normals_to_raw(hipStream_t s) {
  d_raw_nrm.reserve((size_t)std::max<int64_t>(total_raw, 1) * sizeof(float4));
  launch_unsort_normals(d_pts.as<float4>(), d_nrm.as<float4>(), d_cloud_of_sorted.as<int>(),
                        d_cloud_off.as<int64_t>(), total_pts, d_raw_nrm.as<float4>(), total_raw, s);
  has_nrm = true;
}

// ------------------------------------------------------------------ engine

Engine::~Engine() { reset_graph(); }

// Never throws: also runs from the destructor.
void Engine::reset_graph() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (tail_exec) (void)hipGraphExecDestroy(tail_exec);
  if (tail_graph) (void)hipGraphDestroy(tail_graph);
  if (split_exec) (void)hipGraphExecDestroy(split_exec);
  if (split_graph) (void)hipGraphDestroy(split_graph);
  exec = nullptr;
  graph = nullptr;
  tail_exec = nullptr;
  tail_graph = nullptr;
  split_exec = nullptr;
  split_graph = nullptr;
}

// Live 64-point tiles at or below which a chunk runs the four-lanes-per-query
// search; LCR_SPLIT_TILES (read at every setup) overrides it, 0 turns the split off.
static int icp_split_units() {
  const char* e = getenv("LCR_SPLIT_TILES");
  if (e && *e) return std::max(0, atoi(e));
  return 2048;
}

void Engine::setup(const CloudSet& cs, const std::vector<int>& src, const std::vector<int>& tgt,
                   const std::vector<double>& r2, int tile_pts) {
  reset_graph();
  tile = (tile_pts == 64 || tile_pts == 128 || tile_pts == 256) ? tile_pts : 256;
  npairs = (int)src.size();
  hp.assign(npairs, PairDesc{});
  std::vector<int> tp, gt;
  ntiles = 0;
  for (int p = 0; p < npairs; ++p) {
    PairDesc& d = hp[p];
    d.src = src[p];
    d.tgt = tgt[p];
    d.tile_begin = ntiles;
    d.ntiles = (cs.grids[d.src].npts + tile - 1) / tile;
    d.max_corr_d2 = r2[p];
    for (int t = 0; t < d.ntiles; ++t) tp.push_back(p);
    for (int t = 0; t < d.ntiles; t += gate_stride(d.ntiles)) gt.push_back(ntiles + t);
    ntiles += d.ntiles;
  }
  d_pairs.reserve(std::max(npairs, 1) * sizeof(PairDesc));
  d_tile_pair.reserve(std::max(ntiles, 1) * sizeof(int));
  d_state.reserve((size_t)(npairs + 1) * sizeof(PairState));
  d_partials.reserve((size_t)std::max(ntiles, 1) * kMaxVals * sizeof(double));
  d_prev.reserve((size_t)std::max(ntiles, 1) * tile * sizeof(int));
  d_guess.reserve((size_t)std::max(npairs, 1) * 16 * sizeof(double));
  h_state.reserve((size_t)(npairs + 1) * sizeof(PairState));
  h_guess.reserve((size_t)std::max(npairs, 1) * 16 * sizeof(double));
  if (npairs > 0)
    LCR_HIP(hipMemcpy(d_pairs.p, hp.data(), npairs * sizeof(PairDesc), hipMemcpyHostToDevice));
  if (ntiles > 0)
    LCR_HIP(hipMemcpy(d_tile_pair.p, tp.data(), ntiles * sizeof(int), hipMemcpyHostToDevice));
  // live-tile lists: every tile until a run compacts them
  d_it_list.reserve((size_t)std::max(ntiles, 1) * sizeof(int2));
  d_fit_list.reserve((size_t)std::max(ntiles, 1) * sizeof(int2));
  d_fit_order.reserve((size_t)std::max(ntiles, 1) * sizeof(int2));
  d_fit_flags.reserve((size_t)std::max(ntiles, 1) * sizeof(int));
  d_list_count.reserve(2 * sizeof(int));
  d_list_off.reserve((size_t)std::max(npairs, 1) * sizeof(int));
  if (!d_fetch.p) {
    d_fetch.reserve(kFetchCounters * sizeof(int));
    LCR_HIP(hipMemset(d_fetch.p, 0, kFetchCounters * sizeof(int)));
  }
  {
    std::vector<int2> all((size_t)ntiles);
    for (int t = 0; t < ntiles; ++t) all[t] = make_int2(t, tp[t]);
    const int counts[2] = {ntiles, ntiles};
    if (ntiles > 0) {
      LCR_HIP(hipMemcpy(d_it_list.p, all.data(), all.size() * sizeof(int2), hipMemcpyHostToDevice));
      LCR_HIP(hipMemcpy(d_fit_list.p, all.data(), all.size() * sizeof(int2), hipMemcpyHostToDevice));
    }
    LCR_HIP(hipMemcpy(d_list_count.p, counts, sizeof counts, hipMemcpyHostToDevice));
  }
  ngate_tiles = (int)gt.size();
  d_gate_tiles.reserve(std::max<size_t>(gt.size(), 1) * sizeof(int));
  if (ngate_tiles > 0)
    LCR_HIP(hipMemcpy(d_gate_tiles.p, gt.data(), gt.size() * sizeof(int), hipMemcpyHostToDevice));
  const int cus = device_cu_count();
  setup_precision = c.precision;
  for (int ring = 0; ring < 2; ++ring) {
    const int per_cu = icp_max_blocks_per_cu(tile, c.point_to_plane != 0, c.precision, ring != 0);
    int nb = std::min(std::max(ntiles, 1), cus * per_cu);
    nb = std::max(8, (nb + 7) / 8 * 8);
    (ring ? nblocks_ring : nblocks) = nb;
  }
  int nf = std::min(std::max(ntiles, 1), cus * fitness_max_blocks_per_cu(tile));
  nblocks_fit = std::max(8, (nf + 7) / 8 * 8);
  split_units = icp_split_units();
  if (split_units > 0) {
    int ns = std::min(std::max(std::min(ntiles, split_units / (tile / 64)), 1),
                      cus * icp_split_blocks_per_cu(tile, c.point_to_plane != 0));
    nblocks_split = std::max(8, (ns + 7) / 8 * 8);
  }
}

// LCR_FIT_ORDER=fixed: the fitness passes in fixed per-block shares of the tiles
// instead of tiles handed out by counters (read at every run).
static bool fitness_fixed_shares() {
  const char* e = getenv("LCR_FIT_ORDER");
  return e && strcmp(e, "fixed") == 0;
}

IcpBuffers Engine::buffers(const CloudSet& cs) const {
  IcpBuffers b;
  memset(&b, 0, sizeof b);  // padding takes part in the graph key comparison
  b.grids = cs.d_grids.as<GridDesc>();
  b.cell_start = cs.d_cell_start.as<int>();
  b.pts = cs.d_pts.as<float4>();
  b.nrm = cs.has_nrm ? cs.d_nrm.as<float4>() : nullptr;
  b.nodes = cs.d_nodes.as<float4>();
  b.tpts = cs.d_tpts.as<float4>();
  b.tpos = cs.d_tpos.as<int>();
  b.pairs = d_pairs.as<PairDesc>();
  b.tile_pair = d_tile_pair.as<int>();
  b.state = d_state.as<PairState>();
  b.partials = d_partials.as<double>();
  b.prev = d_prev.as<int>();
  b.near = cs.d_near.p ? cs.d_near.as<uint32_t>() : nullptr;
  b.gate_tiles = d_gate_tiles.as<int>();
  b.ngate_tiles = ngate_tiles;
  b.it_list = d_it_list.as<int2>();
  b.it_count = d_list_count.as<int>();
  b.fit_list = d_fit_list.as<int2>();
  b.fit_count = d_list_count.as<int>() + 1;
  b.fetch = fitness_fixed_shares() ? nullptr : d_fetch.as<int>();
  b.fit_order = b.fetch ? d_fit_order.as<int2>() : nullptr;
  b.fit_flags = d_fit_flags.as<int>();
  b.ntiles = ntiles;
  b.npairs = npairs;
  b.tile = tile;
  b.morton = (c.morton && cs.d_tpts.p) ? 1 : 0;
  return b;
}

namespace {
// out[0]: instances still active; out[1]: live tiles (the last compaction's count).
__global__ void k_count_active(const PairState* __restrict__ st, int n,
                               const int* __restrict__ live_tiles, int* __restrict__ out) {
  __shared__ int s[256];
  int a = 0;
  for (int i = threadIdx.x; i < n; i += 256) a += st[i].active ? 1 : 0;
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s[0];
    out[1] = *live_tiles;
  }
}
}  // namespace

// The run is: state init -> chunks of kChunk iterations -> fitness.  Each chunk
// ends with a count of still-active instances copied to pinned memory; the host
// stops launching chunks once it reads zero, so converged cycles do not replay
// empty iterations.  With the profiler off each chunk (and the tail) replays a
// captured hipGraph; HIP events recorded inside a captured graph cannot be timed
// afterwards, so with the profiler on the same launches are issued eagerly with
// an event pair around each kernel.
bool Engine::ring_only(const CloudSet& cs) const {
  if (c.precision != 0 || npairs == 0) return false;
  for (const PairDesc& d : hp) {
    // the same float arithmetic as nn_query's per-query test
    const float r = (float)sqrt(d.max_corr_d2);
    const float rlim = r * 1.00001f;
    if (!std::isfinite(r) || !(rlim * cs.grids[d.tgt].inv_h <= kRingOnlyCells)) return false;
  }
  return true;
}

void Engine::run(hipStream_t s, const CloudSet& cs, const double* guesses, bool use_graph,
                 const std::function<void()>& pre) {
  if (npairs == 0) return;
  Profiler& prof = Profiler::get();
  c.ring_only = ring_only(cs) ? 1 : 0;
  const int nb_search = c.ring_only ? nblocks_ring : nblocks;
  const IcpBuffers b = buffers(cs);
  if (guesses) {
    memcpy(h_guess.p, guesses, (size_t)npairs * 16 * sizeof(double));
  } else {
    double* g = h_guess.as<double>();
    for (int p = 0; p < npairs; ++p)
      for (int i = 0; i < 16; ++i) g[16 * p + i] = (i % 5 == 0) ? 1.0 : 0.0;
  }
  LCR_HIP(hipMemcpyAsync(d_guess.p, h_guess.p, (size_t)npairs * 16 * sizeof(double),
                         hipMemcpyHostToDevice, s));
  if (pre) pre();
  launch_state_init(d_state.as<PairState>(), d_guess.as<double>(), d_pairs.as<PairDesc>(),
                    cs.d_grids.as<GridDesc>(), npairs, s);
  launch_fill_i32(d_prev.as<int>(), (int64_t)ntiles * tile, -1, s);  // no previous matches
  launch_compact_tiles(b, d_list_off.as<int>(), d_it_list.as<int2>(), d_list_count.as<int>(),
                       d_fit_list.as<int2>(), d_list_count.as<int>() + 1, true, s);
  d_active.reserve(2 * sizeof(int));
  h_active.reserve(2 * sizeof(int));
  const bool split_ok = use_split();
  const bool eager = !use_graph || prof.on();
  std::vector<Profiler::Span> spans;
  if (gate_h > 1 && ntiles > 0) {
    const int k = prof.begin("gate", s, &spans);
    // XCD-contiguous: the hypotheses of a pair share its target grid in one L2
    // (74 pairs x 15 hypotheses: 21.4 against 21.6 ms per cycle step in grid-stride
    // order)
    launch_fitness(b, 0.0, nblocks_fit, s, true, 0);
    launch_gate(d_state.as<PairState>(), npairs / gate_h, gate_h, s);
    launch_compact_tiles(b, d_list_off.as<int>(), d_it_list.as<int2>(), d_list_count.as<int>(),
                         d_fit_list.as<int2>(), d_list_count.as<int>() + 1, true, s);
    prof.end(k, s, &spans);
  }
  auto chunk = [&](std::vector<Profiler::Span>* sp, bool split) {
    for (int it = 0; it < kChunk; ++it) {
      // the split launches are also counted on their own
      const int k = prof.begin("icp_iteration", s, sp);
      const int ks = split ? prof.begin("icp_iteration_split", s, sp) : -1;
      if (split)
        launch_icp_search_split(b, c, nblocks_split, s);
      else
        launch_icp_search(b, c, nb_search, s);
      prof.end(ks, s, sp);
      prof.end(k, s, sp);
      const int k2 = prof.begin("icp_solve", s, sp);
      launch_icp_solve(b, c, s);
      prof.end(k2, s, sp);
    }
    launch_compact_tiles(b, d_list_off.as<int>(), d_it_list.as<int2>(), d_list_count.as<int>(),
                         nullptr, nullptr, false, s);
    hipLaunchKernelGGL(k_count_active, dim3(1), dim3(256), 0, s, d_state.as<PairState>(), npairs,
                       d_list_count.as<int>(), d_active.as<int>());
    LCR_HIP(hipMemcpyAsync(h_active.p, d_active.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  };
  auto tail = [&](std::vector<Profiler::Span>* sp) {
    const int k = prof.begin("fitness", s, sp);
    // one launch over every instance (with fixed shares, XCD-contiguous tiles keep
    // each target's tree in one L2)
    launch_fitness(b, DBL_MAX, nblocks_fit, s, false, 0);
    prof.end(k, s, sp);
    LCR_HIP(hipMemcpyAsync(h_state.p, d_state.p, (size_t)npairs * sizeof(PairState),
                           hipMemcpyDeviceToHost, s));
  };
  if (!eager) {
    const GraphKey key = graph_key(b);
    if (!exec || !tail_exec || (split_ok && !split_exec) || !(key == graph_key_)) {
      reset_graph();
      capture(s, [&] { chunk(nullptr, false); }, &graph, &exec);
      if (split_ok) capture(s, [&] { chunk(nullptr, true); }, &split_graph, &split_exec);
      capture(s, [&] { tail(nullptr); }, &tail_graph, &tail_exec);
      graph_key_ = key;
    }
  }
  // the first chunk splits when all of the tiles fit under split_units, later ones
  // when the previous chunk left that few live
  bool split = split_fits(ntiles);
  const int max_chunks = ntiles > 0 ? (c.max_iterations + kChunk - 1) / kChunk : 0;
  for (int k = 0; k < max_chunks; ++k) {
    if (eager)
      chunk(&spans, split);
    else
      LCR_HIP(hipGraphLaunch(split ? split_exec : exec, s));
    LCR_HIP(hipStreamSynchronize(s));
    if (h_active.as<int>()[0] == 0) break;
    split = split_fits(h_active.as<int>()[1]);
  }
  if (ntiles > 0) {
    if (eager)
      tail(&spans);
    else
      LCR_HIP(hipGraphLaunch(tail_exec, s));
  } else {
    LCR_HIP(hipMemcpyAsync(h_state.p, d_state.p, (size_t)npairs * sizeof(PairState),
                           hipMemcpyDeviceToHost, s));
  }
  LCR_HIP(hipStreamSynchronize(s));
  if (!spans.empty()) prof.collect(spans, false);
}

void Engine::capture(hipStream_t s, const std::function<void()>& f, hipGraph_t* g,
                     hipGraphExec_t* e) {
  LCR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    f();
  } catch (...) {
    hipGraph_t t = nullptr;
    (void)hipStreamEndCapture(s, &t);
    if (t) (void)hipGraphDestroy(t);
    throw;
  }
  LCR_HIP(hipStreamEndCapture(s, g));
  LCR_HIP(hipGraphInstantiate(e, *g, nullptr, nullptr, 0));
}

// Everything a captured graph bakes in: kernel arguments and buffer addresses.
Engine::GraphKey Engine::graph_key(const IcpBuffers& b) const {
  GraphKey k;
  memset(&k, 0, sizeof k);
  k.b = b;
  k.c = c;
  k.nblocks = c.ring_only ? nblocks_ring : nblocks;
  k.nblocks_fit = nblocks_fit;
  k.nblocks_split = use_split() ? nblocks_split : 0;
  k.guess = d_guess.p;
  k.hstate = h_state.p;
  k.active = d_active.p;
  k.hactive = h_active.p;
  return k;
}

void Engine::run_fitness(hipStream_t s, const CloudSet& cs, double max_range) {
  if (npairs == 0) return;
  Profiler& prof = Profiler::get();
  std::vector<Profiler::Span> spans;
  const IcpBuffers b = buffers(cs);
  if (ntiles > 0) {
    const int sp = prof.begin("fitness", s, &spans);
    launch_fitness(b, max_range, nblocks_fit, s, false, c.interleave);
    prof.end(sp, s, &spans);
  }
  LCR_HIP(hipMemcpyAsync(h_state.p, d_state.p, (size_t)npairs * sizeof(PairState),
                         hipMemcpyDeviceToHost, s));
  LCR_HIP(hipStreamSynchronize(s));
  if (!spans.empty()) prof.collect(spans, false);
}

}  // namespace lcr
