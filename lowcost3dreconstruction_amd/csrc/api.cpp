// api.cpp -- the C ABI of liblcr (include/lcr.h).  Every entry point catches
// exceptions and turns them into negative return codes plus lcr_last_error().
#include <float.h>
#include <math.h>
#include <string.h>

#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lcr.h"
#include "runtime.hpp"
#include "tune_cache.hpp"

using namespace lcr;

namespace {
thread_local std::string g_err;
}  // namespace

int lcr::set_last_error(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

namespace {

thread_local int g_device = 0;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    return fail(e.code, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(LCR_E_HIP, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(LCR_E_ARG, e.what());
  }
}

bool params_ok(const lcr_icp_params* p, std::string* why) {
  if (!p) {
    *why = "params is NULL";
    return false;
  }
  if (!(p->max_corr_dist > 0.0) || isnan(p->max_corr_dist)) {
    *why = "max_corr_dist must be > 0";
    return false;
  }
  if (p->max_iterations <= 0) {
    *why = "max_iterations needs to be greater than zero";
    return false;
  }
  if (p->max_iterations > 100000) {
    *why = "max_iterations above 100000";
    return false;
  }
  if (p->precision_mode != LCR_PREC_FP32 && p->precision_mode != LCR_PREC_FP16 &&
      p->precision_mode != LCR_PREC_BF16 && p->precision_mode != LCR_PREC_FP32_TILES) {
    *why = "unsupported precision_mode";
    return false;
  }
  if (p->precision_mode != LCR_PREC_FP32 && p->tile != 0 && p->tile != 64) {
    *why = "distance tiles use tile 64";
    return false;
  }
  if (p->tile != 0 && p->tile != 64 && p->tile != 128 && p->tile != 256) {
    *why = "tile must be 0, 64, 128 or 256";
    return false;
  }
  return true;
}

// Tile order of the ICP search launches (read once): by default XCD-contiguous
// ranges while at least half of the tiles are live, grid stride after (2; on
// cycle36 2.34 ms of search per step against 2.37 ms in grid stride throughout and
// 2.53 ms XCD-contiguous throughout); LCR_TILE_ORDER "stride" (1) or "xcd" (0)
// fixes one order.
int tile_order_interleave() {
  static const int v = [] {
    const char* e = getenv("LCR_TILE_ORDER");
    if (e && strcmp(e, "xcd") == 0) return 0;
    if (e && strcmp(e, "stride") == 0) return 1;
    return 2;
  }();
  return v;
}

// The ICP iterations take each tile's points in the source's tree (Morton)
// order (read once); LCR_ICP_ORDER=cell takes them in cell order.
bool icp_point_order_morton() {
  static const bool v = [] {
    const char* e = getenv("LCR_ICP_ORDER");
    return !(e && strcmp(e, "cell") == 0);
  }();
  return v;
}

IcpConst make_const(const lcr_icp_params& p) {
  IcpConst c{};
  c.interleave = tile_order_interleave();
  c.morton = icp_point_order_morton() ? 1 : 0;
  c.max_iterations = p.max_iterations;
  c.point_to_plane = p.point_to_plane ? 1 : 0;
  c.rotation_threshold = 1.0 - p.transformation_epsilon;
  c.translation_threshold = p.transformation_epsilon;
  c.mse_rel = p.euclidean_fitness_epsilon;
  c.mse_abs = 1e-12;
  c.precision = p.precision_mode;
  return c;
}

struct Stream {
  hipStream_t s = nullptr;
  int device = 0;
  Stream() {
    device = g_device;
    LCR_HIP(hipSetDevice(device));
    LCR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  }
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
  void bind() const { LCR_HIP(hipSetDevice(device)); }
};

struct HostCloud {
  std::vector<float> xyz, nrm;  // packed
  int64_t n = 0;
  bool has_nrm = false;
  void set(const float* p, const float* q, int64_t count, int64_t stride) {
    n = count;
    xyz.resize((size_t)std::max<int64_t>(count, 0) * 3);
    has_nrm = q != nullptr;
    nrm.resize(has_nrm ? xyz.size() : 0);
    const char* pb = reinterpret_cast<const char*>(p);
    const char* qb = reinterpret_cast<const char*>(q);
    for (int64_t i = 0; i < count; ++i) {
      const float* a = reinterpret_cast<const float*>(pb + i * stride);
      xyz[3 * i] = a[0];
      xyz[3 * i + 1] = a[1];
      xyz[3 * i + 2] = a[2];
      if (has_nrm) {
        const float* b = reinterpret_cast<const float*>(qb + i * stride);
        nrm[3 * i] = b[0];
        nrm[3 * i + 1] = b[1];
        nrm[3 * i + 2] = b[2];
      }
    }
  }
};

constexpr int64_t kMaxCloudPoints = (int64_t)1 << 24;

}  // namespace

struct lcr_icp {
  lcr_icp_params params;
  Stream stream;
  HostCloud src, tgt;
  bool have_src = false, have_tgt = false, dirty = true;
  CloudSet cs;  // cloud 0 = source, cloud 1 = target
  Engine eng;
  bool aligned = false;
  PairState result{};
  // scratch for correspondences
  DevBuf d_idx, d_d2;
  HostBuf h_idx;
};

struct lcr_cycle {
  lcr_cycle_desc desc;
  Stream stream;
  std::vector<int32_t> src, tgt;
  std::vector<double> guesses;
  std::vector<float> vps;
  CloudSet cs;
  Engine eng;
  bool ran = false;
  bool need_normals = false;
  int nhyp_total = 1;
  DevBuf d_pair_src, d_pair_tgt;
  std::vector<PairState> states;
};

extern "C" {

void lcr_icp_params_default(lcr_icp_params* p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  p->max_corr_dist = sqrt(DBL_MAX);
  p->max_iterations = 10;
  p->transformation_epsilon = 0.0;
  p->euclidean_fitness_epsilon = -DBL_MAX;
  p->point_to_plane = 0;
  p->precision_mode = LCR_PREC_FP32;
  p->cell_factor = 0.0;
  p->tile = 0;
}

int lcr_init(int device) {
  return guarded([&] {
    int n = 0;
    LCR_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(LCR_E_ARG, "no such HIP device");
    LCR_HIP(hipSetDevice(device));
    g_device = device;
    noalloc_init(device);
    return LCR_OK;
  });
}

int lcr_icp_create(const lcr_icp_params* p, lcr_icp** out) {
  return guarded([&] {
    if (!out) return fail(LCR_E_ARG, "out is NULL");
    *out = nullptr;
    std::string why;
    if (!params_ok(p, &why)) return fail(LCR_E_ARG, why);
    std::unique_ptr<lcr_icp> h(new lcr_icp());
    h->params = *p;
    *out = h.release();
    return LCR_OK;
  });
}

void lcr_icp_destroy(lcr_icp* h) {
  if (!h) return;
  try {
    h->stream.bind();
  } catch (...) {
  }
  delete h;
}

int lcr_icp_set_params(lcr_icp* h, const lcr_icp_params* p) {
  return guarded([&] {
    if (!h) return fail(LCR_E_ARG, "handle is NULL");
    std::string why;
    if (!params_ok(p, &why)) return fail(LCR_E_ARG, why);
    const bool regrid = p->cell_factor != h->params.cell_factor || p->tile != h->params.tile;
    h->params = *p;
    if (regrid) h->dirty = true;
    return LCR_OK;
  });
}

static int set_cloud(lcr_icp* h, HostCloud& hc, bool& have, const float* xyz, const float* nrm,
                     int64_t n, int64_t stride) {
  if (!h) return fail(LCR_E_ARG, "handle is NULL");
  if (n < 0 || (n > 0 && !xyz)) return fail(LCR_E_ARG, "invalid cloud");
  if (n > kMaxCloudPoints) return fail(LCR_E_LIMIT, "cloud larger than 2^24 points");
  if (stride == 0) stride = 3 * sizeof(float);
  if (stride < (int64_t)(3 * sizeof(float)) || stride % 4 != 0)
    return fail(LCR_E_ARG, "stride_bytes must be a multiple of 4 and >= 12");
  hc.set(xyz, nrm, n, stride);
  have = true;
  h->dirty = true;
  h->aligned = false;
  return LCR_OK;
}

int lcr_icp_set_source(lcr_icp* h, const float* xyz, const float* nrm, int64_t n,
                       int64_t stride_bytes) {
  return guarded([&] { return set_cloud(h, h->src, h->have_src, xyz, nrm, n, stride_bytes); });
}

int lcr_icp_set_target(lcr_icp* h, const float* xyz, const float* nrm, int64_t n,
                       int64_t stride_bytes) {
  return guarded([&] { return set_cloud(h, h->tgt, h->have_tgt, xyz, nrm, n, stride_bytes); });
}

static void prepare(lcr_icp* h) {
  h->stream.bind();
  h->eng.c = make_const(h->params);
  if (h->dirty) {
    std::vector<const float*> xyz = {h->src.xyz.data(), h->tgt.xyz.data()};
    std::vector<const float*> nrm = {h->src.has_nrm ? h->src.nrm.data() : nullptr,
                                     h->tgt.has_nrm ? h->tgt.nrm.data() : nullptr};
    // normals travel with the target (point-to-plane); the source's are only
    // transformed for the output, from the host copy
    const bool tn = h->tgt.has_nrm;
    std::vector<const float*> nrm_up = {tn ? (h->src.has_nrm ? h->src.nrm.data() : nullptr)
                                           : nullptr,
                                        tn ? h->tgt.nrm.data() : nullptr};
    (void)nrm;
    std::vector<int64_t> cnt = {h->src.n, h->tgt.n};
    // a cloud set has normals only when every cloud has them; give the source
    // zeros if only the target has normals
    std::vector<float> zeros;
    if (tn && !h->src.has_nrm) {
      zeros.assign(h->src.xyz.size(), 0.0f);
      nrm_up[0] = zeros.data();
    }
    h->cs.upload(h->stream.s, xyz, nrm_up, cnt, 3 * sizeof(float));
    h->cs.build(h->stream.s, h->params.cell_factor, nullptr);
    const double r = h->params.max_corr_dist;
    h->cs.build_near_mask(h->stream.s, r);
    h->eng.setup(h->cs, {0}, {1}, {r * r}, (h->params.tile && !h->params.precision_mode) ? h->params.tile : 64);
    h->dirty = false;
  } else {
    // parameters may have changed: refresh the instance radius
    const double r = h->params.max_corr_dist;
    if (h->eng.hp.size() == 1 && h->eng.hp[0].max_corr_d2 != r * r) {
      h->cs.build_near_mask(h->stream.s, r);
      h->eng.setup(h->cs, {0}, {1}, {r * r}, (h->params.tile && !h->params.precision_mode) ? h->params.tile : 64);
    }
  }
}

int lcr_icp_align(lcr_icp* h, const double* guess, float* out_xyz, float* out_nrm) {
  return guarded([&] {
    if (!h) return fail(LCR_E_ARG, "handle is NULL");
    if (!h->have_src || !h->have_tgt) return fail(LCR_E_STATE, "source and target must be set");
    if (h->params.point_to_plane && !h->tgt.has_nrm)
      return fail(LCR_E_NORMALS, "point-to-plane ICP needs target normals");
    prepare(h);
    h->eng.run(h->stream.s, h->cs, guess, true);
    h->result = h->eng.states()[0];
    h->aligned = true;
    if (out_xyz || out_nrm) {
      // K8: transform the raw source by the final transform
      DevBuf out;
      const int64_t n = h->src.n;
      out.reserve((size_t)std::max<int64_t>(n, 1) * 6 * sizeof(float));
      float* dx = out.as<float>();
      float* dn = dx + 3 * std::max<int64_t>(n, 1);
      DevBuf rawn;
      const float4* raw_nrm = nullptr;
      if (out_nrm && h->src.has_nrm) {
        rawn.reserve((size_t)std::max<int64_t>(n, 1) * sizeof(float4));
        std::vector<float4> tmp((size_t)n);
        for (int64_t i = 0; i < n; ++i)
          tmp[i] = make_float4(h->src.nrm[3 * i], h->src.nrm[3 * i + 1], h->src.nrm[3 * i + 2], 0.f);
        LCR_HIP(hipMemcpyAsync(rawn.p, tmp.data(), n * sizeof(float4), hipMemcpyHostToDevice,
                               h->stream.s));
        LCR_HIP(hipStreamSynchronize(h->stream.s));
        raw_nrm = rawn.as<float4>();
      }
      launch_transform_output(h->cs.d_raw.as<float4>(), raw_nrm, n, h->eng.d_state.as<PairState>(),
                              0, dx, raw_nrm ? dn : nullptr, h->stream.s);
      if (out_xyz && n > 0)
        LCR_HIP(hipMemcpyAsync(out_xyz, dx, n * 3 * sizeof(float), hipMemcpyDeviceToHost,
                               h->stream.s));
      if (out_nrm && n > 0) {
        if (raw_nrm) {
          LCR_HIP(hipMemcpyAsync(out_nrm, dn, n * 3 * sizeof(float), hipMemcpyDeviceToHost,
                                 h->stream.s));
        } else {
          for (int64_t i = 0; i < 3 * n; ++i) out_nrm[i] = NAN;
        }
      }
      LCR_HIP(hipStreamSynchronize(h->stream.s));
    }
    return LCR_OK;
  });
}

int lcr_icp_final_transform(const lcr_icp* h, double T[16]) {
  if (!h || !T) return fail(LCR_E_ARG, "NULL argument");
  if (!h->aligned) {
    for (int i = 0; i < 16; ++i) T[i] = (i % 5 == 0) ? 1.0 : 0.0;
    return LCR_OK;
  }
  memcpy(T, h->result.T, 16 * sizeof(double));
  return LCR_OK;
}

int lcr_icp_has_converged(const lcr_icp* h) {
  if (!h) return fail(LCR_E_ARG, "handle is NULL");
  return h->aligned && h->result.converged ? 1 : 0;
}

int lcr_icp_iterations(const lcr_icp* h) {
  if (!h) return fail(LCR_E_ARG, "handle is NULL");
  return h->aligned ? h->result.iterations : 0;
}

int lcr_icp_convergence_state(const lcr_icp* h) {
  if (!h) return fail(LCR_E_ARG, "handle is NULL");
  return h->aligned ? h->result.state : LCR_CONV_NOT_CONVERGED;
}

int lcr_icp_fitness(lcr_icp* h, double max_range, double* out) {
  return guarded([&] {
    if (!h || !out) return fail(LCR_E_ARG, "NULL argument");
    if (!h->aligned) return fail(LCR_E_STATE, "align has not run");
    if (isnan(max_range)) return fail(LCR_E_ARG, "max_range is NaN");
    h->stream.bind();
    if (max_range == DBL_MAX || isinf(max_range)) {
      *out = h->result.fitness;  // computed by align with max_range = DBL_MAX
      return LCR_OK;
    }
    h->eng.run_fitness(h->stream.s, h->cs, max_range);
    *out = h->eng.states()[0].fitness;
    return LCR_OK;
  });
}

int lcr_icp_correspondences(lcr_icp* h, const double* T, int32_t* idx, float* d2) {
  return guarded([&] {
    if (!h || !idx || !d2) return fail(LCR_E_ARG, "NULL argument");
    if (!h->have_src || !h->have_tgt) return fail(LCR_E_STATE, "source and target must be set");
    prepare(h);
    const int64_t n = h->src.n;
    for (int64_t i = 0; i < n; ++i) {
      idx[i] = -1;
      d2[i] = INFINITY;
    }
    // probe slot: state[1] carries T (or the current final transform)
    PairState probe{};
    if (T)
      memcpy(probe.T, T, sizeof probe.T);
    else if (h->aligned)
      memcpy(probe.T, h->result.T, sizeof probe.T);
    else
      for (int i = 0; i < 16; ++i) probe.T[i] = (i % 5 == 0) ? 1.0 : 0.0;
    h->h_idx.reserve(sizeof(PairState));
    memcpy(h->h_idx.p, &probe, sizeof probe);
    PairState* slot = h->eng.d_state.as<PairState>() + 1;
    LCR_HIP(hipMemcpyAsync(slot, h->h_idx.p, sizeof probe, hipMemcpyHostToDevice, h->stream.s));
    IcpBuffers b = h->eng.buffers(h->cs);
    b.state = slot;
    b.prev = nullptr;
    h->d_idx.reserve((size_t)std::max<int64_t>(n, 1) * sizeof(int32_t));
    h->d_d2.reserve((size_t)std::max<int64_t>(n, 1) * sizeof(float));
    if (n > 0) {
      LCR_HIP(hipMemcpyAsync(h->d_idx.p, idx, n * sizeof(int32_t), hipMemcpyHostToDevice,
                             h->stream.s));
      LCR_HIP(hipMemcpyAsync(h->d_d2.p, d2, n * sizeof(float), hipMemcpyHostToDevice, h->stream.s));
      launch_correspondences(b, 0, h->cs.grids[0].npts, h->d_idx.as<int32_t>(),
                             h->d_d2.as<float>(), h->stream.s);
      LCR_HIP(hipMemcpyAsync(idx, h->d_idx.p, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                             h->stream.s));
      LCR_HIP(hipMemcpyAsync(d2, h->d_d2.p, n * sizeof(float), hipMemcpyDeviceToHost, h->stream.s));
    }
    LCR_HIP(hipStreamSynchronize(h->stream.s));
    return LCR_OK;
  });
}

int lcr_normals(const float* xyz, int64_t n, int64_t stride_bytes, int32_t k, const float vp[3],
                int32_t flip_all, float* nrm, float* curv) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!xyz || !nrm))) return fail(LCR_E_ARG, "invalid arguments");
    if (n > kMaxCloudPoints) return fail(LCR_E_LIMIT, "cloud larger than 2^24 points");
    if (k < 1) return fail(LCR_E_ARG, "k must be >= 1");
    if (stride_bytes == 0) stride_bytes = 3 * sizeof(float);
    if (stride_bytes < 12 || stride_bytes % 4 != 0) return fail(LCR_E_ARG, "bad stride_bytes");
    if (n == 0) return LCR_OK;
    Stream st;
    CloudSet cs;
    cs.upload(st.s, {xyz}, {nullptr}, {n}, stride_bytes);
    // cells sized for a k-neighbourhood: ~sqrt(k) point spacings.  0.70 sqrt(k)
    // (3.8 spacings at k = 30) keeps the 3x3x3 blocks of dense capture regions
    // inside LDS more often than 0.75 (Kinect-v2 cycle 20.2 -> 19.5 ms per step, no
    // change on the Kinect-v1 cycle); certification rarely needs a larger block.
    const double factor = std::max(2.5, 0.70 * sqrt((double)k));
    cs.build(st.s, factor, nullptr, false);
    const float v[3] = {vp ? vp[0] : 0.f, vp ? vp[1] : 0.f, vp ? vp[2] : 0.f};
    cs.compute_normals(st.s, k, v, flip_all, nullptr);
    cs.normals_to_raw(st.s);  // input order; NaN for non-finite points
    std::vector<float4> out((size_t)n);
    LCR_HIP(hipMemcpyAsync(out.data(), cs.d_raw_nrm.p, n * sizeof(float4), hipMemcpyDeviceToHost,
                           st.s));
    LCR_HIP(hipStreamSynchronize(st.s));
    for (int64_t i = 0; i < n; ++i) {
      nrm[3 * i] = out[i].x;
      nrm[3 * i + 1] = out[i].y;
      nrm[3 * i + 2] = out[i].z;
      if (curv) curv[i] = out[i].w;
    }
    return LCR_OK;
  });
}

// ------------------------------------------------------------------ cycles

int lcr_cycle_create(const lcr_cycle_desc* d, lcr_cycle** out) {
  return guarded([&] {
    if (!d || !out) return fail(LCR_E_ARG, "NULL argument");
    *out = nullptr;
    std::string why;
    if (!params_ok(&d->params, &why)) return fail(LCR_E_ARG, why);
    if (d->n_clouds <= 0 || !d->xyz || !d->n_points) return fail(LCR_E_ARG, "no clouds");
    if (d->n_pairs < 0 || (d->n_pairs > 0 && (!d->pair_src || !d->pair_tgt)))
      return fail(LCR_E_ARG, "invalid pairs");
    if (d->n_hyp < 1) return fail(LCR_E_ARG, "n_hyp must be >= 1");
    std::unique_ptr<lcr_cycle> c(new lcr_cycle());
    c->desc = *d;
    c->stream.bind();
    bool all_nrm = d->nrm != nullptr;
    for (int i = 0; i < d->n_clouds; ++i) {
      if (d->n_points[i] < 0 || (d->n_points[i] > 0 && !d->xyz[i]))
        return fail(LCR_E_ARG, "invalid cloud");
      if (d->n_points[i] > kMaxCloudPoints) return fail(LCR_E_LIMIT, "cloud larger than 2^24 points");
      if (all_nrm && !d->nrm[i]) all_nrm = false;
    }
    for (int p = 0; p < d->n_pairs; ++p) {
      if (d->pair_src[p] < 0 || d->pair_src[p] >= d->n_clouds || d->pair_tgt[p] < 0 ||
          d->pair_tgt[p] >= d->n_clouds)
        return fail(LCR_E_ARG, "pair refers to a missing cloud");
    }
    std::vector<const float*> xyz(d->xyz, d->xyz + d->n_clouds);
    std::vector<const float*> nrm(d->n_clouds, nullptr);
    if (all_nrm)
      for (int i = 0; i < d->n_clouds; ++i) nrm[i] = d->nrm[i];
    std::vector<int64_t> cnt(d->n_points, d->n_points + d->n_clouds);
    c->need_normals = d->params.point_to_plane && !all_nrm;
    c->cs.upload(c->stream.s, xyz, nrm, cnt, 3 * sizeof(float));
    c->src.assign(d->pair_src, d->pair_src + d->n_pairs);
    c->tgt.assign(d->pair_tgt, d->pair_tgt + d->n_pairs);
    if (d->proposal && d->proposal_points < 0) return fail(LCR_E_ARG, "proposal_points < 0");
    const int H = d->n_hyp + (d->proposal ? 1 : 0);
    c->nhyp_total = H;
    const int ninst = d->n_pairs * H;
    c->guesses.assign((size_t)ninst * 16, 0.0);
    for (int p = 0; p < d->n_pairs; ++p)
      for (int hy = 0; hy < H; ++hy)
        for (int k = 0; k < 16; ++k) {
          const bool given = d->guesses && hy < d->n_hyp;
          c->guesses[16 * ((size_t)p * H + hy) + k] =
              given ? d->guesses[16 * ((size_t)p * d->n_hyp + hy) + k] : ((k % 5 == 0) ? 1.0 : 0.0);
        }
    if (d->n_pairs > 0) {
      c->d_pair_src.reserve((size_t)d->n_pairs * sizeof(int));
      c->d_pair_tgt.reserve((size_t)d->n_pairs * sizeof(int));
      LCR_HIP(hipMemcpy(c->d_pair_src.p, d->pair_src, d->n_pairs * sizeof(int), hipMemcpyHostToDevice));
      LCR_HIP(hipMemcpy(c->d_pair_tgt.p, d->pair_tgt, d->n_pairs * sizeof(int), hipMemcpyHostToDevice));
    }
    c->vps.assign((size_t)d->n_clouds * 3, 0.0f);
    if (d->viewpoints)
      for (int i = 0; i < 3 * d->n_clouds; ++i) c->vps[i] = d->viewpoints[i];
    c->desc.xyz = nullptr;
    c->desc.nrm = nullptr;
    c->desc.n_points = nullptr;
    c->desc.pair_src = nullptr;
    c->desc.pair_tgt = nullptr;
    c->desc.guesses = nullptr;
    c->desc.viewpoints = nullptr;
    *out = c.release();
    return LCR_OK;
  });
}

int lcr_cycle_run(lcr_cycle* c) {
  return guarded([&] {
    if (!c) return fail(LCR_E_ARG, "handle is NULL");
    c->stream.bind();
    const lcr_cycle_desc& d = c->desc;
    Profiler& prof = Profiler::get();
    std::vector<Profiler::Span> spans;
    const int k = d.normals_k > 0 ? d.normals_k : 30;
    if (c->need_normals) {
      // normals on a grid with cells sized for a k-neighbourhood (~sqrt(k) point
      // spacings), then the ICP grid with cells sized for 1-NN
      // the ICP grid is sorted from the normals grid's order (spatially coherent
      // input; the normals stay in sorted order)
      c->cs.has_nrm = false;
      c->cs.build(c->stream.s, std::max(2.5, 0.70 * sqrt((double)k)), &spans, false,
                  CloudSet::kBuildNormals);
      c->cs.compute_normals(c->stream.s, k, c->vps.data(), 0, &spans, true);
      c->cs.build(c->stream.s, d.params.cell_factor, &spans, true, CloudSet::kBuildFromNormals);
    } else {
      c->cs.build(c->stream.s, d.params.cell_factor, &spans);
    }
    c->cs.build_near_mask(c->stream.s, d.params.max_corr_dist);
    if (d.params.point_to_plane && !c->cs.has_nrm)
      return fail(LCR_E_NORMALS, "point-to-plane ICP needs normals");
    const int H = c->nhyp_total;
    const int ninst = d.n_pairs * H;
    std::vector<int> src(ninst), tgt(ninst);
    std::vector<double> r2(ninst, d.params.max_corr_dist * d.params.max_corr_dist);
    for (int p = 0; p < d.n_pairs; ++p)
      for (int hy = 0; hy < H; ++hy) {
        src[p * H + hy] = c->src[p];
        tgt[p * H + hy] = c->tgt[p];
      }
    c->eng.c = make_const(d.params);
    const int tile = (d.params.tile && !d.params.precision_mode) ? d.params.tile : 64;
    bool same = c->eng.npairs == ninst && c->eng.tile == tile && c->eng.hp.size() == (size_t)ninst &&
                c->eng.setup_precision == c->eng.c.precision;
    if (same) {
      for (int i = 0; i < ninst && same; ++i)
        same = c->eng.hp[i].ntiles == (c->cs.grids[src[i]].npts + tile - 1) / tile;
    }
    if (!same) c->eng.setup(c->cs, src, tgt, r2, tile);
    if (!spans.empty()) prof.collect(spans, false);
    std::function<void()> pre;
    if (d.proposal && d.n_pairs > 0) {
      ProposalImpl* P = proposal_impl(d.proposal);
      const int npts = d.proposal_points > 0 ? d.proposal_points : 8192;
      pre = [&, P, npts]() {
        const int sp = prof.begin("proposal", c->stream.s, &spans);
        proposal_guesses(*P, c->stream.s, c->cs, c->d_pair_src.as<int>(), c->d_pair_tgt.as<int>(),
                         d.n_pairs, npts, c->eng.d_guess.as<double>(), H, d.n_hyp);
        prof.end(sp, c->stream.s, &spans);
      };
    }
    c->eng.gate_h = (d.hyp_gate && H > 1) ? H : 0;
    c->eng.run(c->stream.s, c->cs, c->guesses.data(), true, pre);
    if (!spans.empty()) prof.collect(spans, false);
    c->states.assign(c->eng.states(), c->eng.states() + ninst);
    c->ran = true;
    return LCR_OK;
  });
}

int lcr_cycle_instances(lcr_cycle* c, double* T, int32_t* iterations, int32_t* state,
                        double* fitness) {
  if (!c) return fail(LCR_E_ARG, "handle is NULL");
  if (!c->ran) return fail(LCR_E_STATE, "lcr_cycle_run has not run");
  for (size_t i = 0; i < c->states.size(); ++i) {
    const PairState& s = c->states[i];
    if (T) memcpy(T + 16 * i, s.T, 16 * sizeof(double));
    if (iterations) iterations[i] = s.iterations;
    if (state) state[i] = s.state;
    if (fitness) fitness[i] = s.fitness;
  }
  return LCR_OK;
}

int lcr_cycle_initial_poses(lcr_cycle* c, double* T) {
  return guarded([&] {
    if (!c || !T) return fail(LCR_E_ARG, "NULL argument");
    if (!c->ran) return fail(LCR_E_STATE, "lcr_cycle_run has not run");
    c->stream.bind();
    // d_guess still holds the last run's poses (the proposal wrote its slot there)
    const size_t n = c->states.size();
    if (n > 0) {
      LCR_HIP(hipMemcpyAsync(T, c->eng.d_guess.p, n * 16 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream.s));
      LCR_HIP(hipStreamSynchronize(c->stream.s));
    }
    return LCR_OK;
  });
}

int lcr_cycle_results(lcr_cycle* c, lcr_cycle_result* r) {
  if (!c || !r) return fail(LCR_E_ARG, "NULL argument");
  if (!c->ran) return fail(LCR_E_STATE, "lcr_cycle_run has not run");
  const int H = c->nhyp_total;
  for (int p = 0; p < c->desc.n_pairs; ++p) {
    int best = -1;
    for (int h = 0; h < H; ++h) {
      const PairState& s = c->states[p * H + h];
      if (s.gated) continue;
      if (best < 0 || s.fitness < c->states[p * H + best].fitness) best = h;
    }
    if (best < 0) best = 0;
    const PairState& s = c->states[p * H + best];
    if (r->T) memcpy(r->T + 16 * (size_t)p, s.T, 16 * sizeof(double));
    if (r->iterations) r->iterations[p] = s.iterations;
    if (r->converged) r->converged[p] = s.converged;
    if (r->state) r->state[p] = s.state;
    if (r->fitness) r->fitness[p] = s.fitness;
    if (r->best_hyp) r->best_hyp[p] = best;
  }
  return LCR_OK;
}

int lcr_transform_points(const float* xyz, const float* nrm, int64_t n, int64_t stride_bytes,
                         const double T[16], float* out_xyz, float* out_nrm) {
  return guarded([&] {
    if (n < 0 || !T || (n > 0 && (!xyz || !out_xyz))) return fail(LCR_E_ARG, "invalid arguments");
    if (n > kMaxCloudPoints) return fail(LCR_E_LIMIT, "cloud larger than 2^24 points");
    if (stride_bytes == 0) stride_bytes = 3 * sizeof(float);
    if (stride_bytes < 12 || stride_bytes % 4 != 0) return fail(LCR_E_ARG, "bad stride_bytes");
    if (n == 0) return LCR_OK;
    Stream st;
    CloudSet cs;
    cs.upload(st.s, {xyz}, {nrm}, {n}, stride_bytes);
    DevBuf d_out;
    d_out.reserve((size_t)n * 6 * sizeof(float));
    float* dx = d_out.as<float>();
    float* dn = dx + 3 * n;
    launch_transform_points(cs.d_raw.as<float4>(), (nrm && cs.has_nrm) ? cs.d_raw_nrm.as<float4>() : nullptr,
                            n, T, dx, (nrm && out_nrm) ? dn : nullptr, st.s);
    LCR_HIP(hipMemcpyAsync(out_xyz, dx, n * 3 * sizeof(float), hipMemcpyDeviceToHost, st.s));
    if (nrm && out_nrm)
      LCR_HIP(hipMemcpyAsync(out_nrm, dn, n * 3 * sizeof(float), hipMemcpyDeviceToHost, st.s));
    LCR_HIP(hipStreamSynchronize(st.s));
    return LCR_OK;
  });
}

int lcr_voxel_grid(const float* xyz, const float* nrm, const float* curv, const uint8_t* rgb,
                   int64_t n, float leaf, float* out_xyz, float* out_nrm, float* out_curv,
                   uint8_t* out_rgb, int64_t* out_n) {
  return guarded([&] {
    if (n < 0 || !out_n || (n > 0 && (!xyz || !out_xyz))) return fail(LCR_E_ARG, "invalid arguments");
    if (!(leaf > 0.0f) || !std::isfinite(leaf)) return fail(LCR_E_ARG, "leaf must be > 0");
    if (n > kMaxCloudPoints) return fail(LCR_E_LIMIT, "cloud larger than 2^24 points");
    *out_n = 0;
    if (n == 0) return LCR_OK;
    Stream st;
    CloudSet cs;
    cs.upload(st.s, {xyz}, {nrm}, {n}, 3 * sizeof(float));
    DevBuf part, box;
    part.reserve((size_t)kBBoxChunks * 7 * sizeof(float));
    box.reserve(7 * sizeof(float));
    launch_bbox(cs.d_raw.as<float4>(), cs.d_cloud_off.as<int64_t>(), 1, part.as<float>(),
                box.as<float>(), st.s);
    float b[7];
    LCR_HIP(hipMemcpyAsync(b, box.p, sizeof b, hipMemcpyDeviceToHost, st.s));
    LCR_HIP(hipStreamSynchronize(st.s));
    const int64_t nfin = (int64_t)b[6];
    if (nfin == 0) return LCR_OK;
    // voxel_grid.hpp: the lattice from the min/max over finite points, FP32
    const float inv = 1.0f / leaf;
    int64_t d[3];
    for (int a = 0; a < 3; ++a) d[a] = (int64_t)((b[3 + a] - b[a]) * inv) + 1;
    if (d[0] * d[1] * d[2] > (int64_t)INT32_MAX) {
      // PCL leaves the cloud unchanged when the lattice overflows an int
      memcpy(out_xyz, xyz, (size_t)n * 3 * sizeof(float));
      if (out_nrm && nrm) memcpy(out_nrm, nrm, (size_t)n * 3 * sizeof(float));
      if (out_curv && curv) memcpy(out_curv, curv, (size_t)n * sizeof(float));
      if (out_rgb && rgb) memcpy(out_rgb, rgb, (size_t)n * 3);
      *out_n = n;
      return LCR_OK;
    }
    VoxLattice L;
    int64_t nvox = 1;
    L.inv = inv;
    for (int a = 0; a < 3; ++a) {
      L.minb[a] = (int)floorf(b[a] * inv);
      const int maxb = (int)floorf(b[3 + a] * inv);
      L.div[a] = (int64_t)maxb - L.minb[a] + 1;
      nvox *= L.div[a];
    }
    if (nvox > ((int64_t)1 << 30)) return fail(LCR_E_LIMIT, "voxel lattice larger than 2^30");
    DevBuf skey, skey_sorted, idx, sorted_raw, counts, start, flags, outidx, scan_tmp, sort_tmp,
        dcurv, drgb, ox, on, oc, orgb;
    for (DevBuf* v : {&skey, &skey_sorted, &idx, &sorted_raw}) v->reserve((size_t)n * sizeof(int));
    for (DevBuf* v : {&counts, &start, &flags, &outidx}) v->reserve((size_t)nvox * sizeof(int));
    scan_tmp.reserve((size_t)scan_tmp_ints(nvox) * sizeof(int));
    const size_t sort_bytes = voxel_sort_tmp_bytes(n, nvox);
    if (sort_bytes == 0) return fail(LCR_E_HIP, "voxel sort: no temporary size");
    sort_tmp.reserve(sort_bytes);
    if (curv) {
      dcurv.reserve((size_t)n * sizeof(float));
      LCR_HIP(hipMemcpyAsync(dcurv.p, curv, (size_t)n * sizeof(float), hipMemcpyHostToDevice, st.s));
    }
    if (rgb) {
      drgb.reserve((size_t)n * 3);
      LCR_HIP(hipMemcpyAsync(drgb.p, rgb, (size_t)n * 3, hipMemcpyHostToDevice, st.s));
    }
    ox.reserve((size_t)nfin * 3 * sizeof(float));
    if (nrm && out_nrm) on.reserve((size_t)nfin * 3 * sizeof(float));
    if (curv && out_curv) oc.reserve((size_t)nfin * sizeof(float));
    if (rgb && out_rgb) orgb.reserve((size_t)nfin * 3);
    VoxScratch w{skey.as<uint32_t>(), skey_sorted.as<uint32_t>(), idx.as<int>(), sorted_raw.as<int>(),
                 counts.as<int>(), start.as<int>(), flags.as<int>(), outidx.as<int>(),
                 scan_tmp.as<int>(), sort_tmp.p, sort_bytes};
    LCR_HIP(launch_voxel_grid(cs.d_raw.as<float4>(),
                              (nrm && cs.has_nrm) ? cs.d_raw_nrm.as<float4>() : nullptr,
                              curv ? dcurv.as<float>() : nullptr, rgb ? drgb.as<uint8_t>() : nullptr,
                              n, L, nvox, w, ox.as<float>(), on.p ? on.as<float>() : nullptr,
                              oc.p ? oc.as<float>() : nullptr, orgb.p ? orgb.as<uint8_t>() : nullptr,
                              st.s));
    int last[2];
    LCR_HIP(hipMemcpyAsync(&last[0], outidx.as<int>() + nvox - 1, sizeof(int), hipMemcpyDeviceToHost,
                           st.s));
    LCR_HIP(hipMemcpyAsync(&last[1], counts.as<int>() + nvox - 1, sizeof(int), hipMemcpyDeviceToHost,
                           st.s));
    LCR_HIP(hipStreamSynchronize(st.s));
    const int64_t nocc = (int64_t)last[0] + (last[1] > 0 ? 1 : 0);
    LCR_HIP(hipMemcpyAsync(out_xyz, ox.p, (size_t)nocc * 3 * sizeof(float), hipMemcpyDeviceToHost,
                           st.s));
    if (on.p)
      LCR_HIP(hipMemcpyAsync(out_nrm, on.p, (size_t)nocc * 3 * sizeof(float), hipMemcpyDeviceToHost,
                             st.s));
    if (oc.p)
      LCR_HIP(hipMemcpyAsync(out_curv, oc.p, (size_t)nocc * sizeof(float), hipMemcpyDeviceToHost,
                             st.s));
    if (orgb.p)
      LCR_HIP(hipMemcpyAsync(out_rgb, orgb.p, (size_t)nocc * 3, hipMemcpyDeviceToHost, st.s));
    LCR_HIP(hipStreamSynchronize(st.s));
    *out_n = nocc;
    return LCR_OK;
  });
}

int lcr_outlier_removal(const float* xyz, int64_t n, int32_t mean_k, double stddev_mult,
                        uint8_t* keep, float* mean_dist, double* stats, int64_t* kept) {
  return guarded([&] {
    if (n < 0 || (n > 0 && (!xyz || !keep))) return fail(LCR_E_ARG, "invalid arguments");
    if (mean_k < 1) return fail(LCR_E_ARG, "mean_k must be >= 1");
    if (n > kMaxCloudPoints) return fail(LCR_E_LIMIT, "cloud larger than 2^24 points");
    if (kept) *kept = 0;
    if (n == 0) return LCR_OK;
    Stream st;
    CloudSet cs;
    cs.upload(st.s, {xyz}, {nullptr}, {n}, 3 * sizeof(float));
    cs.build(st.s, std::max(2.5, 0.70 * sqrt((double)mean_k + 1.0)), nullptr, false);
    DevBuf dist, fin, sums, nsel, defer, part, dstats, dkeep;
    dist.reserve((size_t)n * sizeof(float));
    fin.reserve((size_t)n);
    const size_t np = (size_t)std::max<int64_t>(cs.total_pts, 1);
    sums.reserve(np * sizeof(double));
    nsel.reserve(np * sizeof(int));
    defer.reserve((np + 1) * sizeof(int));
    part.reserve((size_t)sor_stat_blocks(n) * 3 * sizeof(double));
    dstats.reserve(3 * sizeof(double));
    dkeep.reserve((size_t)n);
    LCR_HIP(hipMemsetAsync(dist.p, 0, (size_t)n * sizeof(float), st.s));
    LCR_HIP(hipMemsetAsync(fin.p, 0, (size_t)n, st.s));
    launch_sor_distances(cs.d_grids.as<GridDesc>(), cs.d_cell_start.as<int>(), cs.d_pts.as<float4>(),
                         cs.d_cloud_of_sorted.as<int>(), cs.d_cloud_off.as<int64_t>(), cs.total_pts,
                         mean_k, sums.as<double>(), nsel.as<int>(), defer.as<int>(),
                         dist.as<float>(), fin.as<uint8_t>(), st.s);
    launch_sor_stats(dist.as<float>(), fin.as<uint8_t>(), n, stddev_mult, part.as<double>(),
                     dstats.as<double>(), dkeep.as<uint8_t>(), st.s);
    LCR_HIP(hipMemcpyAsync(keep, dkeep.p, (size_t)n, hipMemcpyDeviceToHost, st.s));
    if (mean_dist)
      LCR_HIP(hipMemcpyAsync(mean_dist, dist.p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost,
                             st.s));
    double hs[3];
    LCR_HIP(hipMemcpyAsync(hs, dstats.p, sizeof hs, hipMemcpyDeviceToHost, st.s));
    LCR_HIP(hipStreamSynchronize(st.s));
    if (stats) memcpy(stats, hs, sizeof hs);
    if (kept) {
      int64_t k = 0;
      for (int64_t i = 0; i < n; ++i) k += keep[i] ? 1 : 0;
      *kept = k;
    }
    return LCR_OK;
  });
}

int lcr_depth_to_cloud(const uint16_t* depth, int32_t width, int32_t height, const uint8_t* color,
                       const double intrinsics[4], int32_t depth_min, int32_t depth_max,
                       const int32_t crop[4], float* out_xyz, float* out_nrm, uint8_t* out_rgb,
                       int64_t* out_n) {
  return guarded([&] {
    if (width < 0 || height < 0 || !intrinsics || !out_n ||
        ((int64_t)width * height > 0 && (!depth || !out_xyz)))
      return fail(LCR_E_ARG, "invalid arguments");
    const int64_t np = (int64_t)width * height;
    if (np > ((int64_t)1 << 26)) return fail(LCR_E_LIMIT, "depth map larger than 2^26 pixels");
    *out_n = 0;
    if (np == 0) return LCR_OK;
    DepthParams P;
    P.fx = intrinsics[0];
    P.fy = intrinsics[1];
    P.cx = intrinsics[2];
    P.cy = intrinsics[3];
    P.depth_min = depth_min;
    P.depth_max = depth_max;
    P.top = crop ? crop[0] : 0;
    P.bottom = crop ? crop[1] : height - 1;
    P.left = crop ? crop[2] : 0;
    P.right = crop ? crop[3] : width - 1;
    Stream st;
    DevBuf dd, dc, flags, pos, tmp, ox, on, orgb;
    dd.reserve((size_t)np * sizeof(uint16_t));
    LCR_HIP(hipMemcpyAsync(dd.p, depth, (size_t)np * sizeof(uint16_t), hipMemcpyHostToDevice, st.s));
    if (color) {
      dc.reserve((size_t)np * 3);
      LCR_HIP(hipMemcpyAsync(dc.p, color, (size_t)np * 3, hipMemcpyHostToDevice, st.s));
    }
    flags.reserve((size_t)np * sizeof(int));
    pos.reserve((size_t)np * sizeof(int));
    tmp.reserve((size_t)scan_tmp_ints(np) * sizeof(int));
    launch_depth_flags(dd.as<uint16_t>(), width, height, P, flags.as<int>(), st.s);
    launch_exclusive_scan(flags.as<int>(), pos.as<int>(), np, tmp.as<int>(), st.s);
    int last[2];
    LCR_HIP(hipMemcpyAsync(&last[0], pos.as<int>() + np - 1, sizeof(int), hipMemcpyDeviceToHost, st.s));
    LCR_HIP(hipMemcpyAsync(&last[1], flags.as<int>() + np - 1, sizeof(int), hipMemcpyDeviceToHost,
                           st.s));
    LCR_HIP(hipStreamSynchronize(st.s));
    const int64_t m = (int64_t)last[0] + last[1];
    ox.reserve((size_t)std::max<int64_t>(m, 1) * 3 * sizeof(float));
    if (out_nrm) on.reserve((size_t)std::max<int64_t>(m, 1) * 3 * sizeof(float));
    if (out_rgb && color) orgb.reserve((size_t)std::max<int64_t>(m, 1) * 3);
    launch_depth_points(dd.as<uint16_t>(), width, height, color ? dc.as<uint8_t>() : nullptr, P,
                        pos.as<int>(), ox.as<float>(), on.p ? on.as<float>() : nullptr,
                        orgb.p ? orgb.as<uint8_t>() : nullptr, st.s);
    if (m > 0) {
      LCR_HIP(hipMemcpyAsync(out_xyz, ox.p, (size_t)m * 3 * sizeof(float), hipMemcpyDeviceToHost, st.s));
      if (on.p)
        LCR_HIP(hipMemcpyAsync(out_nrm, on.p, (size_t)m * 3 * sizeof(float), hipMemcpyDeviceToHost,
                               st.s));
      if (orgb.p) LCR_HIP(hipMemcpyAsync(out_rgb, orgb.p, (size_t)m * 3, hipMemcpyDeviceToHost, st.s));
    }
    LCR_HIP(hipStreamSynchronize(st.s));
    *out_n = m;
    return LCR_OK;
  });
}

}  // extern "C"

namespace {

std::string cycle_tune_key(const lcr_cycle* c) {
  hipDeviceProp_t prop;
  std::string dev = "unknown";
  if (hipGetDeviceProperties(&prop, c->stream.device) == hipSuccess)
    dev = std::string(prop.gcnArchName) + ":" + std::to_string(prop.multiProcessorCount);
  return tune_key(dev, c->desc.n_clouds, c->desc.n_pairs, c->nhyp_total, c->cs.total_raw,
                  c->desc.params.max_corr_dist, c->desc.params.point_to_plane,
                  c->desc.normals_k > 0 ? c->desc.normals_k : 30);
}

}  // namespace

extern "C" {

int lcr_cycle_autotune(lcr_cycle* c, const char* cache_path, lcr_tune_result* out) {
  return guarded([&] {
    if (!c) return fail(LCR_E_ARG, "handle is NULL");
    lcr_tune_result r;
    memset(&r, 0, sizeof r);
    lcr_icp_params& P = c->desc.params;
    const std::string key = cycle_tune_key(c);
    if (cache_path && tune_lookup_file(cache_path, key, &r.tile, &r.precision_mode, &r.cell_factor)) {
      P.tile = r.tile;
      P.precision_mode = r.precision_mode;
      P.cell_factor = r.cell_factor;
      r.from_cache = 1;
      if (out) *out = r;
      return LCR_OK;
    }
    struct Cand {
      int tile, prec;
      double cf;
    };
    std::vector<Cand> cands;
    for (double cf : {1.5, 2.0, 2.5, 3.0, 4.0}) {
      for (int tile : {64, 128, 256}) cands.push_back({tile, LCR_PREC_FP32, cf});
      cands.push_back({64, LCR_PREC_FP16, cf});
      cands.push_back({64, LCR_PREC_BF16, cf});
      cands.push_back({64, LCR_PREC_FP32_TILES, cf});
    }
    const lcr_icp_params saved = P;
    int rc = LCR_OK;
    auto time_run = [&](const Cand& k) -> double {
      P.tile = k.tile;
      P.precision_mode = k.prec;
      P.cell_factor = k.cf;
      rc = lcr_cycle_run(c);  // the first run of a setting captures its graphs
      if (rc != LCR_OK) return -1.0;
      const auto t0 = std::chrono::steady_clock::now();
      rc = lcr_cycle_run(c);
      const auto t1 = std::chrono::steady_clock::now();
      if (rc != LCR_OK) return -1.0;
      return std::chrono::duration<double, std::milli>(t1 - t0).count();
    };
    const double dms = time_run({saved.tile ? saved.tile : 64, saved.precision_mode,
                                 saved.cell_factor > 0.0 ? saved.cell_factor : 2.5});
    if (dms < 0.0) {
      P = saved;
      return rc;
    }
    r.default_ms = dms;
    double best = 1e300;
    Cand bc{64, 0, 2.5};
    for (const Cand& k : cands) {
      const double ms = time_run(k);
      if (ms < 0.0) {
        P = saved;
        return rc;
      }
      ++r.candidates;
      if (ms < best) {
        best = ms;
        bc = k;
      }
    }
    P = saved;
    P.tile = bc.tile;
    P.precision_mode = bc.prec;
    P.cell_factor = bc.cf;
    r.tile = bc.tile;
    r.precision_mode = bc.prec;
    r.cell_factor = bc.cf;
    r.best_ms = best;
    if (cache_path) {
      FILE* f = fopen(cache_path, "a");
      if (f) {
        fprintf(f, "%s\t%d %d %.17g %.6f\n", key.c_str(), bc.tile, bc.prec, bc.cf, best);
        fclose(f);
      }
    }
    if (out) *out = r;
    return LCR_OK;
  });
}

void lcr_cycle_destroy(lcr_cycle* c) {
  if (!c) return;
  try {
    c->stream.bind();
  } catch (...) {
  }
  delete c;
}

int lcr_icp_batch_align(const lcr_cycle_desc* d, lcr_cycle_result* r) {
  lcr_cycle* c = nullptr;
  int rc = lcr_cycle_create(d, &c);
  if (rc != LCR_OK) return rc;
  rc = lcr_cycle_run(c);
  if (rc == LCR_OK) rc = lcr_cycle_results(c, r);
  lcr_cycle_destroy(c);
  return rc;
}

// ------------------------------------------------------------------ profiling

int lcr_prof_enable(int32_t on) {
  Profiler::get().enable(on != 0);
  return LCR_OK;
}

int lcr_prof_reset(void) {
  Profiler::get().reset();
  return LCR_OK;
}

int lcr_prof_read(const char* name, int64_t* launches, double* total_ms) {
  if (!name || !launches || !total_ms) return fail(LCR_E_ARG, "NULL argument");
  Profiler::get().read(name, launches, total_ms);
  return LCR_OK;
}

const char* lcr_last_error(void) { return g_err.c_str(); }

const char* lcr_version(void) { return "lcr 0.1 (gfx950)"; }

}  // extern "C"
