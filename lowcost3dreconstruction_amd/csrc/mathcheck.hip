// mathcheck.hip -- host entry points over the FP64 solvers of lcr_math.hpp (the
// same __host__ __device__ code the ICP and normals kernels run), so the CPU test
// suite can check them against the oracle without a GPU.  Not part of liblcr's
// C ABI.
#include <string.h>

#include "lcr_common.hpp"
#include "lcr_math.hpp"
#include "lcr_search.hpp"
#include "tune_cache.hpp"

#include <math.h>
#include <algorithm>
#include <vector>

using namespace lcr;

extern "C" {

// v: point-to-point moments (17 doubles, lcr_common.hpp layout) -> row-major T
void lcrm_p2p_increment(const double* v, double* T) { p2p_increment(v, T); }

// v: point-to-plane moments (29 doubles) -> T; returns 1 when solved
int lcrm_p2pl_increment(const double* v, double* T) { return p2pl_increment(v, T) ? 1 : 0; }

// S: 3x3 row-major cross-covariance (source first) -> R row-major
void lcrm_horn(const double* S, double* R) {
  double s[3][3], r[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) s[i][j] = S[3 * i + j];
  horn_rotation(s, r);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[3 * i + j] = r[i][j];
}

// A: symmetric 3x3 -> w ascending, V columns
void lcrm_eigh3(const double* A, double* w, double* V) {
  double a[3][3], v[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = A[3 * i + j];
  eigh3(a, w, v);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[3 * i + j] = v[i][j];
}

// One convergence step.  io[0..3] = iterations, state, similar, converged;
// dio[0] = prev_mse.  cst: max_iterations, then (rot_thr, trans_thr, rel, abs).
int lcrm_conv_step(int* io, double* dio, int max_iterations, const double* thr,
                   const double* Tinc, double mse) {
  PairState st{};
  st.iterations = io[0];
  st.state = io[1];
  st.similar = io[2];
  st.prev_mse = dio[0];
  IcpConst c{};
  c.max_iterations = max_iterations;
  c.rotation_threshold = thr[0];
  c.translation_threshold = thr[1];
  c.mse_rel = thr[2];
  c.mse_abs = thr[3];
  const bool done = conv_step(&st, c, Tinc, mse);
  io[1] = st.state;
  io[2] = st.similar;
  io[3] = done ? 1 : 0;
  dio[0] = st.prev_mse;
  return done ? 1 : 0;
}

// The split ring search with its lanes run in turn: each lane scans its share of
// a ring's rows (ring_rows) from the shared best, then every lane takes the
// minimum by (d2, id).
static NNResult ring_split_host(const GridDesc& g, const int* cs, const float4* P, float qx, float qy,
                                float qz, float r, int seed, int nsub) {
  NNResult res{INFINITY, 0x7fffffff, -1, false};
  if (g.npts == 0 || !(isfinite(qx) && isfinite(qy) && isfinite(qz))) return res;
  const CellPos c = cell_pos(g, qx, qy, qz);
  const float rlim = r * 1.00001f;
  const int k0 = ring_start(g, c, rlim);
  if (k0 < 0) return res;
  const int kmax = k0 + std::max(g.nx, std::max(g.ny, g.nz));
  float best = INFINITY;
  int bid = 0x7fffffff, bj = -1;
  if (seed >= 0) scan_points(P, seed, seed + 1, qx, qy, qz, best, bid, bj);
  const float h2 = g.h * g.h, lim2 = rlim * rlim;
  for (int k = k0; k <= kmax; ++k) {
    float lb[8];
    int lid[8], lj[8];
    for (int l = 0; l < nsub; ++l) {
      lb[l] = best;
      lid[l] = bid;
      lj[l] = bj;
      ring_rows(g, cs, P, c, k, l, nsub, h2, lim2, qx, qy, qz, lb[l], lid[l], lj[l]);
    }
    for (int l = 0; l < nsub; ++l) {
      if (lb[l] < best || (lb[l] == best && lid[l] < bid)) {
        best = lb[l];
        bid = lid[l];
        bj = lj[l];
      }
    }
    if (ring_stop(g, c, k, best, rlim)) break;
  }
  res.d2 = best;
  res.id = bid;
  res.j = bj;
  return res;
}

// Host build of one cloud's grid in the device layout (cell edge h, origin at the
// bounding-box minimum, points sorted by cell then by index), then nn_query for
// every query: exercises the ring search and its certification bound on the CPU.
// Returns 0, or -1 when the grid would be too large.
// seed_ids (nullable): per query an original target index used as the search's
// seed candidate (-1 = none).
// ring_limit < 0: no tree (ring search only); otherwise the tree takes over after
// ring_limit rings, as on the device.
// nsub > 0: the bounded ring-only search with each ring's rows dealt to nsub lanes
// (run here one after another) and the minimum over the lanes taken after every
// ring, as in the ICP search for launches with few live tiles.
int lcrm_nn_grid(const float* tgt, int64_t nt, const float* q, int64_t nq, float h, float r,
                 int bounded, const int32_t* seed_ids, int ring_limit, int32_t* idx, float* d2,
                 int nsub) {
  if (nsub < 0 || nsub > 8 || (nsub > 0 && !bounded)) return -2;
  GridDesc g{};
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  int npts = 0;
  for (int64_t i = 0; i < nt; ++i) {
    const float* p = tgt + 3 * i;
    if (!(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]))) continue;
    ++npts;
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], p[a]);
      hi[a] = fmaxf(hi[a], p[a]);
    }
  }
  g.npts = npts;
  g.h = h;
  g.inv_h = (float)(1.0 / (double)h);
  if (npts == 0) {
    g.nx = g.ny = g.nz = 1;
  } else {
    g.ox = lo[0];
    g.oy = lo[1];
    g.oz = lo[2];
    g.nx = (int)(floor(((double)hi[0] - lo[0]) * (double)g.inv_h) + 1);
    g.ny = (int)(floor(((double)hi[1] - lo[1]) * (double)g.inv_h) + 1);
    g.nz = (int)(floor(((double)hi[2] - lo[2]) * (double)g.inv_h) + 1);
  }
  const int64_t ncells = (int64_t)g.nx * g.ny * g.nz;
  if (ncells > (int64_t)1 << 26) return -1;
  std::vector<int> cnt(ncells + 1, 0), cid(nt, -1);
  for (int64_t i = 0; i < nt; ++i) {
    const float* p = tgt + 3 * i;
    if (!(isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]))) continue;
    int x = (int)floorf((p[0] - g.ox) * g.inv_h), y = (int)floorf((p[1] - g.oy) * g.inv_h),
        z = (int)floorf((p[2] - g.oz) * g.inv_h);
    x = x < 0 ? 0 : (x > g.nx - 1 ? g.nx - 1 : x);
    y = y < 0 ? 0 : (y > g.ny - 1 ? g.ny - 1 : y);
    z = z < 0 ? 0 : (z > g.nz - 1 ? g.nz - 1 : z);
    cid[i] = (z * g.ny + y) * g.nx + x;
    cnt[cid[i]]++;
  }
  std::vector<int> cs(ncells + 1, 0);
  for (int64_t c = 0; c < ncells; ++c) cs[c + 1] = cs[c] + cnt[c];
  std::vector<int> fill(cs.begin(), cs.end() - 1);
  std::vector<float4> P((size_t)(npts > 0 ? npts : 1));
  for (int64_t i = 0; i < nt; ++i) {
    if (cid[i] < 0) continue;
    const float* p = tgt + 3 * i;
    P[fill[cid[i]]++] = make_float4(p[0], p[1], p[2], __builtin_bit_cast(float, (int)i));
  }
  // tree order (Morton keys, stable) and the bounding-volume tree over it, in the
  // device layout
  std::vector<int> tpos((size_t)(npts > 0 ? npts : 1));
  std::vector<uint32_t> tkey((size_t)(npts > 0 ? npts : 1));
  for (int j = 0; j < npts; ++j) {
    tpos[j] = j;
    tkey[j] = tree_key(g, P[j].x, P[j].y, P[j].z);
  }
  std::stable_sort(tpos.begin(), tpos.begin() + npts, [&](int a, int b) { return tkey[a] < tkey[b]; });
  std::vector<float4> TP((size_t)(npts > 0 ? npts : 1));
  for (int i = 0; i < npts; ++i) TP[i] = P[tpos[i]];
  const int64_t leaves = (npts + kLeaf - 1) / kLeaf;
  int depth = 0;
  while (((int64_t)1 << depth) < leaves) ++depth;
  g.depth = depth;
  g.node_off = 0;
  g.pt_off = 0;
  const int64_t nslots = (int64_t)1 << depth;
  std::vector<float4> nodes((size_t)(2 * nslots) * 2);
  for (int64_t i = 0; i < nslots; ++i) {
    float4 lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f), hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
    for (int64_t j = i * kLeaf; j < std::min<int64_t>(i * kLeaf + kLeaf, npts); ++j) {
      lo.x = fminf(lo.x, TP[j].x); lo.y = fminf(lo.y, TP[j].y); lo.z = fminf(lo.z, TP[j].z);
      hi.x = fmaxf(hi.x, TP[j].x); hi.y = fmaxf(hi.y, TP[j].y); hi.z = fmaxf(hi.z, TP[j].z);
    }
    nodes[2 * (nslots + i)] = lo;
    nodes[2 * (nslots + i) + 1] = hi;
  }
  for (int64_t node = nslots - 1; node >= 1; --node) {
    const float4 l0 = nodes[4 * node], l1 = nodes[4 * node + 1], r0 = nodes[4 * node + 2], r1 = nodes[4 * node + 3];
    nodes[2 * node] = make_float4(fminf(l0.x, r0.x), fminf(l0.y, r0.y), fminf(l0.z, r0.z), 0.f);
    nodes[2 * node + 1] = make_float4(fmaxf(l1.x, r1.x), fmaxf(l1.y, r1.y), fmaxf(l1.z, r1.z), 0.f);
  }
  std::vector<int> pos(nt, -1);
  for (int j = 0; j < npts; ++j) pos[__builtin_bit_cast(int, P[j].w)] = j;
  for (int64_t i = 0; i < nq; ++i) {
    const int seed = (seed_ids && seed_ids[i] >= 0) ? pos[seed_ids[i]] : -1;
    const Tree tr{ring_limit >= 0 ? nodes.data() : nullptr, TP.data(), tpos.data()};
    const NNResult res =
        nsub > 0 ? ring_split_host(g, cs.data(), P.data(), q[3 * i], q[3 * i + 1], q[3 * i + 2], r,
                                   seed, nsub)
                 : nn_query(g, cs.data(), P.data(), tr, q[3 * i], q[3 * i + 1], q[3 * i + 2], r,
                            bounded != 0, seed, ring_limit >= 0 ? ring_limit : 0);
    const bool keep = res.j >= 0 && (!bounded || (double)res.d2 <= (double)r * (double)r);
    idx[i] = keep ? res.id : -1;
    d2[i] = res.d2;
  }
  return 0;
}

// C = A * B (4x4 row-major)
void lcrm_mat4_mul(const double* A, const double* B, double* C) { mat4_mul(A, B, C); }

// The autotuner's cache key and lookup (tune_cache.hpp).  key: buffer of `len`
// bytes; returns the key's length.
int lcrm_tune_key(const char* device, int clouds, int pairs, int hyp, int64_t points, double r,
                  int p2pl, int k, char* key, int len) {
  const std::string s = tune_key(device, clouds, pairs, hyp, points, r, p2pl, k);
  snprintf(key, (size_t)len, "%s", s.c_str());
  return (int)s.size();
}

// 1 and the entry's (tile, precision_mode, cell_factor) when `key` is cached in
// the file at `path` (the last valid entry wins), else 0.
int lcrm_tune_lookup(const char* path, const char* key, int* tile, int* prec, double* cf) {
  return tune_lookup_file(path, key, tile, prec, cf) ? 1 : 0;
}

}  // extern "C"
