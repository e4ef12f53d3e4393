/*
 * lcr_oracle.c -- CPU oracle for the fine-registration hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against, and the
 * timed host baseline.  The product library never links or calls this code.
 * Built with -ffp-contract=off so every float/double operation rounds exactly as
 * written (FLANN's float distance is evaluated without FMA contraction).
 *
 * Each function cites the PCL routine it restates (PCL 1.7.2-1.10; PCL is not
 * vendored in the reference, see lcr_oracle.h) and the reference call site.
 */
#include "lcr_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* FLANN L2_Simple<float> (flann/algorithms/dist.h): result accumulates in float
 * starting at 0, diff = query - point.  Used through KdTreeFLANN by
 * CorrespondenceEstimation (fine_registration.cpp:121) and NormalEstimation
 * (normal_estimation.cpp:108). */
static inline float fl_d2(const float *q, const float *p) {
  float r = 0.0f, d;
  d = q[0] - p[0];
  r += d * d;
  d = q[1] - p[1];
  r += d * d;
  d = q[2] - p[2];
  r += d * d;
  return r;
}

static inline int pt_finite(const float *p) {
  return isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]);
}

/* (d, i) < (e, j) lexicographically: the tie rule shared by every search here */
static inline int lex_less(float d, int32_t i, float e, int32_t j) {
  return d < e || (d == e && i < j);
}

/* ------------------------------------------------------------------------- */
/* kd-tree (exact search, like FLANN KDTreeSingleIndex with eps = 0, leaf size 15) */

typedef struct {
  int32_t begin, end;  /* range in the reordered point array */
  int32_t left, right; /* child nodes or -1 */
  float lo[3], hi[3];  /* bounding box of the node's points */
} onode;

struct okd_tree {
  float *p;      /* reordered points, 3 floats each */
  int32_t *id;   /* original index of each reordered point */
  int64_t n;
  onode *nodes;
  int32_t nnodes, cap;
};

#define OKD_LEAF 15

static void select_kth(int32_t *perm, const float *pts, int axis, int32_t lo, int32_t hi,
                       int32_t k) {
  /* quickselect on perm[lo..hi] (inclusive) so perm[k] is the k-th by coordinate */
  while (hi > lo) {
    float pivot = pts[3 * (int64_t)perm[(lo + hi) / 2] + axis];
    int32_t i = lo, j = hi;
    while (i <= j) {
      while (pts[3 * (int64_t)perm[i] + axis] < pivot) i++;
      while (pts[3 * (int64_t)perm[j] + axis] > pivot) j--;
      if (i <= j) {
        int32_t t = perm[i];
        perm[i] = perm[j];
        perm[j] = t;
        i++;
        j--;
      }
    }
    if (k <= j)
      hi = j;
    else if (k >= i)
      lo = i;
    else
      return;
  }
}

static int32_t okd_new_node(okd_tree *t) {
  if (t->nnodes == t->cap) {
    t->cap = t->cap ? 2 * t->cap : 64;
    t->nodes = (onode *)realloc(t->nodes, sizeof(onode) * (size_t)t->cap);
  }
  return t->nnodes++;
}

static int32_t okd_build_rec(okd_tree *t, int32_t *perm, const float *pts, int32_t b,
                             int32_t e) {
  int32_t me = okd_new_node(t);
  onode nd;
  nd.begin = b;
  nd.end = e;
  nd.left = nd.right = -1;
  for (int a = 0; a < 3; a++) {
    nd.lo[a] = INFINITY;
    nd.hi[a] = -INFINITY;
  }
  for (int32_t i = b; i < e; i++) {
    const float *q = pts + 3 * (int64_t)perm[i];
    for (int a = 0; a < 3; a++) {
      if (q[a] < nd.lo[a]) nd.lo[a] = q[a];
      if (q[a] > nd.hi[a]) nd.hi[a] = q[a];
    }
  }
  if (e - b > OKD_LEAF) {
    int axis = 0;
    float ext = nd.hi[0] - nd.lo[0];
    for (int a = 1; a < 3; a++)
      if (nd.hi[a] - nd.lo[a] > ext) {
        ext = nd.hi[a] - nd.lo[a];
        axis = a;
      }
    int32_t mid = b + (e - b) / 2;
    select_kth(perm, pts, axis, b, e - 1, mid);
    int32_t l = okd_build_rec(t, perm, pts, b, mid);
    int32_t r = okd_build_rec(t, perm, pts, mid, e);
    nd.left = l;
    nd.right = r;
  }
  t->nodes[me] = nd;
  return me;
}

okd_tree *okd_build(const float *xyz, int64_t n, int64_t stride) {
  okd_tree *t = (okd_tree *)calloc(1, sizeof(okd_tree));
  float *pts = (float *)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
  int32_t *perm = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int64_t m = 0;
  for (int64_t i = 0; i < n; i++) {
    const float *q = xyz + i * stride;
    pts[3 * i + 0] = q[0];
    pts[3 * i + 1] = q[1];
    pts[3 * i + 2] = q[2];
    if (pt_finite(q)) perm[m++] = (int32_t)i;
  }
  t->n = m;
  if (m > 0) okd_build_rec(t, perm, pts, 0, (int32_t)m);
  t->p = (float *)malloc(sizeof(float) * 3 * (size_t)(m > 0 ? m : 1));
  t->id = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
  for (int64_t i = 0; i < m; i++) {
    memcpy(t->p + 3 * i, pts + 3 * (int64_t)perm[i], 3 * sizeof(float));
    t->id[i] = perm[i];
  }
  free(pts);
  free(perm);
  return t;
}

void okd_free(okd_tree *t) {
  if (!t) return;
  free(t->p);
  free(t->id);
  free(t->nodes);
  free(t);
}

static inline double box_d2(const onode *nd, const float *q) {
  double s = 0.0;
  for (int a = 0; a < 3; a++) {
    double d = 0.0;
    if (q[a] < nd->lo[a])
      d = (double)nd->lo[a] - (double)q[a];
    else if (q[a] > nd->hi[a])
      d = (double)q[a] - (double)nd->hi[a];
    s += d * d;
  }
  return s;
}

/* A float d2 may round below the exact squared distance by a few ulps, so a node
 * is only pruned when its exact lower bound clears the bound by a margin; the
 * result then equals brute force on the float distances. */
#define PRUNE_MARGIN (1.0 + 1e-6)

typedef struct {
  float best;
  int32_t bidx;
  double bound; /* max_d2 */
} nn_state;

static void nn_rec(const okd_tree *t, int32_t ni, const float *q, nn_state *s) {
  const onode *nd = &t->nodes[ni];
  if (nd->left < 0) {
    for (int32_t i = nd->begin; i < nd->end; i++) {
      float d = fl_d2(q, t->p + 3 * (int64_t)i);
      if (lex_less(d, t->id[i], s->best, s->bidx)) {
        s->best = d;
        s->bidx = t->id[i];
      }
    }
    return;
  }
  double dl = box_d2(&t->nodes[nd->left], q);
  double dr = box_d2(&t->nodes[nd->right], q);
  int32_t first = dl <= dr ? nd->left : nd->right;
  int32_t second = dl <= dr ? nd->right : nd->left;
  double d1 = dl <= dr ? dl : dr, d2v = dl <= dr ? dr : dl;
  double thr = (s->bidx >= 0 ? (double)s->best : s->bound) * PRUNE_MARGIN;
  if (d1 <= thr) nn_rec(t, first, q, s);
  thr = (s->bidx >= 0 ? (double)s->best : s->bound) * PRUNE_MARGIN;
  if (d2v <= thr) nn_rec(t, second, q, s);
}

int okd_nn(const okd_tree *t, const float q[3], double max_d2, int32_t *idx, float *d2) {
  *idx = -1;
  *d2 = INFINITY;
  if (t->n == 0 || !pt_finite(q)) return 0;
  nn_state s;
  s.best = INFINITY;
  s.bidx = -1;
  s.bound = max_d2;
  if (box_d2(&t->nodes[0], q) <= max_d2 * PRUNE_MARGIN) nn_rec(t, 0, q, &s);
  if (s.bidx < 0 || !((double)s.best <= max_d2)) {
    if (s.bidx >= 0) *d2 = s.best;
    return 0;
  }
  *idx = s.bidx;
  *d2 = s.best;
  return 1;
}

typedef struct {
  float *hd;
  int32_t *hi;
  int k, cnt;
} knn_heap; /* max-heap by (d, idx) */

static void heap_push(knn_heap *h, float d, int32_t i) {
  if (h->cnt < h->k) {
    int c = h->cnt++;
    while (c > 0) {
      int p = (c - 1) / 2;
      if (!lex_less(h->hd[p], h->hi[p], d, i)) break;
      h->hd[c] = h->hd[p];
      h->hi[c] = h->hi[p];
      c = p;
    }
    h->hd[c] = d;
    h->hi[c] = i;
    return;
  }
  if (!lex_less(d, i, h->hd[0], h->hi[0])) return;
  int c = 0;
  for (;;) {
    int l = 2 * c + 1, r = l + 1, m = c;
    float md = d;
    int32_t mi = i;
    if (l < h->cnt && lex_less(md, mi, h->hd[l], h->hi[l])) {
      m = l;
      md = h->hd[l];
      mi = h->hi[l];
    }
    if (r < h->cnt && lex_less(md, mi, h->hd[r], h->hi[r])) {
      m = r;
    }
    if (m == c) break;
    h->hd[c] = h->hd[m];
    h->hi[c] = h->hi[m];
    c = m;
  }
  h->hd[c] = d;
  h->hi[c] = i;
}

static void knn_rec(const okd_tree *t, int32_t ni, const float *q, knn_heap *h) {
  const onode *nd = &t->nodes[ni];
  if (nd->left < 0) {
    for (int32_t i = nd->begin; i < nd->end; i++)
      heap_push(h, fl_d2(q, t->p + 3 * (int64_t)i), t->id[i]);
    return;
  }
  double dl = box_d2(&t->nodes[nd->left], q);
  double dr = box_d2(&t->nodes[nd->right], q);
  int32_t first = dl <= dr ? nd->left : nd->right;
  int32_t second = dl <= dr ? nd->right : nd->left;
  double d1 = dl <= dr ? dl : dr, d2v = dl <= dr ? dr : dl;
  if (h->cnt < h->k || d1 <= (double)h->hd[0] * PRUNE_MARGIN) knn_rec(t, first, q, h);
  if (h->cnt < h->k || d2v <= (double)h->hd[0] * PRUNE_MARGIN) knn_rec(t, second, q, h);
}

static void sort_pairs(float *d, int32_t *i, int n) { /* insertion sort, small n */
  for (int a = 1; a < n; a++) {
    float dv = d[a];
    int32_t iv = i[a];
    int b = a - 1;
    while (b >= 0 && lex_less(dv, iv, d[b], i[b])) {
      d[b + 1] = d[b];
      i[b + 1] = i[b];
      b--;
    }
    d[b + 1] = dv;
    i[b + 1] = iv;
  }
}

int okd_knn(const okd_tree *t, const float q[3], int k, int32_t *idx, float *d2) {
  if (t->n == 0 || k <= 0 || !pt_finite(q)) return 0;
  knn_heap h;
  h.hd = d2;
  h.hi = idx;
  h.k = k;
  h.cnt = 0;
  knn_rec(t, 0, q, &h);
  sort_pairs(d2, idx, h.cnt);
  return h.cnt;
}

void obf_nn(const float *tgt, int64_t nt, const float *q, int64_t nq, double max_d2,
            int32_t *idx, float *d2) {
  for (int64_t i = 0; i < nq; i++) {
    float best = INFINITY;
    int32_t bi = -1;
    const float *qi = q + 3 * i;
    if (pt_finite(qi)) {
      for (int64_t j = 0; j < nt; j++) {
        const float *p = tgt + 3 * j;
        if (!pt_finite(p)) continue;
        float d = fl_d2(qi, p);
        if (lex_less(d, (int32_t)j, best, bi)) {
          best = d;
          bi = (int32_t)j;
        }
      }
    }
    d2[i] = best;
    idx[i] = (bi >= 0 && (double)best <= max_d2) ? bi : -1;
  }
}

int obf_knn(const float *tgt, int64_t nt, const float q[3], int k, int32_t *idx, float *d2) {
  if (k <= 0 || !pt_finite(q)) return 0;
  knn_heap h;
  h.hd = d2;
  h.hi = idx;
  h.k = k;
  h.cnt = 0;
  for (int64_t j = 0; j < nt; j++) {
    const float *p = tgt + 3 * j;
    if (!pt_finite(p)) continue;
    heap_push(&h, fl_d2(q, p), (int32_t)j);
  }
  sort_pairs(d2, idx, h.cnt);
  return h.cnt;
}

void oracle_nn(const float *tgt, int64_t nt, const float *q, int64_t nq, double max_d2,
               int32_t *idx, float *d2) {
  okd_tree *t = okd_build(tgt, nt, 3);
  for (int64_t i = 0; i < nq; i++) okd_nn(t, q + 3 * i, max_d2, idx + i, d2 + i);
  okd_free(t);
}

/* ------------------------------------------------------------------------- */
/* small dense linear algebra, FP64 */

static void mat3_mul(const double *A, const double *B, double *C) {
  double t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) s += A[3 * i + k] * B[3 * k + j];
      t[3 * i + j] = s;
    }
  memcpy(C, t, sizeof t);
}

static double det3(const double *A) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
         A[2] * (A[3] * A[7] - A[4] * A[6]);
}

static void mat4_mul(const double *A, const double *B, double *C) {
  double t[16];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s += A[4 * i + k] * B[4 * k + j];
      t[4 * i + j] = s;
    }
  memcpy(C, t, sizeof t);
}

static void mat4f_mul(const float *A, const float *B, float *C) {
  float t[16];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      float s = 0.0f;
      for (int k = 0; k < 4; k++) s += A[4 * i + k] * B[4 * k + j];
      t[4 * i + j] = s;
    }
  memcpy(C, t, sizeof t);
}

static void mat4_identity(double *T) {
  memset(T, 0, 16 * sizeof(double));
  T[0] = T[5] = T[10] = T[15] = 1.0;
}

/* One-sided (Hestenes) Jacobi SVD, FP64.  Plays the role of Eigen::JacobiSVD
 * inside Eigen::umeyama; singular values sorted descending as Eigen does. */
void oracle_svd3(const double A[9], double U[9], double s[3], double V[9]) {
  double W[9];
  memcpy(W, A, sizeof W);
  double Vm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int sweep = 0; sweep < 60; sweep++) {
    int rotated = 0;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        double al = 0, be = 0, ga = 0;
        for (int i = 0; i < 3; i++) {
          al += W[3 * i + p] * W[3 * i + p];
          be += W[3 * i + q] * W[3 * i + q];
          ga += W[3 * i + p] * W[3 * i + q];
        }
        if (ga == 0.0 || fabs(ga) <= 1e-300 || fabs(ga) <= 1e-17 * sqrt(al * be)) continue;
        rotated = 1;
        double zeta = (be - al) / (2.0 * ga);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
        for (int i = 0; i < 3; i++) {
          double wp = W[3 * i + p], wq = W[3 * i + q];
          W[3 * i + p] = c * wp - sn * wq;
          W[3 * i + q] = sn * wp + c * wq;
          double vp = Vm[3 * i + p], vq = Vm[3 * i + q];
          Vm[3 * i + p] = c * vp - sn * vq;
          Vm[3 * i + q] = sn * vp + c * vq;
        }
      }
    if (!rotated) break;
  }
  double sv[3];
  for (int j = 0; j < 3; j++)
    sv[j] = sqrt(W[j] * W[j] + W[3 + j] * W[3 + j] + W[6 + j] * W[6 + j]);
  int ord[3] = {0, 1, 2};
  for (int a = 0; a < 3; a++)
    for (int b = a + 1; b < 3; b++)
      if (sv[ord[b]] > sv[ord[a]]) {
        int t = ord[a];
        ord[a] = ord[b];
        ord[b] = t;
      }
  double Uo[9], Vo[9];
  for (int j = 0; j < 3; j++) {
    int c = ord[j];
    s[j] = sv[c];
    for (int i = 0; i < 3; i++) Vo[3 * i + j] = Vm[3 * i + c];
    for (int i = 0; i < 3; i++) Uo[3 * i + j] = sv[c] > 0 ? W[3 * i + c] / sv[c] : 0.0;
  }
  /* complete U for (numerically) rank-deficient A */
  double smax = s[0] > 0 ? s[0] : 1.0;
  if (s[1] <= 1e-14 * smax) {
    /* rank <= 1: build an orthonormal completion of u0 */
    double u0[3] = {Uo[0], Uo[3], Uo[6]};
    if (s[0] == 0.0) {
      u0[0] = 1;
      u0[1] = 0;
      u0[2] = 0;
    }
    double a[3] = {1, 0, 0};
    if (fabs(u0[0]) > 0.9) {
      a[0] = 0;
      a[1] = 1;
    }
    double u1[3] = {u0[1] * a[2] - u0[2] * a[1], u0[2] * a[0] - u0[0] * a[2],
                    u0[0] * a[1] - u0[1] * a[0]};
    double nn = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    for (int i = 0; i < 3; i++) u1[i] /= nn;
    for (int i = 0; i < 3; i++) {
      Uo[3 * i + 0] = u0[i];
      Uo[3 * i + 1] = u1[i];
    }
  }
  if (s[2] <= 1e-14 * smax) {
    double c0[3] = {Uo[0], Uo[3], Uo[6]}, c1[3] = {Uo[1], Uo[4], Uo[7]};
    Uo[2] = c0[1] * c1[2] - c0[2] * c1[1];
    Uo[5] = c0[2] * c1[0] - c0[0] * c1[2];
    Uo[8] = c0[0] * c1[1] - c0[1] * c1[0];
  }
  memcpy(U, Uo, sizeof Uo);
  memcpy(V, Vo, sizeof Vo);
}

/* FP64 cyclic Jacobi eigen-decomposition of a symmetric 3x3 (ascending). */
void oracle_eigh3(const double A[9], double w[3], double V[9]) {
  double a[9];
  memcpy(a, A, sizeof a);
  double v[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
    double dg = a[0] * a[0] + a[4] * a[4] + a[8] * a[8];
    if (off == 0.0 || off <= 1e-34 * dg) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        double apq = a[3 * p + q];
        if (apq == 0.0) continue;
        double app = a[3 * p + p], aqq = a[3 * q + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; k++) { /* A <- A J (columns p,q) */
          double akp = a[3 * k + p], akq = a[3 * k + q];
          a[3 * k + p] = c * akp - s * akq;
          a[3 * k + q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; k++) { /* A <- J^T A (rows p,q) */
          double apk = a[3 * p + k], aqk = a[3 * q + k];
          a[3 * p + k] = c * apk - s * aqk;
          a[3 * q + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; k++) {
          double vkp = v[3 * k + p], vkq = v[3 * k + q];
          v[3 * k + p] = c * vkp - s * vkq;
          v[3 * k + q] = s * vkp + c * vkq;
        }
      }
  }
  double d[3] = {a[0], a[4], a[8]};
  int ord[3] = {0, 1, 2};
  for (int i = 0; i < 3; i++)
    for (int j = i + 1; j < 3; j++)
      if (d[ord[j]] < d[ord[i]]) {
        int t = ord[i];
        ord[i] = ord[j];
        ord[j] = t;
      }
  for (int j = 0; j < 3; j++) {
    w[j] = d[ord[j]];
    for (int i = 0; i < 3; i++) V[3 * i + j] = v[3 * i + ord[j]];
  }
}

/* Eigen::umeyama(src, dst, with_scaling=false) as called by
 * TransformationEstimationSVD::getTransformationFromCorrelation path (use_umeyama_ =
 * true): centred means, sigma = dst_demean * src_demean^T / n, JacobiSVD, S(2) = -1
 * when det(U) det(V) < 0, R = U S V^T, t = mu_dst - R mu_src. */
static void umeyama_from_moments(const double mus[3], const double mud[3],
                                 const double Sig[9], double T[16]) {
  double U[9], s[3], V[9];
  oracle_svd3(Sig, U, s, V);
  double S[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  if (det3(U) * det3(V) < 0) S[8] = -1.0;
  double Vt[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Vt[3 * i + j] = V[3 * j + i];
  double R[9];
  mat3_mul(U, S, R);
  mat3_mul(R, Vt, R);
  mat4_identity(T);
  for (int i = 0; i < 3; i++) {
    for (int j = 0; j < 3; j++) T[4 * i + j] = R[3 * i + j];
    T[4 * i + 3] = mud[i] - (R[3 * i + 0] * mus[0] + R[3 * i + 1] * mus[1] + R[3 * i + 2] * mus[2]);
  }
}

int oracle_umeyama(const double *src, const double *dst, int64_t n, double T[16]) {
  mat4_identity(T);
  if (n < 1) return -1;
  double mus[3] = {0, 0, 0}, mud[3] = {0, 0, 0};
  for (int64_t i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) {
      mus[a] += src[3 * i + a];
      mud[a] += dst[3 * i + a];
    }
  for (int a = 0; a < 3; a++) {
    mus[a] /= (double)n;
    mud[a] /= (double)n;
  }
  double Sig[9] = {0};
  for (int64_t i = 0; i < n; i++) {
    double ds[3], dd[3];
    for (int a = 0; a < 3; a++) {
      ds[a] = src[3 * i + a] - mus[a];
      dd[a] = dst[3 * i + a] - mud[a];
    }
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) Sig[3 * r + c] += dd[r] * ds[c];
  }
  for (int k = 0; k < 9; k++) Sig[k] /= (double)n;
  umeyama_from_moments(mus, mud, Sig, T);
  return 0;
}

/* TransformationEstimationPointToPlaneLLS::constructTransformationMatrix */
void oracle_p2pl_matrix(double al, double be, double ga, double tx, double ty, double tz,
                        double T[16]) {
  mat4_identity(T);
  double ca = cos(al), sa = sin(al), cb = cos(be), sb = sin(be), cg = cos(ga), sg = sin(ga);
  T[0] = cg * cb;
  T[1] = -sg * ca + cg * sb * sa;
  T[2] = sg * sa + cg * sb * ca;
  T[4] = sg * cb;
  T[5] = cg * ca + sg * sb * sa;
  T[6] = -cg * sa + sg * sb * ca;
  T[8] = -sb;
  T[9] = cb * sa;
  T[10] = cb * ca;
  T[3] = tx;
  T[7] = ty;
  T[11] = tz;
}

/* Solve the 6x6 system M x = b by Gaussian elimination with partial pivoting
 * (PCL uses ATA.inverse() * ATb). */
static int solve6(const double M[36], const double b[6], double x[6]) {
  double A[6][7];
  for (int i = 0; i < 6; i++) {
    for (int j = 0; j < 6; j++) A[i][j] = M[6 * i + j];
    A[i][6] = b[i];
  }
  for (int c = 0; c < 6; c++) {
    int piv = c;
    for (int r = c + 1; r < 6; r++)
      if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
    if (!(fabs(A[piv][c]) > 1e-300)) return -1;
    if (piv != c)
      for (int j = 0; j < 7; j++) {
        double t = A[c][j];
        A[c][j] = A[piv][j];
        A[piv][j] = t;
      }
    for (int r = c + 1; r < 6; r++) {
      double f = A[r][c] / A[c][c];
      for (int j = c; j < 7; j++) A[r][j] -= f * A[c][j];
    }
  }
  for (int i = 5; i >= 0; i--) {
    double s = A[i][6];
    for (int j = i + 1; j < 6; j++) s -= A[i][j] * x[j];
    x[i] = s / A[i][i];
  }
  return 0;
}

/* Accumulate one point-to-plane row [s x n ; n] with residual d = n.(dst - src)
 * into the upper triangle of AtA and into Atb (transformation_estimation_point_to_plane_lls.hpp:
 * a = nz*sy - ny*sz, b = nx*sz - nz*sx, c = ny*sx - nx*sy). */
static void p2pl_accumulate(double sx, double sy, double sz, double dx, double dy, double dz,
                            double nx, double ny, double nz, double AtA[36], double Atb[6]) {
  double row[6] = {nz * sy - ny * sz, nx * sz - nz * sx, ny * sx - nx * sy, nx, ny, nz};
  double d = nx * dx + ny * dy + nz * dz - nx * sx - ny * sy - nz * sz;
  for (int i = 0; i < 6; i++) {
    for (int j = i; j < 6; j++) AtA[6 * i + j] += row[i] * row[j];
    Atb[i] += row[i] * d;
  }
}

static int p2pl_solve(double AtA[36], const double Atb[6], double T[16]) {
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < i; j++) AtA[6 * i + j] = AtA[6 * j + i];
  double x[6];
  if (solve6(AtA, Atb, x) != 0) {
    mat4_identity(T);
    return -1;
  }
  oracle_p2pl_matrix(x[0], x[1], x[2], x[3], x[4], x[5], T);
  return 0;
}

int oracle_p2pl(const double *src, const double *dst, const double *nrm, int64_t n,
                double T[16]) {
  double AtA[36] = {0}, Atb[6] = {0};
  for (int64_t i = 0; i < n; i++) {
    const double *s = src + 3 * i, *d = dst + 3 * i, *m = nrm + 3 * i;
    if (!isfinite(s[0]) || !isfinite(s[1]) || !isfinite(s[2]) || !isfinite(d[0]) ||
        !isfinite(d[1]) || !isfinite(d[2]) || !isfinite(m[0]) || !isfinite(m[1]) ||
        !isfinite(m[2]))
      continue;
    p2pl_accumulate(s[0], s[1], s[2], d[0], d[1], d[2], m[0], m[1], m[2], AtA, Atb);
  }
  return p2pl_solve(AtA, Atb, T);
}

/* ------------------------------------------------------------------------- */
/* DefaultConvergenceCriteria (default_convergence_criteria.hpp) */

void oconv_init(oconv *c, int max_it, double trans_eps, double fit_eps) {
  c->max_iterations = max_it;
  c->rotation_threshold = 1.0 - trans_eps;
  c->translation_threshold = trans_eps;
  c->mse_threshold_relative = fit_eps;
  c->mse_threshold_absolute = 1e-12;
  c->prev_mse = DBL_MAX;
  c->iterations_similar = 0;
  c->max_iterations_similar = 0;
  c->state = OR_CONV_NOT_CONVERGED;
}

int oconv_step(oconv *c, int iterations, const double Tinc[16], double mse) {
  int is_similar = 0;
  if (iterations >= c->max_iterations) {
    c->state = OR_CONV_ITERATIONS;
    return 1;
  }
  /* PCL reads the float Matrix4 coefficients; the sums are float arithmetic */
  float t00 = (float)Tinc[0], t11 = (float)Tinc[5], t22 = (float)Tinc[10];
  float t03 = (float)Tinc[3], t13 = (float)Tinc[7], t23 = (float)Tinc[11];
  double cos_angle = 0.5 * (double)(t00 + t11 + t22 - 1.0f);
  double translation_sqr = (double)(t03 * t03 + t13 * t13 + t23 * t23);
  if (cos_angle >= c->rotation_threshold && translation_sqr <= c->translation_threshold) {
    if (c->iterations_similar >= c->max_iterations_similar) {
      c->state = OR_CONV_TRANSFORM;
      return 1;
    }
    is_similar = 1;
  }
  if (fabs(mse - c->prev_mse) < c->mse_threshold_absolute) {
    if (c->iterations_similar >= c->max_iterations_similar) {
      c->state = OR_CONV_ABS_MSE;
      return 1;
    }
    is_similar = 1;
  }
  if (fabs(mse - c->prev_mse) / c->prev_mse < c->mse_threshold_relative) {
    if (c->iterations_similar >= c->max_iterations_similar) {
      c->state = OR_CONV_REL_MSE;
      return 1;
    }
    is_similar = 1;
  }
  if (is_similar)
    c->iterations_similar++;
  else
    c->iterations_similar = 0;
  c->prev_mse = mse;
  return 0;
}

/* ------------------------------------------------------------------------- */
/* transforms */

void oracle_transform_points(const float *src, int64_t n, const double T[16], float *out) {
  for (int64_t i = 0; i < n; i++) {
    double x = src[3 * i], y = src[3 * i + 1], z = src[3 * i + 2];
    for (int r = 0; r < 3; r++) {
      double v = T[4 * r + 0] * x;
      v = v + T[4 * r + 1] * y;
      v = v + T[4 * r + 2] * z;
      v = v + T[4 * r + 3];
      out[3 * i + r] = (float)v;
    }
  }
}

/* IterativeClosestPoint::transformCloud in float (Eigen Matrix4f * Vector4f) */
static void transform_points_f32(const float *src, int64_t n, const float Tf[16], float *out) {
  for (int64_t i = 0; i < n; i++) {
    float x = src[3 * i], y = src[3 * i + 1], z = src[3 * i + 2];
    float o[3];
    for (int r = 0; r < 3; r++)
      o[r] = Tf[4 * r + 0] * x + Tf[4 * r + 1] * y + Tf[4 * r + 2] * z + Tf[4 * r + 3];
    out[3 * i] = o[0];
    out[3 * i + 1] = o[1];
    out[3 * i + 2] = o[2];
  }
}

void oracle_correspondences(const float *src, int64_t ns, const float *tgt, int64_t nt,
                            const double T[16], double max_corr_dist, int32_t *idx,
                            float *d2) {
  float *q = (float *)malloc(sizeof(float) * 3 * (size_t)(ns > 0 ? ns : 1));
  oracle_transform_points(src, ns, T, q);
  oracle_nn(tgt, nt, q, ns, max_corr_dist * max_corr_dist, idx, d2);
  free(q);
}

static double fitness_tree(const okd_tree *t, const float *q, int64_t ns, double max_range) {
  double sum = 0.0;
  int64_t nr = 0;
  for (int64_t i = 0; i < ns; i++) {
    int32_t id;
    float d;
    if (!okd_nn(t, q + 3 * i, INFINITY, &id, &d)) continue;
    if ((double)d <= max_range) {
      sum += (double)d;
      nr++;
    }
  }
  return nr > 0 ? sum / (double)nr : DBL_MAX;
}

double oracle_fitness(const float *src, int64_t ns, const float *tgt, int64_t nt,
                      const double T[16], double max_range, int mode) {
  okd_tree *t = okd_build(tgt, nt, 3);
  float *q = (float *)malloc(sizeof(float) * 3 * (size_t)(ns > 0 ? ns : 1));
  if (mode == LCR_OR_PCL32) {
    float Tf[16];
    for (int k = 0; k < 16; k++) Tf[k] = (float)T[k];
    transform_points_f32(src, ns, Tf, q);
  } else {
    oracle_transform_points(src, ns, T, q);
  }
  double f = fitness_tree(t, q, ns, max_range);
  free(q);
  okd_free(t);
  return f;
}

/* ------------------------------------------------------------------------- */
/* IterativeClosestPoint::computeTransformation (icp.hpp) with the reference's
 * settings (fine_registration.cpp:112-121). */
int oracle_icp(const oicp_params *p, const float *src, int64_t ns, const float *tgt,
               const float *tgt_nrm, int64_t nt, const double *guess, oicp_result *out) {
  memset(out, 0, sizeof *out);
  double final_T[16];
  if (guess)
    memcpy(final_T, guess, sizeof final_T);
  else
    mat4_identity(final_T);
  memcpy(out->T, final_T, sizeof final_T);
  out->state = OR_CONV_NOT_CONVERGED;
  out->fitness = DBL_MAX;
  if (ns <= 0 || nt <= 0) return -1; /* initCompute fails: nothing runs */
  if (p->point_to_plane && !tgt_nrm) return -2;

  const int pcl32 = p->mode == LCR_OR_PCL32;
  const double r2 = p->max_corr_dist * p->max_corr_dist;
  okd_tree *tree = okd_build(tgt, nt, 3);
  float *q = (float *)malloc(sizeof(float) * 3 * (size_t)ns);
  int32_t *cidx = (int32_t *)malloc(sizeof(int32_t) * (size_t)ns);
  float *cd2 = (float *)malloc(sizeof(float) * (size_t)ns);
  float final_f[16];
  for (int k = 0; k < 16; k++) final_f[k] = (float)final_T[k];
  if (pcl32) transform_points_f32(src, ns, final_f, q);

  oconv cv;
  oconv_init(&cv, p->max_iterations, p->transformation_epsilon, p->euclidean_fitness_epsilon);
  int iters = 0, converged = 0;
  int64_t ncorr = 0;
  double mse = 0.0;
  for (;;) {
    if (!pcl32) oracle_transform_points(src, ns, final_T, q);
    ncorr = 0;
    for (int64_t i = 0; i < ns; i++) {
      okd_nn(tree, q + 3 * i, r2, cidx + i, cd2 + i);
      if (cidx[i] >= 0) ncorr++;
    }
    if (ncorr < 3) {
      cv.state = OR_CONV_NO_CORRESPONDENCES;
      converged = 0;
      break;
    }
    double Tinc[16];
    if (p->point_to_plane) {
      double AtA[36] = {0}, Atb[6] = {0};
      for (int64_t i = 0; i < ns; i++) {
        if (cidx[i] < 0) continue;
        const float *s = q + 3 * i, *d = tgt + 3 * (int64_t)cidx[i],
                    *m = tgt_nrm + 3 * (int64_t)cidx[i];
        if (!isfinite(m[0]) || !isfinite(m[1]) || !isfinite(m[2])) continue;
        if (pcl32) {
          /* PCL evaluates a, b, c and d in float, accumulates in double */
          float a = m[2] * s[1] - m[1] * s[2];
          float b = m[0] * s[2] - m[2] * s[0];
          float c = m[1] * s[0] - m[0] * s[1];
          float dd = m[0] * d[0] + m[1] * d[1] + m[2] * d[2] - m[0] * s[0] - m[1] * s[1] -
                     m[2] * s[2];
          double row[6] = {a, b, c, m[0], m[1], m[2]};
          for (int u = 0; u < 6; u++) {
            for (int v = u; v < 6; v++) AtA[6 * u + v] += row[u] * row[v];
            Atb[u] += row[u] * (double)dd;
          }
        } else {
          p2pl_accumulate(s[0], s[1], s[2], d[0], d[1], d[2], m[0], m[1], m[2], AtA, Atb);
        }
      }
      p2pl_solve(AtA, Atb, Tinc);
    } else if (pcl32) {
      /* Eigen::umeyama<float>: float means and float cross-covariance */
      float ms[3] = {0, 0, 0}, md[3] = {0, 0, 0};
      for (int64_t i = 0; i < ns; i++) {
        if (cidx[i] < 0) continue;
        for (int a = 0; a < 3; a++) {
          ms[a] += q[3 * i + a];
          md[a] += tgt[3 * (int64_t)cidx[i] + a];
        }
      }
      float inv = 1.0f / (float)ncorr;
      for (int a = 0; a < 3; a++) {
        ms[a] *= inv;
        md[a] *= inv;
      }
      float Sf[9] = {0};
      for (int64_t i = 0; i < ns; i++) {
        if (cidx[i] < 0) continue;
        float ds[3], dd[3];
        for (int a = 0; a < 3; a++) {
          ds[a] = q[3 * i + a] - ms[a];
          dd[a] = tgt[3 * (int64_t)cidx[i] + a] - md[a];
        }
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) Sf[3 * r + c] += dd[r] * ds[c];
      }
      double Sig[9], mus[3], mud[3];
      for (int k = 0; k < 9; k++) Sig[k] = (double)(Sf[k] * inv);
      for (int a = 0; a < 3; a++) {
        mus[a] = ms[a];
        mud[a] = md[a];
      }
      umeyama_from_moments(mus, mud, Sig, Tinc);
    } else {
      double mus[3] = {0, 0, 0}, mud[3] = {0, 0, 0};
      for (int64_t i = 0; i < ns; i++) {
        if (cidx[i] < 0) continue;
        for (int a = 0; a < 3; a++) {
          mus[a] += (double)q[3 * i + a];
          mud[a] += (double)tgt[3 * (int64_t)cidx[i] + a];
        }
      }
      for (int a = 0; a < 3; a++) {
        mus[a] /= (double)ncorr;
        mud[a] /= (double)ncorr;
      }
      double Sig[9] = {0};
      for (int64_t i = 0; i < ns; i++) {
        if (cidx[i] < 0) continue;
        double ds[3], dd[3];
        for (int a = 0; a < 3; a++) {
          ds[a] = (double)q[3 * i + a] - mus[a];
          dd[a] = (double)tgt[3 * (int64_t)cidx[i] + a] - mud[a];
        }
        for (int r = 0; r < 3; r++)
          for (int c = 0; c < 3; c++) Sig[3 * r + c] += dd[r] * ds[c];
      }
      for (int k = 0; k < 9; k++) Sig[k] /= (double)ncorr;
      umeyama_from_moments(mus, mud, Sig, Tinc);
    }
    if (pcl32) {
      float Tf[16];
      for (int k = 0; k < 16; k++) Tf[k] = (float)Tinc[k];
      transform_points_f32(q, ns, Tf, q);
      mat4f_mul(Tf, final_f, final_f);
      for (int k = 0; k < 16; k++) final_T[k] = final_f[k];
    } else {
      mat4_mul(Tinc, final_T, final_T);
    }
    iters++;
    mse = 0.0;
    for (int64_t i = 0; i < ns; i++)
      if (cidx[i] >= 0) mse += (double)cd2[i];
    mse /= (double)ncorr;
    if (oconv_step(&cv, iters, Tinc, mse)) {
      converged = 1;
      break;
    }
  }
  memcpy(out->T, final_T, sizeof final_T);
  out->iterations = iters;
  out->converged = converged;
  out->state = cv.state;
  out->last_ncorr = ncorr;
  out->last_mse = mse;
  /* getFitnessScore(): the original source transformed once by the final matrix */
  if (pcl32) {
    float Tf[16];
    for (int k = 0; k < 16; k++) Tf[k] = (float)final_T[k];
    transform_points_f32(src, ns, Tf, q);
  } else {
    oracle_transform_points(src, ns, final_T, q);
  }
  out->fitness = fitness_tree(tree, q, ns, DBL_MAX);
  free(q);
  free(cidx);
  free(cd2);
  okd_free(tree);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* host threading helpers */

typedef struct {
  const oicp_params *p;
  const float *const *pts;
  const float *const *nrm;
  const int64_t *n;
  int npairs;
  const int32_t *src, *tgt;
  const double *guesses;
  oicp_result *out;
  int next; /* atomic work counter */
} batch_ctx;

static void *batch_worker(void *arg) {
  batch_ctx *c = (batch_ctx *)arg;
  for (;;) {
    int i = __atomic_fetch_add(&c->next, 1, __ATOMIC_RELAXED);
    if (i >= c->npairs) break;
    int s = c->src[i], t = c->tgt[i];
    oracle_icp(c->p, c->pts[s], c->n[s], c->pts[t], c->nrm ? c->nrm[t] : NULL, c->n[t],
               c->guesses ? c->guesses + 16 * (int64_t)i : NULL, c->out + i);
  }
  return NULL;
}

int oracle_icp_batch(const oicp_params *p, int nclouds, const float *const *pts,
                     const float *const *nrm, const int64_t *n, int npairs,
                     const int32_t *src, const int32_t *tgt, const double *guesses,
                     int nthreads, oicp_result *out) {
  (void)nclouds;
  batch_ctx c = {p, pts, nrm, n, npairs, src, tgt, guesses, out, 0};
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, batch_worker, &c);
  for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* NormalEstimation::computeFeature (normal_3d.hpp): k nearest neighbours
 * including the point itself (pcl::search::KdTree, normal_estimation.cpp:89-96),
 * covariance of the neighbourhood, the eigenvector of the smallest eigenvalue
 * (solvePlaneParameters), curvature |l0|/(l0+l1+l2), flipNormalTowardsViewpoint;
 * then the CLI negates all normals when asked (normal_estimation.cpp:112-118). */

typedef struct {
  const float *xyz;
  int64_t n;
  int k;
  const float *vp;
  int flip_all;
  float *nrm, *curv;
  const okd_tree *tree;
  int64_t next;
} normals_ctx;

static void normal_one(const normals_ctx *c, int64_t i, int32_t *ids, float *ds) {
  const float *p = c->xyz + 3 * i;
  int cnt = okd_knn(c->tree, p, c->k, ids, ds);
  float *o = c->nrm + 3 * i;
  if (cnt < 3) {
    o[0] = o[1] = o[2] = NAN;
    c->curv[i] = NAN;
    return;
  }
  double mu[3] = {0, 0, 0};
  for (int j = 0; j < cnt; j++)
    for (int a = 0; a < 3; a++) mu[a] += (double)c->xyz[3 * (int64_t)ids[j] + a];
  for (int a = 0; a < 3; a++) mu[a] /= cnt;
  double C[9] = {0};
  for (int j = 0; j < cnt; j++) {
    double d[3];
    for (int a = 0; a < 3; a++) d[a] = (double)c->xyz[3 * (int64_t)ids[j] + a] - mu[a];
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++) C[3 * r + s] += d[r] * d[s];
  }
  for (int k = 0; k < 9; k++) C[k] /= cnt;
  double w[3], V[9];
  oracle_eigh3(C, w, V);
  double nx = V[0], ny = V[3], nz = V[6];
  double sum = w[0] + w[1] + w[2];
  double cv = sum != 0.0 ? fabs(w[0] / sum) : 0.0;
  double dot = ((double)c->vp[0] - p[0]) * nx + ((double)c->vp[1] - p[1]) * ny +
               ((double)c->vp[2] - p[2]) * nz;
  if (dot < 0) {
    nx = -nx;
    ny = -ny;
    nz = -nz;
  }
  if (c->flip_all) {
    nx = -nx;
    ny = -ny;
    nz = -nz;
  }
  o[0] = (float)nx;
  o[1] = (float)ny;
  o[2] = (float)nz;
  c->curv[i] = (float)cv;
}

static void *normals_worker(void *arg) {
  normals_ctx *c = (normals_ctx *)arg;
  int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * (size_t)c->k);
  float *ds = (float *)malloc(sizeof(float) * (size_t)c->k);
  for (;;) {
    int64_t b = __atomic_fetch_add(&c->next, 256, __ATOMIC_RELAXED);
    if (b >= c->n) break;
    int64_t e = b + 256 < c->n ? b + 256 : c->n;
    for (int64_t i = b; i < e; i++) normal_one(c, i, ids, ds);
  }
  free(ids);
  free(ds);
  return NULL;
}

void oracle_normals(const float *xyz, int64_t n, int k, const float vp[3], int flip_all,
                    float *nrm, float *curv, int nthreads) {
  okd_tree *t = okd_build(xyz, n, 3);
  normals_ctx c = {xyz, n, k < 1 ? 1 : k, vp, flip_all, nrm, curv, t, 0};
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, normals_worker, &c);
  for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
  okd_free(t);
}
