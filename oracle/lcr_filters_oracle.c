/*
 * lcr_filters_oracle.c -- CPU oracle for the pre-registration filters and the
 * depth back-projection.
 *
 * TEST INFRASTRUCTURE ONLY (see lcr_oracle.h): only tests/ load it.
 *
 * Restates, from PCL's and OpenCV's published behaviour:
 *   oracle_voxel_grid -- pcl::VoxelGrid<PointXYZRGBNormal>::filter with
 *       setLeafSize(l, l, l) (pcl_tools/cloud_downsampling.cpp:73-75;
 *       PCL filters/impl/voxel_grid.hpp, common/impl/centroid.hpp CentroidPoint)
 *   oracle_sor        -- pcl::StatisticalOutlierRemoval with setMeanK /
 *       setStddevMulThresh (pcl_tools/outlier_removal.cpp:80-84;
 *       filters/impl/statistical_outlier_removal.hpp)
 *   oracle_depth_to_cloud -- the depth tool's cross-product normals and
 *       back-projection (capture/depth_capture/depth_capture.cpp:230-305)
 *
 * Choices where PCL's result depends on unpinnable details:
 *   - VoxelGrid sorts (voxel, point) pairs with std::sort, which is not stable, so
 *     PCL's summation order inside a voxel is unspecified; here points of a voxel
 *     are summed in increasing input index.  Float sums differ by rounding only.
 *   - StatisticalOutlierRemoval adds sqrt of the float squared distances into a
 *     double; whether that sqrt runs in float or double depends on overload
 *     resolution in the PCL build.  Here it runs in double.
 * Parity of both is checked against closed forms and brute force in tests/ (regular
 * lattices, planted outliers); PCL itself is absent.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lcr_oracle.h"

typedef struct {
  int64_t key;
  int64_t idx;
} vkey;

static int cmp_vkey(const void* a, const void* b) {
  const vkey* x = (const vkey*)a;
  const vkey* y = (const vkey*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

static int finite3(const float* p) { return isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]); }

/* voxel_grid.hpp applyFilter: min/max over finite points; lattice coordinates
 * floor(p * inverse_leaf) (float) relative to floor(min * inverse_leaf); a lattice
 * of more than INT32_MAX voxels returns the input unchanged (PCL's warning path).
 * One output point per occupied voxel, in increasing voxel index (x fastest):
 * xyz mean, normal mean normalised, curvature mean, rgb per-channel mean
 * truncated.  Returns the output count. */
int64_t oracle_voxel_grid(const float* xyz, const float* nrm, const float* curv,
                          const uint8_t* rgb, int64_t n, float leaf, float* out_xyz,
                          float* out_nrm, float* out_curv, uint8_t* out_rgb) {
  const float inv = 1.0f / leaf;
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  int64_t nf = 0;
  for (int64_t i = 0; i < n; ++i) {
    const float* p = xyz + 3 * i;
    if (!finite3(p)) continue;
    ++nf;
    for (int a = 0; a < 3; ++a) {
      if (p[a] < mn[a]) mn[a] = p[a];
      if (p[a] > mx[a]) mx[a] = p[a];
    }
  }
  if (nf == 0) return 0;
  int64_t d[3];
  for (int a = 0; a < 3; ++a) d[a] = (int64_t)((mx[a] - mn[a]) * inv) + 1;
  if (d[0] * d[1] * d[2] > (int64_t)INT32_MAX) {
    /* PCL copies the input through unchanged */
    for (int64_t i = 0; i < n; ++i) {
      memcpy(out_xyz + 3 * i, xyz + 3 * i, 3 * sizeof(float));
      if (out_nrm && nrm) memcpy(out_nrm + 3 * i, nrm + 3 * i, 3 * sizeof(float));
      if (out_curv && curv) out_curv[i] = curv[i];
      if (out_rgb && rgb) memcpy(out_rgb + 3 * i, rgb + 3 * i, 3);
    }
    return n;
  }
  int minb[3], maxb[3];
  int64_t div[3];
  for (int a = 0; a < 3; ++a) {
    minb[a] = (int)floorf(mn[a] * inv);
    maxb[a] = (int)floorf(mx[a] * inv);
    div[a] = (int64_t)(maxb[a] - minb[a] + 1);
  }
  vkey* v = (vkey*)malloc((size_t)nf * sizeof(vkey));
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i) {
    const float* p = xyz + 3 * i;
    if (!finite3(p)) continue;
    int ijk[3];
    for (int a = 0; a < 3; ++a) ijk[a] = (int)(floorf(p[a] * inv) - (float)minb[a]);
    v[m].key = (int64_t)ijk[0] + (int64_t)ijk[1] * div[0] + (int64_t)ijk[2] * div[0] * div[1];
    v[m].idx = i;
    ++m;
  }
  qsort(v, (size_t)m, sizeof(vkey), cmp_vkey);
  int64_t out = 0;
  for (int64_t s = 0; s < m;) {
    int64_t e = s + 1;
    while (e < m && v[e].key == v[s].key) ++e;
    float sx = 0.f, sy = 0.f, sz = 0.f, snx = 0.f, sny = 0.f, snz = 0.f, sc = 0.f;
    float sr = 0.f, sg = 0.f, sb = 0.f;
    for (int64_t t = s; t < e; ++t) {
      const int64_t i = v[t].idx;
      sx = sx + xyz[3 * i];
      sy = sy + xyz[3 * i + 1];
      sz = sz + xyz[3 * i + 2];
      if (nrm) {
        snx = snx + nrm[3 * i];
        sny = sny + nrm[3 * i + 1];
        snz = snz + nrm[3 * i + 2];
      }
      if (curv) sc = sc + curv[i];
      if (rgb) {
        sr = sr + (float)rgb[3 * i];
        sg = sg + (float)rgb[3 * i + 1];
        sb = sb + (float)rgb[3 * i + 2];
      }
    }
    const float cnt = (float)(e - s);
    out_xyz[3 * out] = sx / cnt;
    out_xyz[3 * out + 1] = sy / cnt;
    out_xyz[3 * out + 2] = sz / cnt;
    if (out_nrm && nrm) {
      /* Eigen normalized() of the summed 4-vector (w = 0): v / sqrt(v . v) */
      float nn = snx * snx;
      nn = nn + sny * sny;
      nn = nn + snz * snz;
      const float len = sqrtf(nn);
      out_nrm[3 * out] = len > 0.f ? snx / len : snx;
      out_nrm[3 * out + 1] = len > 0.f ? sny / len : sny;
      out_nrm[3 * out + 2] = len > 0.f ? snz / len : snz;
    }
    if (out_curv && curv) out_curv[out] = sc / cnt;
    if (out_rgb && rgb) {
      out_rgb[3 * out] = (uint8_t)(uint32_t)(sr / cnt);
      out_rgb[3 * out + 1] = (uint8_t)(uint32_t)(sg / cnt);
      out_rgb[3 * out + 2] = (uint8_t)(uint32_t)(sb / cnt);
    }
    ++out;
    s = e;
  }
  free(v);
  return out;
}

/* statistical_outlier_removal.hpp applyFilterIndices: per finite point, its
 * mean_k + 1 nearest (itself included, FLANN float d2), the mean of sqrt(d2) over
 * all but the first, stored as float; non-finite points get 0 and are not counted.
 * mean and sample variance over the stored values (double sums, the square taken
 * in float), threshold = mean + mul * stddev; a point is kept when its value is
 * <= threshold (so non-finite points are kept).  keep[i] = 1/0; returns the number
 * kept.  mean_dist (nullable) receives the per-point values; stats (nullable) the
 * mean, stddev and threshold. */
int64_t oracle_sor(const float* xyz, int64_t n, int mean_k, double mul, uint8_t* keep,
                   float* mean_dist, double* stats) {
  okd_tree* t = okd_build(xyz, n, 3);
  float* dist = (float*)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
  int32_t* idx = (int32_t*)malloc((size_t)(mean_k + 1) * sizeof(int32_t));
  float* d2 = (float*)malloc((size_t)(mean_k + 1) * sizeof(float));
  int64_t valid = 0;
  for (int64_t i = 0; i < n; ++i) {
    const float* p = xyz + 3 * i;
    if (!finite3(p)) {
      dist[i] = 0.f;
      continue;
    }
    const int got = okd_knn(t, p, mean_k + 1, idx, d2);
    double s = 0.0;
    for (int j = 1; j < got; ++j) s += sqrt((double)d2[j]);
    dist[i] = (float)(s / (double)mean_k);
    ++valid;
  }
  double sum = 0.0, sq = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    sum += dist[i];
    sq += (double)(dist[i] * dist[i]);
  }
  const double mean = valid > 0 ? sum / (double)valid : 0.0;
  const double var = valid > 1 ? (sq - sum * sum / (double)valid) / ((double)valid - 1.0) : 0.0;
  const double sd = sqrt(var);
  const double thr = mean + mul * sd;
  int64_t kept = 0;
  for (int64_t i = 0; i < n; ++i) {
    keep[i] = (double)dist[i] > thr ? 0 : 1;
    kept += keep[i];
    if (mean_dist) mean_dist[i] = dist[i];
  }
  if (stats) {
    stats[0] = mean;
    stats[1] = sd;
    stats[2] = thr;
  }
  free(d2);
  free(idx);
  free(dist);
  okd_free(t);
  return kept;
}

/* depth_capture.cpp:230-305: per pixel (i row, j column) of a 16-bit depth map
 * in millimetres, neighbours "right" = (i+1, j), "left" = (i-1, j), "down" =
 * (i, j+1), "up" = (i, j-1) (0 on the border), each set to 0 when it differs
 * from the centre by more than 20; n = v * (1 / |v|) for v = (-(right-left)/2,
 * -(down-up)/2, 1) in double (cv::normalize), stored as float.  A pixel with depth z != 0, depth_min <= z <=
 * depth_max and inside the crop [top, bottom] x [left, right] becomes the point
 * (z (j - cx) (1/fx), -z (i - cy) (1/fy), -z) (double arithmetic in that order,
 * rounded to float), in row-major pixel order, with its colour (3 bytes from a
 * W x H x 3 image, copied as stored).  Returns the number of points. */
int64_t oracle_depth_to_cloud(const uint16_t* depth, int w, int h, const uint8_t* color, double fx,
                              double fy, double cx, double cy, int depth_min, int depth_max,
                              int top, int bottom, int left, int right, float* out_xyz,
                              float* out_nrm, uint8_t* out_rgb) {
  int64_t m = 0;
  for (int i = 0; i < h; ++i) {
    for (int j = 0; j < w; ++j) {
      const int z = depth[(int64_t)i * w + j];
      if (!(z != 0 && z >= depth_min && z <= depth_max && i >= top && i <= bottom && j >= left &&
            j <= right))
        continue;
      int rt = 0, lf = 0, dn = 0, up = 0;
      if (i > 0 && j > 0 && i < h - 1 && j < w - 1) {
        rt = depth[(int64_t)(i + 1) * w + j];
        lf = depth[(int64_t)(i - 1) * w + j];
        dn = depth[(int64_t)i * w + j + 1];
        up = depth[(int64_t)i * w + j - 1];
      }
      if (abs(z - rt) > 20) rt = 0;
      if (abs(z - lf) > 20) lf = 0;
      if (abs(z - dn) > 20) dn = 0;
      if (abs(z - up) > 20) up = 0;
      const double dzdx = (rt - lf) / 2.0, dzdy = (dn - up) / 2.0;
      /* cv::normalize(Vec3d): v * (1 / |v|) */
      const double vx = -dzdx, vy = -dzdy, vz = 1.0;
      const double inv = 1.0 / sqrt(vx * vx + vy * vy + vz * vz);
      const double vxp = (double)z * (j - cx) * (1 / fx);
      const double vyp = -(double)z * (i - cy) * (1 / fy);
      out_xyz[3 * m] = (float)vxp;
      out_xyz[3 * m + 1] = (float)vyp;
      out_xyz[3 * m + 2] = (float)(-z);
      if (out_nrm) {
        out_nrm[3 * m] = (float)(vx * inv);
        out_nrm[3 * m + 1] = (float)(vy * inv);
        out_nrm[3 * m + 2] = (float)(vz * inv);
      }
      if (out_rgb && color) memcpy(out_rgb + 3 * m, color + 3 * ((int64_t)i * w + j), 3);
      ++m;
    }
  }
  return m;
}

/* Lanczos-4 weights of the 8 taps around a sample point at fractional offset fr
 * (taps at -3..4 from the integer part): the windowed sinc
 * L(d) = sin(pi d) sin(pi d / 4) / (pi^2 d^2 / 4), d = fr + 3 - i, normalised to sum
 * 1 in double and rounded to float; fr == 0 gives the unit tap. */
void oracle_lanczos4_weights(double fr, float w[8]) {
  const double pi = 3.14159265358979323846;
  if (fr == 0.0) {
    for (int i = 0; i < 8; i++) w[i] = i == 3 ? 1.0f : 0.0f;
    return;
  }
  double c[8], s = 0.0;
  for (int i = 0; i < 8; i++) {
    const double d = fr + 3.0 - i;
    c[i] = sin(pi * d) * sin(pi * d / 4.0) / (pi * pi * d * d / 4.0);
    s += c[i];
  }
  for (int i = 0; i < 8; i++) w[i] = (float)(c[i] / s);
}

/* The burst fusion of capture/super_resolution/super_resolution.cpp:283-314 for
 * frames already aligned (identity registration): each frame upsampled by `factor`
 * with nearest neighbours (:283), the zero-eliminating mean per pixel of the
 * upsampled frames (:295-306: the mean of the positive values, accumulated in
 * double, NaN when there are none), the Lanczos-4 resize back to width x height
 * (:311: source position (x + 0.5) f - 0.5, 8 taps clamped to the image, the
 * horizontal pass then the vertical, in float) and the conversion to 16 bits
 * (:312: round to nearest even, clamped to [0, 65535], NaN -> 0).  frames:
 * n_frames x height x width, row-major. */
void oracle_depth_fuse(const uint16_t *frames, int n_frames, int w, int h, int factor,
                       uint16_t *out) {
  const int64_t np = (int64_t)w * h;
  float *mean = (float *)malloc(sizeof(float) * (size_t)(np > 0 ? np : 1));
  for (int64_t p = 0; p < np; p++) {
    double s = 0.0;
    int c = 0;
    for (int f = 0; f < n_frames; f++) {
      const uint16_t v = frames[(int64_t)f * np + p];
      if (v > 0) {
        s += (double)(float)v;
        c++;
      }
    }
    mean[p] = c > 0 ? (float)(s / c) : NAN;
  }
  /* the upsampled image (factor w x factor h) repeats each mean; sample it */
  const int W = w * factor, H = h * factor;
  const double pos = (factor - 1) / 2.0; /* (x + 0.5) f - 0.5 = f x + (f - 1) / 2 */
  const int base = (int)floor(pos);
  float wx[8];
  oracle_lanczos4_weights(pos - base, wx);
  for (int y = 0; y < h; y++) {
    const int sy = factor * y + base;
    for (int x = 0; x < w; x++) {
      const int sx = factor * x + base;
      float acc = 0.0f;
      for (int ky = 0; ky < 8; ky++) {
        int r = sy - 3 + ky;
        r = r < 0 ? 0 : (r >= H ? H - 1 : r);
        const float *mr = mean + (int64_t)(r / factor) * w;
        float hs = 0.0f;
        for (int kx = 0; kx < 8; kx++) {
          int q = sx - 3 + kx;
          q = q < 0 ? 0 : (q >= W ? W - 1 : q);
          hs = hs + wx[kx] * mr[q / factor];
        }
        acc = acc + wx[ky] * hs;
      }
      uint16_t v = 0;
      if (!isnan(acc)) {
        const float rv = rintf(acc);
        v = rv <= 0.0f ? 0 : (rv >= 65535.0f ? 65535 : (uint16_t)rv);
      }
      out[(int64_t)y * w + x] = v;
    }
  }
  free(mean);
}
