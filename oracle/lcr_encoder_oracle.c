/*
 * lcr_encoder_oracle.c -- FP64 CPU oracle of the learned initial-pose proposal.
 *
 * TEST INFRASTRUCTURE ONLY (see lcr_oracle.h).  The network has no counterpart in
 * the reference: it takes the role of the coarse-alignment step that precedes
 * fine registration (Super4PCS at scripts/alignment.sh:107).
 *
 * Network (one weight set shared by both clouds of a pair):
 *   per point  x (3, normalised)      a1 = x W1 + b1,  h1 = relu(a1)   (64)
 *                                     a2 = h1 W2 + b2, h2 = relu(a2)   (128)
 *                                     a3 = h2 W3 + b3, h3 = relu(a3)   (256)
 *   per cloud  f[c] = max_p h3[p][c], ties -> the lowest point index
 *   per pair   g = [f_src, f_tgt] (512), z = relu(g Hw1 + hb1) (256),
 *              out = z Hw2 + hb2 (6) = (axis-angle omega, translation tau)
 * Weights are row-major [in][out] in one flat array (ENC_* offsets below).
 *
 * quantize = 1 reproduces the device's rounding points: h1 and h2 rounded to
 * bfloat16 (round to nearest even), W2 and W3 used as bfloat16; every sum in
 * FP64.  quantize = 0 is the plain FP64 network (used for the finite-difference
 * check of the analytic backward).
 *
 * Loss per batch of B pairs: sum over pairs and components of (out - label)^2 / B.
 * Backward: exact for quantize = 0; with quantize = 1 the roundings are passed
 * through (straight-through), as on the device.
 *
 * Pinned by: finite differences of the FP64 loss (tests/test_encoder_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lcr_oracle.h"

static inline double bf16(double v) {
  float x = (float)v;
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) {
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
  }
  memcpy(&x, &u, 4);
  return (double)x;
}

typedef struct {
  const double *W1, *b1, *W2, *b2, *W3, *b3, *Hw1, *hb1, *Hw2, *hb2;
} enc_w;

static enc_w unpack(const double *w) {
  enc_w e;
  e.W1 = w + ENC_W1;
  e.b1 = w + ENC_B1;
  e.W2 = w + ENC_W2;
  e.b2 = w + ENC_B2;
  e.W3 = w + ENC_W3;
  e.b3 = w + ENC_B3;
  e.Hw1 = w + ENC_HW1;
  e.hb1 = w + ENC_HB1;
  e.Hw2 = w + ENC_HW2;
  e.hb2 = w + ENC_HB2;
  return e;
}

/* per point activations of one cloud */
static void point_forward(const enc_w *e, const float *x, int q, double *a1, double *h1,
                          double *a2, double *h2, double *a3) {
  for (int j = 0; j < 64; j++) {
    double s = e->b1[j];
    for (int i = 0; i < 3; i++) s += (double)x[i] * e->W1[i * 64 + j];
    a1[j] = s;
    h1[j] = s > 0 ? (q ? bf16(s) : s) : 0.0;
  }
  for (int j = 0; j < 128; j++) {
    double s = e->b2[j];
    for (int i = 0; i < 64; i++) s += h1[i] * (q ? bf16(e->W2[i * 128 + j]) : e->W2[i * 128 + j]);
    a2[j] = s;
    h2[j] = s > 0 ? (q ? bf16(s) : s) : 0.0;
  }
  for (int j = 0; j < 256; j++) {
    double s = e->b3[j];
    for (int i = 0; i < 128; i++) s += h2[i] * (q ? bf16(e->W3[i * 256 + j]) : e->W3[i * 256 + j]);
    a3[j] = s;
  }
}

static void cloud_forward(const enc_w *e, const float *pts, int n, int q, double *f, int *arg) {
  double a1[64], h1[64], a2[128], h2[128], a3[256];
  for (int c = 0; c < 256; c++) {
    f[c] = -1.0;
    arg[c] = 0;
  }
  for (int p = 0; p < n; p++) {
    point_forward(e, pts + 3 * (size_t)p, q, a1, h1, a2, h2, a3);
    for (int c = 0; c < 256; c++) {
      const double v = a3[c] > 0 ? a3[c] : 0.0;
      if (v > f[c]) {
        f[c] = v;
        arg[c] = p;
      }
    }
  }
}

static void head_forward(const enc_w *e, const double *g, double *z, double *out) {
  for (int j = 0; j < 256; j++) {
    double s = e->hb1[j];
    for (int i = 0; i < 512; i++) s += g[i] * e->Hw1[i * 256 + j];
    z[j] = s > 0 ? s : 0.0;
  }
  for (int j = 0; j < 6; j++) {
    double s = e->hb2[j];
    for (int i = 0; i < 256; i++) s += z[i] * e->Hw2[i * 6 + j];
    out[j] = s;
  }
}

void oenc_forward(const double *w, const float *src, const float *tgt, int n, int quantize,
                  double out[6], double feat[512], int32_t argmax[512]) {
  enc_w e = unpack(w);
  double g[512], z[256];
  int a[512];
  cloud_forward(&e, src, n, quantize, g, a);
  cloud_forward(&e, tgt, n, quantize, g + 256, a + 256);
  head_forward(&e, g, z, out);
  if (feat) memcpy(feat, g, sizeof g);
  if (argmax)
    for (int i = 0; i < 512; i++) argmax[i] = a[i];
}

/* backward of one cloud given dL/df (256); accumulates into grad */
static void cloud_backward(const enc_w *e, const float *pts, int q, const double *f,
                           const int *arg, const double *df, double *grad) {
  double a1[64], h1[64], a2[128], h2[128], a3[256];
  for (int c = 0; c < 256; c++) {
    if (df[c] == 0.0 || !(f[c] > 0.0)) continue; /* relu'(0) = 0 */
    const int p = arg[c];
    point_forward(e, pts + 3 * (size_t)p, q, a1, h1, a2, h2, a3);
    const double d3 = df[c];
    grad[ENC_B3 + c] += d3;
    double dh2[128];
    for (int i = 0; i < 128; i++) {
      grad[ENC_W3 + i * 256 + c] += h2[i] * d3;
      dh2[i] = d3 * (q ? bf16(e->W3[i * 256 + c]) : e->W3[i * 256 + c]);
    }
    double dh1[64];
    for (int i = 0; i < 64; i++) dh1[i] = 0.0;
    for (int j = 0; j < 128; j++) {
      if (!(a2[j] > 0)) continue;
      const double d2 = dh2[j];
      grad[ENC_B2 + j] += d2;
      for (int i = 0; i < 64; i++) {
        grad[ENC_W2 + i * 128 + j] += h1[i] * d2;
        dh1[i] += d2 * (q ? bf16(e->W2[i * 128 + j]) : e->W2[i * 128 + j]);
      }
    }
    const float *x = pts + 3 * (size_t)p;
    for (int j = 0; j < 64; j++) {
      if (!(a1[j] > 0)) continue;
      grad[ENC_B1 + j] += dh1[j];
      for (int i = 0; i < 3; i++) grad[ENC_W1 + i * 64 + j] += (double)x[i] * dh1[j];
    }
  }
}

double oenc_loss_grad(const double *w, const float *src, const float *tgt, int n, int batch,
                      const double *labels, int quantize, double *grad) {
  enc_w e = unpack(w);
  if (grad) memset(grad, 0, sizeof(double) * ENC_NPARAM);
  double loss = 0.0;
  for (int b = 0; b < batch; b++) {
    const float *s = src + (size_t)b * n * 3, *t = tgt + (size_t)b * n * 3;
    double g[512], z[256], out[6];
    int a[512];
    cloud_forward(&e, s, n, quantize, g, a);
    cloud_forward(&e, t, n, quantize, g + 256, a + 256);
    head_forward(&e, g, z, out);
    double dout[6];
    for (int j = 0; j < 6; j++) {
      const double d = out[j] - labels[6 * b + j];
      loss += d * d / batch;
      dout[j] = 2.0 * d / batch;
    }
    if (!grad) continue;
    double dz[256];
    for (int i = 0; i < 256; i++) {
      double s2 = 0.0;
      for (int j = 0; j < 6; j++) {
        grad[ENC_HW2 + i * 6 + j] += z[i] * dout[j];
        s2 += dout[j] * e.Hw2[i * 6 + j];
      }
      dz[i] = z[i] > 0 ? s2 : 0.0;
    }
    for (int j = 0; j < 6; j++) grad[ENC_HB2 + j] += dout[j];
    double dg[512];
    for (int i = 0; i < 512; i++) {
      double s2 = 0.0;
      for (int j = 0; j < 256; j++) {
        grad[ENC_HW1 + i * 256 + j] += g[i] * dz[j];
        s2 += dz[j] * e.Hw1[i * 256 + j];
      }
      dg[i] = s2;
    }
    for (int j = 0; j < 256; j++) grad[ENC_HB1 + j] += dz[j];
    cloud_backward(&e, s, quantize, g, a, dg, grad);
    cloud_backward(&e, t, quantize, g + 256, a + 256, dg + 256, grad);
  }
  return loss;
}
