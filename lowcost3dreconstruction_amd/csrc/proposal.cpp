// proposal.cpp -- host side of the learned initial-pose proposal (K9): weights,
// device-resident training set, forward / backward / Adam on the GPU, and the
// per-pair inference the batched cycle uses as an extra pose hypothesis.
#include <float.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <random>
#include <vector>

#include "../../include/lcr.h"
#include "lcr_kernels.hpp"
#include "runtime.hpp"

namespace lcr {

namespace {

// ---- device helpers local to the proposal

__global__ void k_gather_batch(const float4* __restrict__ ds, const float* __restrict__ ds_lab,
                               const int* __restrict__ idx, int npts, int batch,
                               float4* __restrict__ X, float* __restrict__ lab) {
  const int64_t per_pair = 2 * (int64_t)npts;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (int64_t)batch * per_pair) {
    const int b = (int)(i / per_pair);
    const int64_t o = i % per_pair;
    X[i] = ds[(int64_t)idx[b] * per_pair + o];
  }
  if (i < (int64_t)batch * 6) lab[i] = ds_lab[(int64_t)idx[i / 6] * 6 + i % 6];
}

__global__ void k_sum_loss(const float* __restrict__ part, int n, float* __restrict__ out) {
  __shared__ float s[256];
  float a = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) a += part[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// Inputs of the network for every pair of a cycle: n_out points subsampled evenly
// from each raw cloud, the source moved by the pair's seed T0, both normalised by
// the target sample's centroid and RMS radius.  One block per pair; the sums are
// folded by a fixed LDS tree, so the frame does not depend on timing.
constexpr int kPrepThreads = 1024;

// Sample i of n_out taken evenly from a cloud of n points: floor(i * n / n_out),
// in 32-bit arithmetic when the product fits.
__device__ __forceinline__ int64_t sample_at(int i, int64_t n, int n_out) {
  if (n < ((int64_t)1 << 31) / n_out) return (int64_t)((unsigned)(i * (int)n) / (unsigned)n_out);
  return (int64_t)i * n / n_out;
}

template <int N>
__device__ __forceinline__ void prep_sum(double (&v)[N], double (*red)[N]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < N; ++k) red[t][k] = v[k];
  __syncthreads();
  for (int o = kPrepThreads / 2; o > 0; o >>= 1) {
    if (t < o) {
#pragma unroll
      for (int k = 0; k < N; ++k) red[t][k] += red[t + o][k];
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < N; ++k) v[k] = red[0][k];
  __syncthreads();  // red is reused
}

__global__ void __launch_bounds__(kPrepThreads)
k_prop_prepare(const float4* __restrict__ raw, const int64_t* __restrict__ cloud_off,
               const int* __restrict__ src, const int* __restrict__ tgt,
               const double* __restrict__ guesses, int guess_stride, int n_out,
               float4* __restrict__ X, float* __restrict__ frame) {
  __shared__ double red[kPrepThreads][4];
  const int pr = blockIdx.x, t = threadIdx.x;
  const int64_t s0 = cloud_off[src[pr]], ns = cloud_off[src[pr] + 1] - s0;
  const int64_t t0 = cloud_off[tgt[pr]], nt = cloud_off[tgt[pr] + 1] - t0;
  float4* Xs = X + (int64_t)pr * 2 * n_out;
  float4* Xt = Xs + n_out;
  double a[4] = {0, 0, 0, 0};
  for (int i = t; i < n_out; i += kPrepThreads) {
    const float4 q = raw[t0 + sample_at(i, nt, n_out)];
    if (isfinite(q.x) && isfinite(q.y) && isfinite(q.z)) {
      a[0] += q.x;
      a[1] += q.y;
      a[2] += q.z;
      a[3] += 1.0;
    }
  }
  prep_sum<4>(a, red);
  const double n = a[3] > 0 ? a[3] : 1.0;
  const double cx = a[0] / n, cy = a[1] / n, cz = a[2] / n;
  double b[4] = {0, 0, 0, 0};
  for (int i = t; i < n_out; i += kPrepThreads) {
    const float4 q = raw[t0 + sample_at(i, nt, n_out)];
    if (isfinite(q.x) && isfinite(q.y) && isfinite(q.z)) {
      const double dx = q.x - cx, dy = q.y - cy, dz = q.z - cz;
      b[0] += dx * dx + dy * dy + dz * dz;
      b[1] += 1.0;
    }
  }
  prep_sum<4>(b, red);
  const double rr = b[1] > 0 ? sqrt(b[0] / b[1]) : 1.0;
  const double rad = rr > 1e-12 ? rr : 1.0;
  if (t == 0) {
    frame[pr * 4 + 0] = (float)cx;
    frame[pr * 4 + 1] = (float)cy;
    frame[pr * 4 + 2] = (float)cz;
    frame[pr * 4 + 3] = (float)rad;
  }
  const double* T = guesses + (int64_t)pr * guess_stride * 16;
  const float inv = (float)(1.0 / rad);
  const float c0 = (float)cx, c1 = (float)cy, c2 = (float)cz;
  for (int i = t; i < n_out; i += kPrepThreads) {
    const float4 q = raw[t0 + sample_at(i, nt, n_out)];
    Xt[i] = (isfinite(q.x) && isfinite(q.y) && isfinite(q.z))
                ? make_float4((q.x - c0) * inv, (q.y - c1) * inv, (q.z - c2) * inv, 0.f)
                : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 p = raw[s0 + sample_at(i, ns, n_out)];
    if (isfinite(p.x) && isfinite(p.y) && isfinite(p.z)) {
      float m[3];
      for (int r = 0; r < 3; ++r)
        m[r] = (float)(T[4 * r] * p.x + T[4 * r + 1] * p.y + T[4 * r + 2] * p.z + T[4 * r + 3]);
      Xs[i] = make_float4((m[0] - c0) * inv, (m[1] - c1) * inv, (m[2] - c2) * inv, 0.f);
    } else {
      Xs[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
}

// learned guess = [R(omega) | s tau + c - R c] * T0, written to the pair's learned slot
__global__ void k_prop_guess(const float* __restrict__ out, const float* __restrict__ frame,
                             int npairs, int guess_stride, int slot, double* __restrict__ guesses) {
  const int pr = blockIdx.x * blockDim.x + threadIdx.x;
  if (pr >= npairs) return;
  const double w0 = out[pr * 6], w1 = out[pr * 6 + 1], w2 = out[pr * 6 + 2];
  const double th = sqrt(w0 * w0 + w1 * w1 + w2 * w2);
  double R[9];
  if (th < 1e-12) {
    for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? 1.0 : 0.0;
  } else {
    const double kx = w0 / th, ky = w1 / th, kz = w2 / th, s = sin(th), c = cos(th), v = 1.0 - c;
    R[0] = c + kx * kx * v;
    R[1] = kx * ky * v - kz * s;
    R[2] = kx * kz * v + ky * s;
    R[3] = ky * kx * v + kz * s;
    R[4] = c + ky * ky * v;
    R[5] = ky * kz * v - kx * s;
    R[6] = kz * kx * v - ky * s;
    R[7] = kz * ky * v + kx * s;
    R[8] = c + kz * kz * v;
  }
  const double cx = frame[pr * 4], cy = frame[pr * 4 + 1], cz = frame[pr * 4 + 2], sg = frame[pr * 4 + 3];
  double D[16];
  for (int r = 0; r < 3; ++r) {
    for (int k = 0; k < 3; ++k) D[4 * r + k] = R[3 * r + k];
    D[4 * r + 3] = sg * out[pr * 6 + 3 + r] + (r == 0 ? cx : (r == 1 ? cy : cz)) -
                   (R[3 * r] * cx + R[3 * r + 1] * cy + R[3 * r + 2] * cz);
  }
  D[12] = D[13] = D[14] = 0.0;
  D[15] = 1.0;
  const double* T0 = guesses + (int64_t)pr * guess_stride * 16;
  double* G = guesses + ((int64_t)pr * guess_stride + slot) * 16;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += D[4 * i + k] * T0[4 * k + j];
      G[4 * i + j] = s;
    }
}

}  // namespace

struct ProposalImpl {
  int device = 0;
  hipStream_t s = nullptr;
  int cap_pairs = 0, cap_pts = 0;
  int step = 0;
  DevBuf W, M, V, G, W2t, W3t;
  DevBuf X, part_max, part_arg, F, ARG, Z, OUT, DOUT, loss_part, DZ, DG, XA, H1A, DA1, H2A, DA2,
      DF, part_grad, lab, idx, losses, frame;
  DevBuf ds_X, ds_lab;
  int ds_pairs = 0, ds_pts = 0;
  HostBuf h_io;

  ProposalImpl() {
    LCR_HIP(hipGetDevice(&device));
    LCR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    W.reserve(kEncNParam * sizeof(float));
    M.reserve(kEncNParam * sizeof(float));
    V.reserve(kEncNParam * sizeof(float));
    G.reserve(kEncNParam * sizeof(float));
    W2t.reserve(64 * 128 * sizeof(__bf16));
    W3t.reserve(128 * 256 * sizeof(__bf16));
    LCR_HIP(hipMemsetAsync(W.p, 0, kEncNParam * sizeof(float), s));
    reset_optimizer();
  }
  ~ProposalImpl() {
    if (s) (void)hipStreamDestroy(s);
  }
  void bind() const { LCR_HIP(hipSetDevice(device)); }
  void reset_optimizer() {
    LCR_HIP(hipMemsetAsync(M.p, 0, kEncNParam * sizeof(float), s));
    LCR_HIP(hipMemsetAsync(V.p, 0, kEncNParam * sizeof(float), s));
    step = 0;
  }
  void reserve(int npairs, int npts) {
    if (npairs <= cap_pairs && npts <= cap_pts) return;
    cap_pairs = std::max(cap_pairs, npairs);
    cap_pts = std::max(cap_pts, npts);
    const int64_t P = cap_pairs, C = 2 * P, N = cap_pts;
    const int64_t chunks = enc_forward_chunks((int)N);
    X.reserve(C * N * sizeof(float4));
    part_max.reserve(C * chunks * 256 * sizeof(float));
    part_arg.reserve(C * chunks * 256 * sizeof(int));
    F.reserve(C * 256 * sizeof(float));
    ARG.reserve(C * 256 * sizeof(int));
    Z.reserve(P * 256 * sizeof(float));
    OUT.reserve(P * 6 * sizeof(float));
    DOUT.reserve(P * 6 * sizeof(float));
    loss_part.reserve(P * 6 * sizeof(float));
    DZ.reserve(P * 256 * sizeof(float));
    DG.reserve(P * 512 * sizeof(float));
    XA.reserve(C * 256 * 3 * sizeof(float));
    H1A.reserve(C * 256 * 64 * sizeof(float));
    DA1.reserve(C * 256 * 64 * sizeof(float));
    H2A.reserve(C * 256 * 128 * sizeof(float));
    DA2.reserve(C * 256 * 128 * sizeof(float));
    DF.reserve(C * 256 * sizeof(float));
    part_grad.reserve(enc_part_grad_floats((int)P) * sizeof(float));
    lab.reserve(P * 6 * sizeof(float));
    idx.reserve(P * sizeof(int));
  }
  EncBuffers bufs() const {
    EncBuffers e;
    e.X = X.as<float4>();
    e.W = W.as<float>();
    e.W2t = W2t.as<__bf16>();
    e.W3t = W3t.as<__bf16>();
    e.part_max = part_max.as<float>();
    e.part_arg = part_arg.as<int>();
    e.F = F.as<float>();
    e.ARG = ARG.as<int>();
    e.Z = Z.as<float>();
    e.OUT = OUT.as<float>();
    e.DOUT = DOUT.as<float>();
    e.loss_part = loss_part.as<float>();
    e.DZ = DZ.as<float>();
    e.DG = DG.as<float>();
    e.XA = XA.as<float>();
    e.H1A = H1A.as<float>();
    e.DA1 = DA1.as<float>();
    e.H2A = H2A.as<float>();
    e.DA2 = DA2.as<float>();
    e.DF = DF.as<float>();
    e.part_grad = part_grad.as<float>();
    e.G = G.as<float>();
    return e;
  }
  // host clouds (npairs * npts * 3 floats each) -> X as (src0, tgt0, src1, tgt1, ...)
  void upload_pairs(const float* src, const float* tgt, int npairs, int npts) {
    reserve(npairs, npts);
    const int64_t n = (int64_t)npairs * 2 * npts;
    h_io.reserve((size_t)n * sizeof(float4));
    float4* h = h_io.as<float4>();
    for (int b = 0; b < npairs; ++b)
      for (int side = 0; side < 2; ++side) {
        const float* in = (side == 0 ? src : tgt) + (int64_t)b * npts * 3;
        float4* o = h + ((int64_t)b * 2 + side) * npts;
        for (int p = 0; p < npts; ++p) o[p] = make_float4(in[3 * p], in[3 * p + 1], in[3 * p + 2], 0.f);
      }
    LCR_HIP(hipMemcpyAsync(X.p, h, (size_t)n * sizeof(float4), hipMemcpyHostToDevice, s));
    LCR_HIP(hipStreamSynchronize(s));
  }
};

// Runs the network on every pair of a cycle (on the cycle's stream) and writes each
// pair's learned guess into its slot `slot` of the instance guesses (instances
// are pair-major with `stride` slots; slot 0 holds the plain seed).
void proposal_guesses(ProposalImpl& P, hipStream_t s, const CloudSet& cs, const int* d_src,
                      const int* d_tgt, int npairs, int npts, double* d_guess, int stride, int slot) {
  P.reserve(npairs, npts);
  P.frame.reserve((size_t)npairs * 4 * sizeof(float));
  hipLaunchKernelGGL(k_prop_prepare, dim3(npairs), dim3(kPrepThreads), 0, s, cs.d_raw.as<float4>(),
                     cs.d_cloud_off.as<int64_t>(), d_src, d_tgt, d_guess, stride, npts,
                     P.X.as<float4>(), P.frame.as<float>());
  launch_enc_forward(P.bufs(), npairs, npts, nullptr, s);
  hipLaunchKernelGGL(k_prop_guess, dim3((npairs + 63) / 64), dim3(64), 0, s, P.OUT.as<float>(),
                     P.frame.as<float>(), npairs, stride, slot, d_guess);
}

}  // namespace lcr

using namespace lcr;

struct lcr_proposal {
  ProposalImpl impl;
};

lcr::ProposalImpl* lcr::proposal_impl(lcr_proposal* p) { return p ? &p->impl : nullptr; }

static inline int lcr_set_error_(int code, const char* msg) { return set_last_error(code, msg); }

template <class F>
static int pguard(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    return lcr_set_error_(e.code, e.msg.c_str());
  } catch (const std::bad_alloc&) {
    return lcr_set_error_(LCR_E_HIP, "out of memory");
  }
}

extern "C" {

int lcr_proposal_create(lcr_proposal** out) {
  return pguard([&]() -> int {
    if (!out) return lcr_set_error_(LCR_E_ARG, "out is NULL");
    *out = new lcr_proposal();
    return LCR_OK;
  });
}

void lcr_proposal_destroy(lcr_proposal* p) {
  if (!p) return;
  try {
    p->impl.bind();
  } catch (...) {
  }
  delete p;
}

int lcr_proposal_set_weights(lcr_proposal* p, const float* w, int64_t n) {
  return pguard([&]() -> int {
    if (!p || !w || n != kEncNParam) return lcr_set_error_(LCR_E_ARG, "expected 174470 weights");
    ProposalImpl& P = p->impl;
    P.bind();
    LCR_HIP(hipMemcpyAsync(P.W.p, w, n * sizeof(float), hipMemcpyHostToDevice, P.s));
    launch_enc_bf16_weights(P.W.as<float>(), P.W2t.as<__bf16>(), P.W3t.as<__bf16>(), P.s);
    P.reset_optimizer();
    LCR_HIP(hipStreamSynchronize(P.s));
    return LCR_OK;
  });
}

int lcr_proposal_get_weights(lcr_proposal* p, float* w, int64_t n) {
  return pguard([&]() -> int {
    if (!p || !w || n != kEncNParam) return lcr_set_error_(LCR_E_ARG, "expected 174470 weights");
    ProposalImpl& P = p->impl;
    P.bind();
    LCR_HIP(hipMemcpyAsync(w, P.W.p, n * sizeof(float), hipMemcpyDeviceToHost, P.s));
    LCR_HIP(hipStreamSynchronize(P.s));
    return LCR_OK;
  });
}

int lcr_proposal_forward(lcr_proposal* p, const float* src, const float* tgt, int32_t n_pairs,
                         int32_t n_points, float* out) {
  return pguard([&]() -> int {
    if (!p || !src || !tgt || !out || n_pairs < 1 || n_points < 1)
      return lcr_set_error_(LCR_E_ARG, "invalid arguments");
    ProposalImpl& P = p->impl;
    P.bind();
    P.upload_pairs(src, tgt, n_pairs, n_points);
    launch_enc_forward(P.bufs(), n_pairs, n_points, nullptr, P.s);
    LCR_HIP(hipMemcpyAsync(out, P.OUT.p, (size_t)n_pairs * 6 * sizeof(float), hipMemcpyDeviceToHost,
                           P.s));
    LCR_HIP(hipStreamSynchronize(P.s));
    return LCR_OK;
  });
}

int lcr_proposal_loss_grad(lcr_proposal* p, const float* src, const float* tgt, const float* labels,
                           int32_t n_pairs, int32_t n_points, float* loss, float* grad) {
  return pguard([&]() -> int {
    if (!p || !src || !tgt || !labels || n_pairs < 1 || n_points < 1)
      return lcr_set_error_(LCR_E_ARG, "invalid arguments");
    ProposalImpl& P = p->impl;
    P.bind();
    P.upload_pairs(src, tgt, n_pairs, n_points);
    LCR_HIP(hipMemcpyAsync(P.lab.p, labels, (size_t)n_pairs * 6 * sizeof(float),
                           hipMemcpyHostToDevice, P.s));
    const EncBuffers e = P.bufs();
    launch_enc_forward(e, n_pairs, n_points, P.lab.as<float>(), P.s);
    launch_enc_backward(e, n_pairs, n_points, P.s);
    std::vector<float> lp((size_t)n_pairs * 6);
    LCR_HIP(hipMemcpyAsync(lp.data(), P.loss_part.p, lp.size() * sizeof(float),
                           hipMemcpyDeviceToHost, P.s));
    if (grad)
      LCR_HIP(hipMemcpyAsync(grad, P.G.p, kEncNParam * sizeof(float), hipMemcpyDeviceToHost, P.s));
    LCR_HIP(hipStreamSynchronize(P.s));
    if (loss) {
      double s = 0.0;
      for (float v : lp) s += v;
      *loss = (float)s;
    }
    return LCR_OK;
  });
}

int lcr_proposal_dataset(lcr_proposal* p, const float* src, const float* tgt, const float* labels,
                         int32_t n_pairs, int32_t n_points) {
  return pguard([&]() -> int {
    if (!p || !src || !tgt || !labels || n_pairs < 1 || n_points < 1)
      return lcr_set_error_(LCR_E_ARG, "invalid arguments");
    ProposalImpl& P = p->impl;
    P.bind();
    const int64_t per = 2 * (int64_t)n_points;
    P.ds_X.reserve((size_t)n_pairs * per * sizeof(float4));
    P.ds_lab.reserve((size_t)n_pairs * 6 * sizeof(float));
    // upload in slices through the pinned staging buffer
    const int slice = std::max(1, (int)std::min<int64_t>(n_pairs, (256ll << 20) / (per * 16)));
    std::vector<float4> tmp;
    for (int b0 = 0; b0 < n_pairs; b0 += slice) {
      const int nb = std::min(slice, n_pairs - b0);
      P.h_io.reserve((size_t)nb * per * sizeof(float4));
      float4* h = P.h_io.as<float4>();
      for (int b = 0; b < nb; ++b)
        for (int side = 0; side < 2; ++side) {
          const float* in = (side == 0 ? src : tgt) + (int64_t)(b0 + b) * n_points * 3;
          float4* o = h + ((int64_t)b * 2 + side) * n_points;
          for (int q = 0; q < n_points; ++q)
            o[q] = make_float4(in[3 * q], in[3 * q + 1], in[3 * q + 2], 0.f);
        }
      LCR_HIP(hipMemcpyAsync(P.ds_X.as<float4>() + (int64_t)b0 * per, h, (size_t)nb * per * 16,
                             hipMemcpyHostToDevice, P.s));
      LCR_HIP(hipStreamSynchronize(P.s));
    }
    LCR_HIP(hipMemcpyAsync(P.ds_lab.p, labels, (size_t)n_pairs * 6 * sizeof(float),
                           hipMemcpyHostToDevice, P.s));
    LCR_HIP(hipStreamSynchronize(P.s));
    P.ds_pairs = n_pairs;
    P.ds_pts = n_points;
    return LCR_OK;
  });
}

int lcr_proposal_synthesize(lcr_proposal* p, const float* shapes, const float* weights,
                            const double* perturb, int32_t n_pairs, int32_t n_points, float noise,
                            uint64_t seed, float* raw, float* frame) {
  return pguard([&]() -> int {
    if (!p || !shapes || !weights || !perturb || n_pairs < 1 || n_points < 1)
      return lcr_set_error_(LCR_E_ARG, "invalid arguments");
    ProposalImpl& P = p->impl;
    P.bind();
    const int64_t per = 2 * (int64_t)n_points;
    P.ds_X.reserve((size_t)n_pairs * per * sizeof(float4));
    P.ds_lab.reserve((size_t)n_pairs * 6 * sizeof(float));
    std::vector<double> Pinv((size_t)n_pairs * 16);
    for (int b = 0; b < n_pairs; ++b) {
      const double* M = perturb + 16 * (int64_t)b;
      double* I = &Pinv[16 * (size_t)b];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) I[4 * r + c] = M[4 * c + r];  // R^T
      for (int r = 0; r < 3; ++r)
        I[4 * r + 3] = -(I[4 * r] * M[3] + I[4 * r + 1] * M[7] + I[4 * r + 2] * M[11]);
      I[12] = I[13] = I[14] = 0.0;
      I[15] = 1.0;
    }
    DevBuf d_shapes, d_w, d_P, d_Pi, d_raw, d_frame;
    d_shapes.reserve((size_t)n_pairs * 45 * sizeof(float));
    d_w.reserve((size_t)n_pairs * 3 * sizeof(float));
    d_P.reserve((size_t)n_pairs * 16 * sizeof(double));
    d_Pi.reserve((size_t)n_pairs * 16 * sizeof(double));
    LCR_HIP(hipMemcpyAsync(d_shapes.p, shapes, (size_t)n_pairs * 45 * 4, hipMemcpyHostToDevice, P.s));
    LCR_HIP(hipMemcpyAsync(d_w.p, weights, (size_t)n_pairs * 3 * 4, hipMemcpyHostToDevice, P.s));
    LCR_HIP(hipMemcpyAsync(d_P.p, perturb, (size_t)n_pairs * 16 * 8, hipMemcpyHostToDevice, P.s));
    LCR_HIP(hipMemcpyAsync(d_Pi.p, Pinv.data(), (size_t)n_pairs * 16 * 8, hipMemcpyHostToDevice, P.s));
    if (raw) d_raw.reserve((size_t)n_pairs * per * 3 * sizeof(float));
    if (frame) d_frame.reserve((size_t)n_pairs * 4 * sizeof(float));
    launch_synth_pairs(d_shapes.as<float>(), d_w.as<float>(), d_P.as<double>(), d_Pi.as<double>(),
                       n_pairs, n_points, noise, seed, P.ds_X.as<float4>(), P.ds_lab.as<float>(),
                       raw ? d_raw.as<float>() : nullptr, frame ? d_frame.as<float>() : nullptr, P.s);
    if (raw)
      LCR_HIP(hipMemcpyAsync(raw, d_raw.p, (size_t)n_pairs * per * 3 * 4, hipMemcpyDeviceToHost, P.s));
    if (frame)
      LCR_HIP(hipMemcpyAsync(frame, d_frame.p, (size_t)n_pairs * 4 * 4, hipMemcpyDeviceToHost, P.s));
    LCR_HIP(hipStreamSynchronize(P.s));
    P.ds_pairs = n_pairs;
    P.ds_pts = n_points;
    return LCR_OK;
  });
}

int lcr_proposal_predict(lcr_proposal* p, int32_t first, int32_t count, float* out, float* labels) {
  return pguard([&]() -> int {
    if (!p || !out || count < 1 || first < 0) return lcr_set_error_(LCR_E_ARG, "invalid arguments");
    ProposalImpl& P = p->impl;
    if (first + count > P.ds_pairs) return lcr_set_error_(LCR_E_ARG, "range beyond the resident set");
    P.bind();
    const int chunk = 512;
    for (int b0 = 0; b0 < count; b0 += chunk) {
      const int nb = std::min(chunk, count - b0);
      P.reserve(nb, P.ds_pts);
      std::vector<int> ids(nb);
      for (int i = 0; i < nb; ++i) ids[i] = first + b0 + i;
      LCR_HIP(hipMemcpyAsync(P.idx.p, ids.data(), nb * sizeof(int), hipMemcpyHostToDevice, P.s));
      const int64_t nx = (int64_t)nb * 2 * P.ds_pts;
      hipLaunchKernelGGL(k_gather_batch, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, P.s,
                         P.ds_X.as<float4>(), P.ds_lab.as<float>(), P.idx.as<int>(), P.ds_pts, nb,
                         P.X.as<float4>(), P.lab.as<float>());
      launch_enc_forward(P.bufs(), nb, P.ds_pts, nullptr, P.s);
      LCR_HIP(hipMemcpyAsync(out + (int64_t)b0 * 6, P.OUT.p, (size_t)nb * 6 * 4, hipMemcpyDeviceToHost, P.s));
      if (labels)
        LCR_HIP(hipMemcpyAsync(labels + (int64_t)b0 * 6, P.lab.p, (size_t)nb * 6 * 4,
                               hipMemcpyDeviceToHost, P.s));
      LCR_HIP(hipStreamSynchronize(P.s));
    }
    return LCR_OK;
  });
}

int lcr_proposal_train(lcr_proposal* p, int32_t steps, int32_t batch, float lr, uint64_t seed,
                       float* losses) {
  return pguard([&]() -> int {
    if (!p || steps < 1 || batch < 1) return lcr_set_error_(LCR_E_ARG, "invalid arguments");
    ProposalImpl& P = p->impl;
    if (P.ds_pairs < 1) return lcr_set_error_(LCR_E_STATE, "no training set");
    P.bind();
    batch = std::min(batch, P.ds_pairs);
    P.reserve(batch, P.ds_pts);
    P.losses.reserve((size_t)steps * sizeof(float));
    std::mt19937_64 rng(seed);
    std::vector<int> all((size_t)steps * batch);
    std::vector<int> perm(P.ds_pairs);
    for (int i = 0; i < P.ds_pairs; ++i) perm[i] = i;
    size_t pos = perm.size();
    for (auto& v : all) {
      if (pos == perm.size()) {
        std::shuffle(perm.begin(), perm.end(), rng);
        pos = 0;
      }
      v = perm[pos++];
    }
    DevBuf d_all;
    d_all.reserve(all.size() * sizeof(int));
    LCR_HIP(hipMemcpyAsync(d_all.p, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice,
                           P.s));
    const EncBuffers e = P.bufs();
    const int64_t nx = (int64_t)batch * 2 * P.ds_pts;
    for (int st = 0; st < steps; ++st) {
      hipLaunchKernelGGL(k_gather_batch, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, P.s,
                         P.ds_X.as<float4>(), P.ds_lab.as<float>(), d_all.as<int>() + (int64_t)st * batch,
                         P.ds_pts, batch, P.X.as<float4>(), P.lab.as<float>());
      launch_enc_forward(e, batch, P.ds_pts, P.lab.as<float>(), P.s);
      hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, P.s, P.loss_part.as<float>(), batch * 6,
                         P.losses.as<float>() + st);
      launch_enc_backward(e, batch, P.ds_pts, P.s);
      P.step += 1;
      launch_adam(P.W.as<float>(), P.G.as<float>(), P.M.as<float>(), P.V.as<float>(), kEncNParam, lr,
                  P.step, P.s);
      launch_enc_bf16_weights(P.W.as<float>(), P.W2t.as<__bf16>(), P.W3t.as<__bf16>(), P.s);
    }
    if (losses)
      LCR_HIP(hipMemcpyAsync(losses, P.losses.p, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost,
                             P.s));
    LCR_HIP(hipStreamSynchronize(P.s));
    return LCR_OK;
  });
}

}  // extern "C"
