/*
 * lcr_oracle.h -- CPU oracle for the fine-registration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it.  The product library (liblcr.so) never links,
 * loads or calls it, and has no CPU fallback.
 *
 * It restates, in plain C, the semantics of the PCL calls made by the
 * reference's fine-registration and normal-estimation tools:
 *   pcl_tools/fine_registration.cpp:105-126  (IterativeClosestPoint: setInputSource,
 *       setInputTarget, setMaxCorrespondenceDistance, setMaximumIterations,
 *       setTransformationEpsilon, setEuclideanFitnessEpsilon, align,
 *       getFinalTransformation, hasConverged, getFitnessScore)
 *   pcl_tools/normal_estimation.cpp:84-118   (NormalEstimation with setKSearch,
 *       setViewPoint / useSensorOriginAsViewPoint, and the CLI's flip rule)
 * PCL itself is not vendored in the reference and is not installed here, so the
 * PCL internals (icp.hpp, correspondence_estimation.hpp,
 * transformation_estimation_svd.hpp, transformation_estimation_point_to_plane_lls.hpp,
 * default_convergence_criteria.hpp, normal_3d.hpp, kdtree_flann.hpp, PCL 1.7.2-1.10)
 * are restated from their published behaviour; every function below names the
 * PCL routine it follows.
 *
 * Pinning: every function is checked in tests/ against numpy-generated golden
 * vectors (tests/golden/), closed forms, invariants, brute force or finite
 * differences.  Parity unpinned: oracle_icp in LCR_OR_PCL32 mode (PCL's float
 * rounding cannot be checked without PCL; it serves only as the timed host
 * baseline and agrees with LCR_OR_FP64 to 1e-3).
 *
 * Two arithmetic modes:
 *   LCR_OR_FP64  -- the FP64 oracle: the source is transformed from the original
 *                   points by the FP64-composed transform (rounded to float once),
 *                   correspondence distances are FLANN's float L2 form, all
 *                   reductions, solves and compositions are FP64.
 *   LCR_OR_PCL32 -- a PCL-semantics FP32 restatement: incremental float
 *                   transforms of the source, float Umeyama statistics, float
 *                   4x4 composition, as PCL does with Scalar=float.  Used as the
 *                   host CPU baseline ("PCL-semantics restatement, not PCL").
 */
#ifndef LCR_ORACLE_H
#define LCR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { LCR_OR_FP64 = 0, LCR_OR_PCL32 = 1 };

/* pcl::registration::DefaultConvergenceCriteria::ConvergenceState values. */
enum {
  OR_CONV_NOT_CONVERGED = 0,
  OR_CONV_ITERATIONS = 1,
  OR_CONV_TRANSFORM = 2,
  OR_CONV_ABS_MSE = 3,
  OR_CONV_REL_MSE = 4,
  OR_CONV_NO_CORRESPONDENCES = 5
};

/* ---- exact nearest-neighbour search (replaces KdTreeFLANN, kdtree_flann.hpp) ---- */
typedef struct okd_tree okd_tree;
/* xyz: n points, 3 floats each, stride `stride` floats.  Non-finite points are
 * left out of the index, as KdTreeFLANN::setInputCloud does for non-dense clouds. */
okd_tree *okd_build(const float *xyz, int64_t n, int64_t stride);
void okd_free(okd_tree *t);
/* 1-NN.  Returns 1 and fills idx/d2 if a point with d2 <= max_d2 exists, else 0.
 * Distance is FLANN L2_Simple in float: ((0+dx*dx)+dy*dy)+dz*dz, dx = q - p.
 * Ties in d2 resolve to the smallest index. */
int okd_nn(const okd_tree *t, const float q[3], double max_d2, int32_t *idx, float *d2);
/* k-NN sorted by (d2, index) ascending; returns the number found (min(k, n)). */
int okd_knn(const okd_tree *t, const float q[3], int k, int32_t *idx, float *d2);

/* brute force (same distance form and tie rule); for small inputs */
void obf_nn(const float *tgt, int64_t nt, const float *q, int64_t nq, double max_d2,
            int32_t *idx, float *d2);
int obf_knn(const float *tgt, int64_t nt, const float q[3], int k, int32_t *idx, float *d2);

/* batch NN with the tree: per query idx (-1 if none within max_d2) and d2 */
void oracle_nn(const float *tgt, int64_t nt, const float *q, int64_t nq, double max_d2,
               int32_t *idx, float *d2);

/* ---- transformation estimation ---- */
/* Umeyama (TransformationEstimationSVD, use_umeyama=true): src/dst paired, 3 doubles
 * each.  T is row-major 4x4 mapping src onto dst.  Returns 0, or -1 if n < 1. */
int oracle_umeyama(const double *src, const double *dst, int64_t n, double T[16]);
/* FP64 one-sided Jacobi SVD of a 3x3 row-major A = U diag(s) V^T, s descending. */
void oracle_svd3(const double A[9], double U[9], double s[3], double V[9]);
/* TransformationEstimationPointToPlaneLLS: src, dst, dst normals (3 doubles each).
 * Returns 0, or -1 if the 6x6 system is singular (T set to identity). */
int oracle_p2pl(const double *src, const double *dst, const double *nrm, int64_t n,
                double T[16]);
/* the S4 matrix: T = [Rz(g)Ry(b)Rx(a) | t] */
void oracle_p2pl_matrix(double a, double b, double g, double tx, double ty, double tz,
                        double T[16]);

/* ---- convergence (DefaultConvergenceCriteria::hasConverged) ---- */
typedef struct {
  int max_iterations;
  double rotation_threshold;     /* 1 - transformation_epsilon */
  double translation_threshold;  /* transformation_epsilon */
  double mse_threshold_relative; /* euclidean_fitness_epsilon */
  double mse_threshold_absolute; /* 1e-12 */
  double prev_mse;               /* starts at DBL_MAX */
  int iterations_similar;
  int max_iterations_similar;    /* 0 */
  int state;
} oconv;
void oconv_init(oconv *c, int max_it, double trans_eps, double fit_eps);
/* Tinc: this iteration's incremental transform, row-major; its coefficients are
 * tested as float (PCL's Matrix4f).  mse: mean correspondence d2 of the iteration.
 * Returns 1 when converged; c->state says why. */
int oconv_step(oconv *c, int iterations, const double Tinc[16], double mse);

/* ---- ICP (IterativeClosestPoint::computeTransformation + getFitnessScore) ---- */
typedef struct {
  double max_corr_dist;
  int max_iterations;
  double transformation_epsilon;
  double euclidean_fitness_epsilon;
  int point_to_plane;
  int mode; /* LCR_OR_FP64 or LCR_OR_PCL32 */
} oicp_params;

typedef struct {
  double T[16];
  int iterations;
  int converged;
  int state;
  double fitness; /* getFitnessScore(DBL_MAX) */
  int64_t last_ncorr;
  double last_mse;
} oicp_result;

/* src/tgt: 3 floats per point (stride 3).  tgt_nrm required when point_to_plane.
 * guess: row-major 4x4 or NULL (identity). */
int oracle_icp(const oicp_params *p, const float *src, int64_t ns, const float *tgt,
               const float *tgt_nrm, int64_t nt, const double *guess, oicp_result *out);

/* ICP over many independent pairs on `nthreads` host threads (one ICP per thread).
 * clouds: cloud c has pts[c] (3 floats per point), nrm[c] (or NULL), n[c].
 * Pair i aligns cloud src[i] onto cloud tgt[i] from guesses[16*i] (NULL = identity). */
int oracle_icp_batch(const oicp_params *p, int nclouds, const float *const *pts,
                     const float *const *nrm, const int64_t *n, int npairs,
                     const int32_t *src, const int32_t *tgt, const double *guesses,
                     int nthreads, oicp_result *out);

/* Registration::getFitnessScore(max_range): src transformed once by T, unbounded
 * 1-NN, mean of d2 over points with d2 <= max_range; DBL_MAX if none. */
double oracle_fitness(const float *src, int64_t ns, const float *tgt, int64_t nt,
                      const double T[16], double max_range, int mode);

/* Correspondences of a given source transform (CorrespondenceEstimation::
 * determineCorrespondences): per source point idx (-1 if d2 > r^2) and d2. */
void oracle_correspondences(const float *src, int64_t ns, const float *tgt, int64_t nt,
                            const double T[16], double max_corr_dist, int32_t *idx,
                            float *d2);

/* The FP64 source transform used by the FP64 oracle and the HIP path:
 * x' = (float)(((R00*x + R01*y) + R02*z) + t0), evaluated in double, no FMA. */
void oracle_transform_points(const float *src, int64_t n, const double T[16], float *out);

/* ---- normals (NormalEstimation::computeFeature + the CLI flip rule) ---- */
/* FP64 symmetric 3x3 Jacobi eigen-decomposition: w ascending, V columns = vectors. */
void oracle_eigh3(const double A[9], double w[3], double V[9]);
/* k nearest neighbours (query point included), FP64 centred covariance, smallest
 * eigenvector, curvature = |l0| / (l0+l1+l2), flipped so (vp - p).n >= 0, then all
 * flipped when flip_all.  Fewer than 3 neighbours -> NaN normal and curvature. */
void oracle_normals(const float *xyz, int64_t n, int k, const float vp[3], int flip_all,
                    float *nrm, float *curv, int nthreads);

/* ---- learned initial-pose proposal (lcr_encoder_oracle.c) ---- */
/* flat FP64 weight layout, row-major [in][out] */
#define ENC_W1 0
#define ENC_B1 (ENC_W1 + 3 * 64)
#define ENC_W2 (ENC_B1 + 64)
#define ENC_B2 (ENC_W2 + 64 * 128)
#define ENC_W3 (ENC_B2 + 128)
#define ENC_B3 (ENC_W3 + 128 * 256)
#define ENC_HW1 (ENC_B3 + 256)
#define ENC_HB1 (ENC_HW1 + 512 * 256)
#define ENC_HW2 (ENC_HB1 + 256)
#define ENC_HB2 (ENC_HW2 + 256 * 6)
#define ENC_NPARAM (ENC_HB2 + 6) /* 174470 */
/* forward of one pair of n-point normalised clouds */
void oenc_forward(const double *w, const float *src, const float *tgt, int n, int quantize,
                  double out[6], double feat[512], int32_t argmax[512]);
/* mean squared loss over `batch` pairs (src/tgt: batch * n * 3 floats, labels:
 * batch * 6) and its gradient (nullable) */
double oenc_loss_grad(const double *w, const float *src, const float *tgt, int n, int batch,
                      const double *labels, int quantize, double *grad);

/* ---- filters before ICP and depth back-projection (lcr_filters_oracle.c) ---- */
int64_t oracle_voxel_grid(const float *xyz, const float *nrm, const float *curv,
                          const uint8_t *rgb, int64_t n, float leaf, float *out_xyz,
                          float *out_nrm, float *out_curv, uint8_t *out_rgb);
int64_t oracle_sor(const float *xyz, int64_t n, int mean_k, double mul, uint8_t *keep,
                   float *mean_dist, double *stats);
int64_t oracle_depth_to_cloud(const uint16_t *depth, int w, int h, const uint8_t *color, double fx,
                              double fy, double cx, double cy, int depth_min, int depth_max,
                              int top, int bottom, int left, int right, float *out_xyz,
                              float *out_nrm, uint8_t *out_rgb);

/* Burst depth fusion of super_resolution.cpp:283-314 (aligned frames): nearest
 * upsampling by `factor`, zero-eliminating mean, Lanczos-4 resize back, 16 bits. */
void oracle_lanczos4_weights(double fr, float w[8]);
void oracle_depth_fuse(const uint16_t *frames, int n_frames, int w, int h, int factor,
                       uint16_t *out);

#ifdef __cplusplus
}
#endif
#endif
