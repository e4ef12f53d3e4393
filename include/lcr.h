/*
 * lcr.h -- C ABI of liblcr, the MI355X fine-registration library.
 *
 * Drop-in boundary for the reference's fine-registration hot path:
 *   pcl::IterativeClosestPoint<PointXYZRGBNormal, PointXYZRGBNormal>
 *     (pcl_tools/fine_registration.cpp:105-126), and
 *   pcl::NormalEstimation<PointXYZRGB, Normal>
 *     (pcl_tools/normal_estimation.cpp:84-118).
 * Plain pointers and sizes only.  Every call runs on the GPU through
 * hand-written HIP kernels; there is no CPU fallback.  Errors: functions return
 * 0 (or a documented non-negative value) on success and a negative LCR_E_* code on
 * failure; lcr_last_error() describes the last failure on the calling thread.
 * Handles are not thread-safe (like a pcl::Registration object): use one per
 * thread.  Input arrays are copied to the device; the caller keeps ownership.
 */
#ifndef LCR_H
#define LCR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define LCR_OK 0
#define LCR_E_ARG (-1)      /* invalid argument */
#define LCR_E_HIP (-2)      /* HIP runtime error (no device, launch failure, ...) */
#define LCR_E_STATE (-3)    /* call out of order (e.g. no source set) */
#define LCR_E_NORMALS (-4)  /* point-to-plane without target normals */
#define LCR_E_LIMIT (-5)    /* size beyond a documented limit */

/* pcl::registration::DefaultConvergenceCriteria::ConvergenceState */
#define LCR_CONV_NOT_CONVERGED 0
#define LCR_CONV_ITERATIONS 1
#define LCR_CONV_TRANSFORM 2
#define LCR_CONV_ABS_MSE 3
#define LCR_CONV_REL_MSE 4
#define LCR_CONV_NO_CORRESPONDENCES 5

/* correspondence search arithmetic */
#define LCR_PREC_FP32 0 /* exact FP32 scan (FLANN distance form) */
#define LCR_PREC_FP16 1 /* FP16 matrix-core distance tiles + exact FP32 rerank (same matches) */
#define LCR_PREC_BF16 2 /* BF16 matrix-core distance tiles + exact FP32 rerank (same matches) */
#define LCR_PREC_FP32_TILES 3 /* wavefront-staged tiles scanned with exact FP32 distances */

/* ICP parameters.  Mirrors the setters called at fine_registration.cpp:112-118:
 * setMaxCorrespondenceDistance, setMaximumIterations, setTransformationEpsilon,
 * setEuclideanFitnessEpsilon; point_to_plane selects
 * TransformationEstimationPointToPlaneLLS instead of TransformationEstimationSVD. */
typedef struct {
  double max_corr_dist;
  int32_t max_iterations;
  double transformation_epsilon;
  double euclidean_fitness_epsilon;
  int32_t point_to_plane;
  int32_t precision_mode; /* LCR_PREC_* */
  double cell_factor;     /* grid cell edge / estimated point spacing; <= 0 = default */
  int32_t tile;           /* source points per work tile: 64, 128 or 256; 0 = default
                             (LCR_PREC_FP16/BF16/FP32_TILES always use 64) */
} lcr_icp_params;

typedef struct lcr_icp lcr_icp;

/* PCL's defaults: max_iterations 10, max_corr_dist sqrt(DBL_MAX), epsilons 0 and
 * -DBL_MAX (pcl/registration/registration.h constructor). */
void lcr_icp_params_default(lcr_icp_params *p);

/* Selects the HIP device used by handles created afterwards on this thread. */
int lcr_init(int device);

/* `IterativeClosestPoint icp;` (fine_registration.cpp:105) */
int lcr_icp_create(const lcr_icp_params *p, lcr_icp **out);
void lcr_icp_destroy(lcr_icp *h);
/* setMaxCorrespondenceDistance / setMaximumIterations / setTransformationEpsilon /
 * setEuclideanFitnessEpsilon (fine_registration.cpp:112-118) */
int lcr_icp_set_params(lcr_icp *h, const lcr_icp_params *p);

/* setInputSource (fine_registration.cpp:108).  xyz: n points, x,y,z floats at the
 * start of each record of `stride_bytes` (12 for packed xyz, 48 for PCL's
 * PointXYZRGBNormal); nrm: nullable, same stride. */
int lcr_icp_set_source(lcr_icp *h, const float *xyz, const float *nrm, int64_t n,
                       int64_t stride_bytes);
/* setInputTarget (fine_registration.cpp:109).  The target's search grid is built
 * once per new target, on the next align. */
int lcr_icp_set_target(lcr_icp *h, const float *xyz, const float *nrm, int64_t n,
                       int64_t stride_bytes);

/* align(output) / align(output, guess) (fine_registration.cpp:121).  guess:
 * row-major 4x4 or NULL (identity).  out_xyz / out_nrm (nullable): the source
 * transformed by the final transform, 3 floats per source point in input order
 * (IterativeClosestPoint::transformCloud; non-finite points copied). */
int lcr_icp_align(lcr_icp *h, const double *guess, float *out_xyz, float *out_nrm);

/* getFinalTransformation (fine_registration.cpp:124): row-major 4x4. */
int lcr_icp_final_transform(const lcr_icp *h, double T[16]);
/* hasConverged (fine_registration.cpp:125): 1, 0, or < 0 on error. */
int lcr_icp_has_converged(const lcr_icp *h);
/* getFitnessScore(max_range) (fine_registration.cpp:126): mean squared distance
 * from each source point, transformed by the final transform, to its nearest
 * target point, over points whose squared distance is <= max_range (PCL compares
 * the squared distance with max_range as given); DBL_MAX if none. */
int lcr_icp_fitness(lcr_icp *h, double max_range, double *out);
/* nr_iterations_ and DefaultConvergenceCriteria::getConvergenceState */
int lcr_icp_iterations(const lcr_icp *h);
int lcr_icp_convergence_state(const lcr_icp *h);
/* CorrespondenceEstimation::determineCorrespondences for the source transformed by
 * T (row-major, NULL = the current final transform): per source point the target
 * index of its nearest neighbour, or -1 when d2 > max_corr_dist^2 or the point is
 * not finite; d2 is that nearest distance (FLANN float form) for kept points, and
 * for rejected ones either the nearest distance found or +inf when no target
 * point lies within max_corr_dist of the query's grid cell (nothing searched). */
int lcr_icp_correspondences(lcr_icp *h, const double *T, int32_t *idx, float *d2);

/* pcl::NormalEstimation with setKSearch(k) and a viewpoint, then the CLI's flip
 * of all normals when flip_all != 0 (normal_estimation.cpp:84-118).  xyz: n
 * records of stride_bytes.  nrm: 3 floats per point, curv: 1 float per point
 * (nullable); NaN where fewer than 3 neighbours exist or the point is not finite. */
int lcr_normals(const float *xyz, int64_t n, int64_t stride_bytes, int32_t k, const float vp[3],
                int32_t flip_all, float *nrm, float *curv);

/* ---- batched alignment: every adjacent pair of a capture cycle (and optional
 * pose hypotheses per pair) in one launch per ICP iteration.  Replaces the
 * per-pair process loop of scripts/alignment.sh:106-113. */
struct lcr_proposal;

typedef struct {
  int32_t n_clouds;
  const float *const *xyz;      /* per cloud: packed x,y,z floats */
  const float *const *nrm;      /* per cloud normals (packed) or NULL: computed on
                                   the device when point_to_plane is set */
  const int64_t *n_points;
  int32_t n_pairs;
  const int32_t *pair_src;      /* source cloud of each pair (capture i) */
  const int32_t *pair_tgt;      /* target cloud of each pair (capture i-1) */
  int32_t n_hyp;                /* initial-pose hypotheses per pair, >= 1 */
  const double *guesses;        /* n_pairs * n_hyp row-major 4x4, or NULL = identity */
  lcr_icp_params params;
  int32_t normals_k;            /* k of the on-device normal estimation (default 30) */
  const float *viewpoints;      /* per cloud 3 floats, or NULL = origin */
  struct lcr_proposal *proposal; /* nullable: adds one learned hypothesis per pair,
                                   seeded from the pair's first guess; the best
                                   hypothesis by fitness wins, so a worse learned
                                   guess falls back to the plain one */
  int32_t proposal_points;      /* points per cloud fed to the network (default 8192) */
  int32_t hyp_gate;             /* 0: refine every hypothesis, keep the best by final
                                   fitness; 1: first score every hypothesis's initial
                                   pose by its source points within max_corr_dist
                                   (ties: smaller mean d2) on a subsample of the
                                   source, then refine the pair's first hypothesis
                                   (the plain seed) and, when another one scores
                                   higher, that one too; the lower final fitness
                                   of the two wins (ties: the first) */
} lcr_cycle_desc;

typedef struct {
  double *T;            /* n_pairs * 16: best hypothesis's final transform */
  int32_t *iterations;  /* n_pairs */
  int32_t *converged;   /* n_pairs */
  int32_t *state;       /* n_pairs */
  double *fitness;      /* n_pairs: getFitnessScore() of the best hypothesis */
  int32_t *best_hyp;    /* n_pairs: index of the chosen hypothesis (nullable); n_hyp
                           means the learned one */
} lcr_cycle_result;

typedef struct lcr_cycle lcr_cycle;
/* Uploads the clouds once and prepares the batched iteration. */
int lcr_cycle_create(const lcr_cycle_desc *d, lcr_cycle **out);
/* normals (when needed) -> grids -> ICP iterations (one hipGraph) -> fitness ->
 * hypothesis selection.  Synchronous. */
int lcr_cycle_run(lcr_cycle *c);
int lcr_cycle_results(lcr_cycle *c, lcr_cycle_result *r);
/* per-instance (pair x hypothesis) results before selection; arrays of
 * n_pairs * (n_hyp + 1 if a proposal is set, else n_hyp) entries (each nullable) */
int lcr_cycle_instances(lcr_cycle *c, double *T, int32_t *iterations, int32_t *state,
                        double *fitness);
/* The initial pose of every instance of the last run, in the same order as
 * lcr_cycle_instances (row-major 4x4 each): the given guesses and, with a
 * proposal, the learned pose the network predicted.  Lets a caller rerun any
 * instance through the single-pair path (alignment.sh:105-113 from that pose). */
int lcr_cycle_initial_poses(lcr_cycle *c, double *T);
void lcr_cycle_destroy(lcr_cycle *c);

/* Autotuner (no counterpart in the reference, which has no launch shape to
 * choose).  Times repeat runs of the cycle under candidate launch settings --
 * tile (64/128/256), cell_factor (1.5..4) and precision_mode (FP32 scan or
 * FP16/BF16 distance tiles) -- and keeps the fastest in the cycle's parameters.
 * Every candidate gives the same correspondences; only speed differs.  With a
 * cache_path (nullable) the choice is looked up first and appended after
 * tuning; entries are keyed by the device and the workload's shape
 * (clouds, pairs, hypotheses, total points, radius, estimator). */
typedef struct {
  int32_t tile;
  int32_t precision_mode;
  double cell_factor;
  double best_ms;      /* run time of the chosen setting (0 when from the cache) */
  double default_ms;   /* run time of the default setting (0 when from the cache) */
  int32_t candidates;  /* settings timed (0 when from the cache) */
  int32_t from_cache;
} lcr_tune_result;
int lcr_cycle_autotune(lcr_cycle *c, const char *cache_path, lcr_tune_result *out);

/* ---- the cycle driver's pose bookkeeping (host arithmetic, no device work) ---- */

/* In-memory pose chaining, replacing alignment.sh:105-113's in-place rewrites of
 * every later capture: poses[j] = T_1 T_2 ... T_j maps capture j into capture 0's
 * frame (poses[0] = I), T (row-major 4x4 each, n_captures - 1 of them) holding
 * T_i, the transform of capture i onto capture i - 1.  With T_close (nullable:
 * capture 0 onto capture n_captures - 1) the loop residual E = poses[n-1] T_close
 * -- the identity for a consistent cycle -- is spread evenly over the loop's
 * n_captures edges: poses[j] <- E^(-j/n) poses[j], E^s = exp(s log E) in SE(3);
 * every edge of the loop, the closing one included, then carries E^(1/n). */
int lcr_chain_poses(int32_t n_captures, const double *T, const double *T_close, double *poses);
/* out = T^s = exp(s log T) for a rigid row-major 4x4 T (the screw motion scaled by s). */
int lcr_se3_pow(const double T[16], double s, double out[16]);
/* centroid_align's transform (pcl_tools/centroid_align.cpp:104-130), the seed
 * alignment.sh:117-118 gives the top and bottom views: move the source's centroid
 * to the origin, rotate by Rx(roll) Ry(pitch) Rz(yaw) (degrees), move it to the
 * target's centroid, then by `elevation` along y.  Packed xyz; centroids over the
 * finite points. */
int lcr_centroid_align(const float *src, int64_t n_src, const float *tgt, int64_t n_tgt,
                       double roll_deg, double pitch_deg, double yaw_deg, double elevation,
                       double T[16]);

/* Rigid transform of a cloud on the device: xyz' = T xyz (FP64, rounded once to
 * float) and normal' = R normal, non-finite entries passed through.  Replaces
 * pcl::transformPointCloudWithNormals as the transform tool applies a 4x4
 * (pcl_tools/transform.cpp:84-90) and as alignment.sh:109-110 chains a pair's
 * result onto the later captures.  nrm / out_nrm nullable; T row-major. */
int lcr_transform_points(const float *xyz, const float *nrm, int64_t n, int64_t stride_bytes,
                         const double T[16], float *out_xyz, float *out_nrm);

/* ---- the steps before ICP: VoxelGrid, outlier removal, depth to cloud ---- */

/* pcl::VoxelGrid<PointXYZRGBNormal> with setLeafSize(leaf, leaf, leaf) and filter
 * (pcl_tools/cloud_downsampling.cpp:73-75): one point per occupied voxel of the
 * lattice floor(p / leaf), in increasing voxel index (x fastest): xyz mean, normal
 * mean normalised, curvature mean, per-channel rgb mean truncated (nrm, curv, rgb
 * and their outputs nullable; xyz/nrm/rgb packed, 3 per point).  Output buffers
 * hold up to n points; *out_n receives the count.  A lattice of more than
 * INT32_MAX voxels returns the input unchanged, as PCL does; one of more than
 * 2^30 voxels fails with LCR_E_LIMIT. */
int lcr_voxel_grid(const float *xyz, const float *nrm, const float *curv, const uint8_t *rgb,
                   int64_t n, float leaf, float *out_xyz, float *out_nrm, float *out_curv,
                   uint8_t *out_rgb, int64_t *out_n);

/* pcl::StatisticalOutlierRemoval with setMeanK(mean_k), setStddevMulThresh(mult)
 * and filter (pcl_tools/outlier_removal.cpp:80-84): per finite point the mean
 * distance to its mean_k nearest neighbours; mean and sample stddev over finite
 * points; keep[i] = 1 when the point's value is <= mean + mult * stddev
 * (non-finite points count 0 and are kept, as PCL's loop leaves them).
 * mean_dist (nullable): per-point values; stats (nullable): mean, stddev,
 * threshold; kept (nullable): number kept.  setNegative(true) is keep[i] == 0. */
int lcr_outlier_removal(const float *xyz, int64_t n, int32_t mean_k, double stddev_mult,
                        uint8_t *keep, float *mean_dist, double *stats, int64_t *kept);

/* The depth tool's back-projection (capture/depth_capture/depth_capture.cpp:230-305):
 * a width x height 16-bit depth map (millimetres, row-major) becomes points
 * (z (j - cx) / fx, -z (i - cy) / fy, -z) for pixels with depth_min <= z <=
 * depth_max, z != 0 and (i, j) inside crop = {top, bottom, left, right} (NULL =
 * whole image), in row-major order, with cross-product normals (neighbours that
 * differ by more than 20 mm ignored) and the pixel's colour (color: width x height
 * x 3 bytes, nullable).  intrinsics = {fx, fy, cx, cy}.  Outputs hold up to
 * width * height points. */
int lcr_depth_to_cloud(const uint16_t *depth, int32_t width, int32_t height, const uint8_t *color,
                       const double intrinsics[4], int32_t depth_min, int32_t depth_max,
                       const int32_t crop[4], float *out_xyz, float *out_nrm, uint8_t *out_rgb,
                       int64_t *out_n);

/* Burst depth fusion (capture/super_resolution/super_resolution.cpp:283-314) for
 * frames already aligned (the ECC registration and warp of :255-279 are the
 * identity for a static rig): each frame upsampled by `factor` with nearest
 * neighbours, the zero-eliminating mean per pixel (the mean of the positive
 * depths, in double), the Lanczos-4 resize back to width x height (source position
 * (x + 0.5) factor - 0.5, 8 taps per axis clamped to the image, horizontal then
 * vertical, float) and the 16-bit conversion (round to nearest even, clamped, NaN
 * -> 0).  frames: n_frames x height x width uint16 millimetres, row-major; out:
 * height x width.  factor in 1..16. */
int lcr_depth_fuse(const uint16_t *frames, int32_t n_frames, int32_t width, int32_t height,
                   int32_t factor, uint16_t *out);

/* lcr_depth_to_cloud for callers that cannot allocate device memory.  No
 * counterpart in the reference: it is for code that may not call hipMalloc --
 * work under HIP graph capture, a per-frame capture loop, harnesses that forbid
 * allocation.  Same computation and outputs as lcr_depth_to_cloud, but every
 * array is a DEVICE pointer (depth, color, out_xyz, out_nrm, out_rgb; color,
 * out_nrm and out_rgb nullable), *out_n is written on the device (one int64), and
 * the work is enqueued on `stream` (a hipStream_t; NULL = the null stream)
 * without waiting for it; intrinsics and crop stay host arrays.  Nothing is
 * allocated: the kernels' scratch is a fixed array of 8 MiB per device, taken by
 * lcr_init or on the first call on a device.  width * height must be at most
 * LCR_NOALLOC_MAX_PIXELS (else LCR_E_LIMIT).  Every call on a device shares its
 * scratch, so calls that overlap on one device (other streams, other threads)
 * race and must be serialised by the caller. */
#define LCR_NOALLOC_MAX_PIXELS (1 << 20)
int lcr_depth_to_cloud_noalloc(const uint16_t *depth, int32_t width, int32_t height,
                               const uint8_t *color, const double intrinsics[4], int32_t depth_min,
                               int32_t depth_max, const int32_t crop[4], float *out_xyz,
                               float *out_nrm, uint8_t *out_rgb, int64_t *out_n, void *stream);

/* One-shot: create, run, results, destroy. */
int lcr_icp_batch_align(const lcr_cycle_desc *d, lcr_cycle_result *r);

/* ---- learned initial-pose proposal.  No counterpart in the reference: it takes
 * the role of the coarse alignment before fine registration (Super4PCS at
 * scripts/alignment.sh:107).  A Siamese point encoder (3->64->128->256, max-pool,
 * shared by both clouds) and a head (512->256->6) map a pair of normalised clouds
 * to an increment (axis-angle omega, translation tau).  Layers 2-3 run on BF16
 * MFMA with FP32 accumulation; weights are FP32 (174470 floats, row-major
 * [in][out]: W1 3x64, b1, W2 64x128, b2, W3 128x256, b3, Hw1 512x256, hb1, Hw2
 * 256x6, hb2).  Clouds are n_pairs * n_points * 3 floats each. */
typedef struct lcr_proposal lcr_proposal;
int lcr_proposal_create(lcr_proposal **out);
void lcr_proposal_destroy(lcr_proposal *p);
int lcr_proposal_set_weights(lcr_proposal *p, const float *w, int64_t n);
int lcr_proposal_get_weights(lcr_proposal *p, float *w, int64_t n);
/* out: n_pairs * 6 */
int lcr_proposal_forward(lcr_proposal *p, const float *src, const float *tgt, int32_t n_pairs,
                         int32_t n_points, float *out);
/* mean over pairs of the summed squared error against labels (n_pairs * 6) and
 * its gradient with respect to the weights (grad: 174470 floats, nullable) */
int lcr_proposal_loss_grad(lcr_proposal *p, const float *src, const float *tgt,
                           const float *labels, int32_t n_pairs, int32_t n_points, float *loss,
                           float *grad);
/* uploads a training set to the device (kept resident for lcr_proposal_train) */
int lcr_proposal_dataset(lcr_proposal *p, const float *src, const float *tgt, const float *labels,
                         int32_t n_pairs, int32_t n_points);
/* Generates a training set of n_pairs on the device and keeps it resident.  Per
 * pair: shapes = 3 ellipsoids x (centre 3, row-major rotation 9, semi-axes 3) = 45
 * floats, weights = the 3 ellipsoids' surface-area shares, perturb = the motion
 * applied to the source (row-major 4x4); both clouds are independent area-uniform
 * samples with N(0, noise^2) per coordinate, normalised by the target's centroid
 * and RMS radius; labels = the inverse motion in that frame (omega, tau).
 * raw (nullable): the un-normalised clouds, n_pairs * 2 * n_points * 3 floats
 * (source, target per pair); frame (nullable): n_pairs * 4 (centroid, radius). */
int lcr_proposal_synthesize(lcr_proposal *p, const float *shapes, const float *weights,
                            const double *perturb, int32_t n_pairs, int32_t n_points, float noise,
                            uint64_t seed, float *raw, float *frame);
/* forward on resident pairs [first, first + count): out (and labels, nullable)
 * count * 6 floats */
int lcr_proposal_predict(lcr_proposal *p, int32_t first, int32_t count, float *out, float *labels);
/* `steps` Adam steps on random batches of the resident set; losses: per step
 * (nullable) */
int lcr_proposal_train(lcr_proposal *p, int32_t steps, int32_t batch, float lr, uint64_t seed,
                       float *losses);

/* ---- profiling: per-kernel GPU time measured with HIP events on the stream the
 * kernel runs on.  names: "icp_iteration", "fitness", "grid_build", "normals". */
int lcr_prof_enable(int32_t on);
int lcr_prof_reset(void);
int lcr_prof_read(const char *name, int64_t *launches, double *total_ms);

const char *lcr_last_error(void);
const char *lcr_version(void);

#ifdef __cplusplus
}
#endif
#endif
