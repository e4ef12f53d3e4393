"""lowcost3dreconstruction_amd -- MI355X fine registration (ICP + kNN/PCA normals).

Python mirror of the PCL interface the reference drives
(pcl_tools/fine_registration.cpp:105-126, pcl_tools/normal_estimation.cpp:84-118),
bound with ctypes to the C ABI of ``liblcr.so`` (include/lcr.h).  Every call runs
the hand-written HIP kernels in ``csrc/``; there is no CPU fallback: importing
this package fails if ``liblcr.so`` has not been built, and calls fail with
``LcrError`` when no GPU is present.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblcr.so")

DBL_MAX = sys.float_info.max

# pcl::registration::DefaultConvergenceCriteria::ConvergenceState
CONV_NOT_CONVERGED = 0
CONV_ITERATIONS = 1
CONV_TRANSFORM = 2
CONV_ABS_MSE = 3
CONV_REL_MSE = 4
CONV_NO_CORRESPONDENCES = 5

PREC_FP32 = 0
PREC_FP16 = 1  # matrix-core FP16 distance tiles + exact FP32 rerank
PREC_BF16 = 2  # matrix-core BF16 distance tiles + exact FP32 rerank
PREC_FP32_TILES = 3  # wavefront-staged tiles scanned with exact FP32 distances

# symbols declared in include/lcr.h
EXPORTS = (
    "lcr_icp_params_default", "lcr_init", "lcr_icp_create", "lcr_icp_destroy",
    "lcr_icp_set_params", "lcr_icp_set_source", "lcr_icp_set_target", "lcr_icp_align",
    "lcr_icp_final_transform", "lcr_icp_has_converged", "lcr_icp_fitness",
    "lcr_icp_iterations", "lcr_icp_convergence_state", "lcr_icp_correspondences",
    "lcr_normals", "lcr_cycle_create", "lcr_cycle_run", "lcr_cycle_results",
    "lcr_cycle_instances", "lcr_cycle_initial_poses", "lcr_cycle_destroy",
    "lcr_icp_batch_align", "lcr_prof_enable",
    "lcr_prof_reset", "lcr_prof_read", "lcr_last_error", "lcr_version",
    "lcr_proposal_create", "lcr_proposal_destroy", "lcr_proposal_set_weights",
    "lcr_proposal_get_weights", "lcr_proposal_forward", "lcr_proposal_loss_grad",
    "lcr_proposal_dataset", "lcr_proposal_train", "lcr_proposal_synthesize",
    "lcr_proposal_predict", "lcr_cycle_autotune", "lcr_transform_points", "lcr_voxel_grid",
    "lcr_outlier_removal", "lcr_depth_to_cloud", "lcr_depth_to_cloud_noalloc",
    "lcr_chain_poses", "lcr_se3_pow", "lcr_centroid_align", "lcr_depth_fuse",
)


class LcrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"liblcr error {code}: {msg}")
        self.code = code


class IcpParams(C.Structure):
    _fields_ = [
        ("max_corr_dist", C.c_double),
        ("max_iterations", C.c_int32),
        ("transformation_epsilon", C.c_double),
        ("euclidean_fitness_epsilon", C.c_double),
        ("point_to_plane", C.c_int32),
        ("precision_mode", C.c_int32),
        ("cell_factor", C.c_double),
        ("tile", C.c_int32),
    ]


class CycleDesc(C.Structure):
    _fields_ = [
        ("n_clouds", C.c_int32),
        ("xyz", C.POINTER(C.POINTER(C.c_float))),
        ("nrm", C.POINTER(C.POINTER(C.c_float))),
        ("n_points", C.POINTER(C.c_int64)),
        ("n_pairs", C.c_int32),
        ("pair_src", C.POINTER(C.c_int32)),
        ("pair_tgt", C.POINTER(C.c_int32)),
        ("n_hyp", C.c_int32),
        ("guesses", C.POINTER(C.c_double)),
        ("params", IcpParams),
        ("normals_k", C.c_int32),
        ("viewpoints", C.POINTER(C.c_float)),
        ("proposal", C.c_void_p),
        ("proposal_points", C.c_int32),
        ("hyp_gate", C.c_int32),
    ]


class CycleResult(C.Structure):
    _fields_ = [
        ("T", C.POINTER(C.c_double)),
        ("iterations", C.POINTER(C.c_int32)),
        ("converged", C.POINTER(C.c_int32)),
        ("state", C.POINTER(C.c_int32)),
        ("fitness", C.POINTER(C.c_double)),
        ("best_hyp", C.POINTER(C.c_int32)),
    ]


class TuneResult(C.Structure):
    _fields_ = [
        ("tile", C.c_int32),
        ("precision_mode", C.c_int32),
        ("cell_factor", C.c_double),
        ("best_ms", C.c_double),
        ("default_ms", C.c_double),
        ("candidates", C.c_int32),
        ("from_cache", C.c_int32),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C lowcost3dreconstruction_amd` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    fp, dp, ip = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int32)
    vp = C.c_void_p
    L.lcr_icp_params_default.argtypes = [C.POINTER(IcpParams)]
    L.lcr_icp_params_default.restype = None
    L.lcr_init.argtypes = [C.c_int]
    L.lcr_icp_create.argtypes = [C.POINTER(IcpParams), C.POINTER(vp)]
    L.lcr_icp_destroy.argtypes = [vp]
    L.lcr_icp_destroy.restype = None
    L.lcr_icp_set_params.argtypes = [vp, C.POINTER(IcpParams)]
    L.lcr_icp_set_source.argtypes = [vp, fp, fp, C.c_int64, C.c_int64]
    L.lcr_icp_set_target.argtypes = [vp, fp, fp, C.c_int64, C.c_int64]
    L.lcr_icp_align.argtypes = [vp, dp, fp, fp]
    L.lcr_icp_final_transform.argtypes = [vp, dp]
    L.lcr_icp_has_converged.argtypes = [vp]
    L.lcr_icp_fitness.argtypes = [vp, C.c_double, dp]
    L.lcr_icp_iterations.argtypes = [vp]
    L.lcr_icp_convergence_state.argtypes = [vp]
    L.lcr_icp_correspondences.argtypes = [vp, dp, ip, fp]
    L.lcr_normals.argtypes = [fp, C.c_int64, C.c_int64, C.c_int32, fp, C.c_int32, fp, fp]
    L.lcr_cycle_create.argtypes = [C.POINTER(CycleDesc), C.POINTER(vp)]
    L.lcr_cycle_run.argtypes = [vp]
    L.lcr_cycle_results.argtypes = [vp, C.POINTER(CycleResult)]
    L.lcr_cycle_instances.argtypes = [vp, dp, ip, ip, dp]
    L.lcr_cycle_initial_poses.argtypes = [vp, dp]
    L.lcr_cycle_destroy.argtypes = [vp]
    L.lcr_cycle_destroy.restype = None
    L.lcr_icp_batch_align.argtypes = [C.POINTER(CycleDesc), C.POINTER(CycleResult)]
    L.lcr_prof_enable.argtypes = [C.c_int32]
    L.lcr_prof_reset.argtypes = []
    L.lcr_prof_read.argtypes = [C.c_char_p, C.POINTER(C.c_int64), dp]
    L.lcr_proposal_create.argtypes = [C.POINTER(vp)]
    L.lcr_proposal_destroy.argtypes = [vp]
    L.lcr_proposal_destroy.restype = None
    L.lcr_proposal_set_weights.argtypes = [vp, fp, C.c_int64]
    L.lcr_proposal_get_weights.argtypes = [vp, fp, C.c_int64]
    L.lcr_proposal_forward.argtypes = [vp, fp, fp, C.c_int32, C.c_int32, fp]
    L.lcr_proposal_loss_grad.argtypes = [vp, fp, fp, fp, C.c_int32, C.c_int32, fp, fp]
    L.lcr_proposal_dataset.argtypes = [vp, fp, fp, fp, C.c_int32, C.c_int32]
    L.lcr_proposal_train.argtypes = [vp, C.c_int32, C.c_int32, C.c_float, C.c_uint64, fp]
    L.lcr_proposal_synthesize.argtypes = [vp, fp, fp, dp, C.c_int32, C.c_int32, C.c_float,
                                          C.c_uint64, fp, fp]
    L.lcr_proposal_predict.argtypes = [vp, C.c_int32, C.c_int32, fp, fp]
    L.lcr_cycle_autotune.argtypes = [vp, C.c_char_p, C.POINTER(TuneResult)]
    L.lcr_transform_points.argtypes = [fp, fp, C.c_int64, C.c_int64, dp, fp, fp]
    u8 = C.POINTER(C.c_uint8)
    i64p = C.POINTER(C.c_int64)
    L.lcr_voxel_grid.argtypes = [fp, fp, fp, u8, C.c_int64, C.c_float, fp, fp, fp, u8, i64p]
    L.lcr_outlier_removal.argtypes = [fp, C.c_int64, C.c_int32, C.c_double, u8, fp, dp, i64p]
    L.lcr_depth_to_cloud.argtypes = [C.POINTER(C.c_uint16), C.c_int32, C.c_int32, u8, dp,
                                     C.c_int32, C.c_int32, C.POINTER(C.c_int32), fp, fp, u8, i64p]
    L.lcr_depth_to_cloud_noalloc.argtypes = [vp, C.c_int32, C.c_int32, vp, dp, C.c_int32,
                                             C.c_int32, C.POINTER(C.c_int32), vp, vp, vp, vp,
                                             vp]
    L.lcr_depth_fuse.argtypes = [C.POINTER(C.c_uint16), C.c_int32, C.c_int32, C.c_int32,
                                 C.c_int32, C.POINTER(C.c_uint16)]
    L.lcr_chain_poses.argtypes = [C.c_int32, dp, dp, dp]
    L.lcr_se3_pow.argtypes = [dp, C.c_double, dp]
    L.lcr_centroid_align.argtypes = [fp, C.c_int64, fp, C.c_int64] + [C.c_double] * 4 + [dp]
    L.lcr_last_error.restype = C.c_char_p
    L.lcr_version.restype = C.c_char_p
    return L


lib = _load()


def _check(rc):
    if rc < 0:
        raise LcrError(rc, lib.lcr_last_error().decode(errors="replace"))
    return rc


def _f32(a, cols=3):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1, cols))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def version():
    return lib.lcr_version().decode()


def init(device=0):
    _check(lib.lcr_init(device))


def default_params():
    p = IcpParams()
    lib.lcr_icp_params_default(C.byref(p))
    return p


class IterativeClosestPoint:
    """pcl::IterativeClosestPoint (point_to_plane=False) or
    IterativeClosestPointWithNormals-style point-to-plane ICP (point_to_plane=True),
    with PCL's defaults (max_iterations 10, no distance limit, epsilons off)."""

    def __init__(self, point_to_plane=False, cell_factor=0.0, tile=0, precision_mode=PREC_FP32):
        self._p = default_params()
        self._p.point_to_plane = int(bool(point_to_plane))
        self._p.cell_factor = float(cell_factor)
        self._p.tile = int(tile)
        self._p.precision_mode = int(precision_mode)
        h = C.c_void_p()
        _check(lib.lcr_icp_create(C.byref(self._p), C.byref(h)))
        self._h = h
        self._src = None
        self._src_nrm = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.lcr_icp_destroy(h)
            self._h = None

    def _push(self):
        _check(lib.lcr_icp_set_params(self._h, C.byref(self._p)))

    # ---- pcl::Registration setters (fine_registration.cpp:108-118)
    def setInputSource(self, cloud, normals=None):
        self._src = _f32(cloud)
        self._src_nrm = _f32(normals) if normals is not None else None
        _check(lib.lcr_icp_set_source(self._h, _fp(self._src), _fp(self._src_nrm),
                                      len(self._src), 12))

    def setInputTarget(self, cloud, normals=None):
        t = _f32(cloud)
        n = _f32(normals) if normals is not None else None
        _check(lib.lcr_icp_set_target(self._h, _fp(t), _fp(n), len(t), 12))

    def setMaxCorrespondenceDistance(self, d):
        self._p.max_corr_dist = float(d)
        self._push()

    def setMaximumIterations(self, n):
        self._p.max_iterations = int(n)
        self._push()

    def setTransformationEpsilon(self, e):
        self._p.transformation_epsilon = float(e)
        self._push()

    def setEuclideanFitnessEpsilon(self, e):
        self._p.euclidean_fitness_epsilon = float(e)
        self._push()

    # ---- align and getters (fine_registration.cpp:121-126)
    def align(self, guess=None):
        """Returns the registered source cloud (N, 3) (and its normals when the source
        has normals, as a second value)."""
        n = len(self._src) if self._src is not None else 0
        out = np.empty((n, 3), np.float32)
        outn = np.empty((n, 3), np.float32) if self._src_nrm is not None else None
        g = np.ascontiguousarray(np.asarray(guess, np.float64).reshape(16)) \
            if guess is not None else None
        _check(lib.lcr_icp_align(self._h, _dp(g), _fp(out), _fp(outn)))
        return (out, outn) if outn is not None else out

    def getFinalTransformation(self):
        T = np.empty(16, np.float64)
        _check(lib.lcr_icp_final_transform(self._h, _dp(T)))
        return T.reshape(4, 4)

    def hasConverged(self):
        return bool(_check(lib.lcr_icp_has_converged(self._h)))

    def getFitnessScore(self, max_range=DBL_MAX):
        out = C.c_double()
        _check(lib.lcr_icp_fitness(self._h, float(max_range), C.byref(out)))
        return out.value

    def nr_iterations(self):
        return _check(lib.lcr_icp_iterations(self._h))

    def getConvergenceState(self):
        return _check(lib.lcr_icp_convergence_state(self._h))

    def correspondences(self, T=None):
        """CorrespondenceEstimation::determineCorrespondences for the source under T
        (default: the final transform): (target index or -1, d2) per source point."""
        n = len(self._src)
        idx = np.empty(n, np.int32)
        d2 = np.empty(n, np.float32)
        TT = np.ascontiguousarray(np.asarray(T, np.float64).reshape(16)) if T is not None else None
        _check(lib.lcr_icp_correspondences(self._h, _dp(TT),
                                           idx.ctypes.data_as(C.POINTER(C.c_int32)), _fp(d2)))
        return idx, d2


def normals(xyz, k, viewpoint=(0.0, 0.0, 0.0), flip_all=False):
    """pcl::NormalEstimation with setKSearch(k) and a viewpoint, then the CLI's
    negation of all normals when flip_all (normal_estimation.cpp:84-118).
    Returns (normals (N, 3) float32, curvature (N,) float32)."""
    x = _f32(xyz)
    vp = np.asarray(viewpoint, np.float32)
    nrm = np.empty_like(x)
    curv = np.empty(len(x), np.float32)
    _check(lib.lcr_normals(_fp(x), len(x), 12, int(k), _fp(vp), int(bool(flip_all)), _fp(nrm),
                           _fp(curv)))
    return nrm, curv


def transform_points(xyz, T, nrm=None):
    """xyz' = T xyz and normals' = R normals on the device (FP64 transform rounded
    once), as pcl::transformPointCloudWithNormals in the transform tool
    (transform.cpp:84-90).  Returns (xyz', normals' or None)."""
    x = _f32(xyz)
    M = np.ascontiguousarray(np.asarray(T, np.float64).reshape(16))
    out = np.empty_like(x)
    n = _f32(nrm) if nrm is not None else None
    out_n = np.empty_like(x) if n is not None else None
    _check(lib.lcr_transform_points(_fp(x), _fp(n) if n is not None else None, len(x), 12, _dp(M),
                                    _fp(out), _fp(out_n) if out_n is not None else None))
    return out, out_n


def depth_fuse(frames, factor=2):
    """Burst depth fusion of super_resolution.cpp:283-314 on the device, for aligned
    frames: nearest upsampling by `factor`, zero-eliminating mean, Lanczos-4 resize
    back, 16 bits.  frames: (K, H, W) uint16 millimetres.  Returns (H, W) uint16."""
    f = np.ascontiguousarray(frames, dtype=np.uint16)
    k, h, w = f.shape
    out = np.empty((h, w), np.uint16)
    u16 = C.POINTER(C.c_uint16)
    _check(lib.lcr_depth_fuse(f.ctypes.data_as(u16), k, w, h, int(factor), out.ctypes.data_as(u16)))
    return out


def chain_poses(T, T_close=None):
    """Poses of the captures of a cycle in capture 0's frame from the pair
    transforms T[i - 1] (capture i onto i - 1), with the closing pair's residual
    (T_close: capture 0 onto the last capture) spread evenly over the loop.
    Replaces alignment.sh:105-113's in-place chaining.  Returns (n, 4, 4)."""
    Ts = np.ascontiguousarray(np.asarray(T, np.float64).reshape(-1, 16))
    n = len(Ts) + 1
    out = np.empty((n, 4, 4), np.float64)
    tc = np.ascontiguousarray(np.asarray(T_close, np.float64).reshape(16)) \
        if T_close is not None else None
    _check(lib.lcr_chain_poses(n, _dp(Ts), _dp(tc) if tc is not None else None, _dp(out)))
    return out


def se3_pow(T, s):
    """T^s = exp(s log T) for a rigid 4x4."""
    M = np.ascontiguousarray(np.asarray(T, np.float64).reshape(16))
    out = np.empty((4, 4), np.float64)
    _check(lib.lcr_se3_pow(_dp(M), float(s), _dp(out)))
    return out


def centroid_align(src, tgt, roll=0.0, pitch=0.0, yaw=0.0, elevation=0.0):
    """centroid_align's transform (centroid_align.cpp:104-130): source centroid to
    the origin, Rx(roll) Ry(pitch) Rz(yaw) in degrees, onto the target centroid,
    then `elevation` along y.  Returns the 4x4."""
    s, t = _f32(src), _f32(tgt)
    out = np.empty((4, 4), np.float64)
    _check(lib.lcr_centroid_align(_fp(s), len(s), _fp(t), len(t), float(roll), float(pitch),
                                  float(yaw), float(elevation), _dp(out)))
    return out


def _u8p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8)) if a is not None else None


def voxel_grid(xyz, leaf, nrm=None, curv=None, rgb=None):
    """pcl::VoxelGrid with setLeafSize(leaf, leaf, leaf) (cloud_downsampling.cpp:73-75)
    on the device.  Returns a dict: xyz and, when given, nrm / curv / rgb of the
    voxel centroids in increasing voxel index."""
    x = _f32(xyz)
    n = len(x)
    nm = _f32(nrm) if nrm is not None else None
    cv = np.ascontiguousarray(curv, np.float32).reshape(-1) if curv is not None else None
    cl = np.ascontiguousarray(rgb, np.uint8).reshape(-1, 3) if rgb is not None else None
    ox = np.empty((max(n, 1), 3), np.float32)
    on = np.empty_like(ox) if nm is not None else None
    oc = np.empty(max(n, 1), np.float32) if cv is not None else None
    orgb = np.empty((max(n, 1), 3), np.uint8) if cl is not None else None
    m = C.c_int64()
    _check(lib.lcr_voxel_grid(_fp(x), _fp(nm), _fp(cv), _u8p(cl), n, float(leaf), _fp(ox), _fp(on),
                              _fp(oc), _u8p(orgb), C.byref(m)))
    k = m.value
    out = {"xyz": ox[:k]}
    if on is not None:
        out["nrm"] = on[:k]
    if oc is not None:
        out["curv"] = oc[:k]
    if orgb is not None:
        out["rgb"] = orgb[:k]
    return out


def outlier_removal(xyz, mean_k=50, stddev_mult=1.0):
    """pcl::StatisticalOutlierRemoval (outlier_removal.cpp:80-84) on the device:
    (keep (N,) bool, per-point mean neighbour distance (N,) float32,
    (mean, stddev, threshold))."""
    x = _f32(xyz)
    keep = np.empty(len(x), np.uint8)
    dist = np.empty(len(x), np.float32)
    stats = np.empty(3)
    kept = C.c_int64()
    _check(lib.lcr_outlier_removal(_fp(x), len(x), int(mean_k), float(stddev_mult), _u8p(keep),
                                   _fp(dist), _dp(stats), C.byref(kept)))
    return keep.astype(bool), dist, tuple(stats)


def depth_to_cloud(depth, intrinsics=(572.882768, 542.739980, 314.649173, 240.160459),
                   depth_min=0, depth_max=65535, crop=None, color=None):
    """The depth tool's back-projection with cross-product normals
    (depth_capture.cpp:230-305) on the device.  depth: (H, W) uint16 millimetres;
    the default intrinsics are the tool's.  Returns (xyz, normals, rgb or None)."""
    d = np.ascontiguousarray(depth, dtype=np.uint16)
    h, w = d.shape
    col = np.ascontiguousarray(color, np.uint8).reshape(h, w, 3) if color is not None else None
    intr = np.ascontiguousarray(intrinsics, np.float64)
    cr = np.ascontiguousarray(crop, np.int32) if crop is not None else None
    ox = np.empty((max(h * w, 1), 3), np.float32)
    on = np.empty_like(ox)
    orgb = np.empty((max(h * w, 1), 3), np.uint8) if col is not None else None
    m = C.c_int64()
    _check(lib.lcr_depth_to_cloud(d.ctypes.data_as(C.POINTER(C.c_uint16)), w, h, _u8p(col),
                                  _dp(intr), int(depth_min), int(depth_max),
                                  cr.ctypes.data_as(C.POINTER(C.c_int32)) if cr is not None else None,
                                  _fp(ox), _fp(on), _u8p(orgb), C.byref(m)))
    k = m.value
    return ox[:k], on[:k], (orgb[:k] if orgb is not None else None)


class NormalEstimation:
    """pcl::NormalEstimation<PointXYZRGB, Normal> as used at normal_estimation.cpp:84-108."""

    def __init__(self):
        self._cloud = None
        self._k = 50
        self._vp = (0.0, 0.0, 0.0)

    def setInputCloud(self, cloud):
        self._cloud = _f32(cloud)

    def setKSearch(self, k):
        self._k = int(k)

    def setViewPoint(self, x, y, z):
        self._vp = (float(x), float(y), float(z))

    def useSensorOriginAsViewPoint(self):
        self._vp = (0.0, 0.0, 0.0)  # PLY clouds carry the sensor origin (0, 0, 0)

    def compute(self):
        return normals(self._cloud, self._k, self._vp)


class Cycle:
    """All pairs (and pose hypotheses) of a capture cycle aligned in one batched
    ICP: one launch per iteration for every instance, replayed from a hipGraph.
    Replaces the per-pair loop of scripts/alignment.sh:106-113."""

    def __init__(self, clouds, pairs, normals=None, guesses=None, n_hyp=1, max_corr_dist=0.05,
                 max_iterations=60, transformation_epsilon=1e-9, euclidean_fitness_epsilon=1e-3,
                 point_to_plane=True, normals_k=30, viewpoints=None, cell_factor=0.0, tile=0,
                 proposal=None, proposal_points=8192, precision_mode=0, hyp_gate=False):
        self.clouds = [_f32(c) for c in clouds]
        self.normals = [_f32(n) for n in normals] if normals is not None else None
        pr = np.asarray(pairs, np.int32).reshape(-1, 2)
        self.n_pairs = len(pr)
        self.n_hyp = int(n_hyp)
        self._src = np.ascontiguousarray(pr[:, 0])
        self._tgt = np.ascontiguousarray(pr[:, 1])
        self._guess = np.ascontiguousarray(np.asarray(guesses, np.float64).reshape(-1)) \
            if guesses is not None else None
        if self._guess is not None and len(self._guess) != 16 * self.n_pairs * self.n_hyp:
            raise ValueError("guesses must hold n_pairs * n_hyp 4x4 matrices")
        fpp = C.POINTER(C.c_float)
        self._xyz_arr = (fpp * len(self.clouds))(*[_fp(c) for c in self.clouds])
        self._nrm_arr = (fpp * len(self.clouds))(*[_fp(n) for n in self.normals]) \
            if self.normals is not None else None
        self._n = np.asarray([len(c) for c in self.clouds], np.int64)
        self._vps = _f32(viewpoints).reshape(-1) if viewpoints is not None else None
        p = default_params()
        p.max_corr_dist = float(max_corr_dist)
        p.max_iterations = int(max_iterations)
        p.transformation_epsilon = float(transformation_epsilon)
        p.euclidean_fitness_epsilon = float(euclidean_fitness_epsilon)
        p.point_to_plane = int(bool(point_to_plane))
        p.cell_factor = float(cell_factor)
        p.tile = int(tile)
        p.precision_mode = int(precision_mode)
        d = CycleDesc()
        d.n_clouds = len(self.clouds)
        d.xyz = self._xyz_arr
        d.nrm = self._nrm_arr
        d.n_points = self._n.ctypes.data_as(C.POINTER(C.c_int64))
        d.n_pairs = self.n_pairs
        d.pair_src = self._src.ctypes.data_as(C.POINTER(C.c_int32))
        d.pair_tgt = self._tgt.ctypes.data_as(C.POINTER(C.c_int32))
        d.n_hyp = self.n_hyp
        d.guesses = _dp(self._guess)
        d.params = p
        d.normals_k = int(normals_k)
        d.viewpoints = _fp(self._vps) if self._vps is not None else None
        self._proposal = proposal  # kept alive while the cycle uses it
        d.proposal = proposal._h if proposal is not None else None
        d.proposal_points = int(proposal_points)
        d.hyp_gate = int(bool(hyp_gate))
        self.n_total_hyp = self.n_hyp + (1 if proposal is not None else 0)
        self._desc = d
        h = C.c_void_p()
        _check(lib.lcr_cycle_create(C.byref(d), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.lcr_cycle_destroy(h)
            self._h = None

    def run(self):
        _check(lib.lcr_cycle_run(self._h))

    def autotune(self, cache_path=None):
        """Times candidate (tile, cell_factor, precision_mode) settings on this cycle
        and keeps the fastest (results are the same for all of them)."""
        r = TuneResult()
        path = None if cache_path is None else os.fsencode(cache_path)
        _check(lib.lcr_cycle_autotune(self._h, path, C.byref(r)))
        return {f: getattr(r, f) for f, _ in TuneResult._fields_}

    def results(self):
        P = self.n_pairs
        T = np.empty((P, 4, 4), np.float64)
        it = np.empty(P, np.int32)
        cv = np.empty(P, np.int32)
        st = np.empty(P, np.int32)
        fit = np.empty(P, np.float64)
        best = np.empty(P, np.int32)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        r = CycleResult(_dp(T), ip(it), ip(cv), ip(st), _dp(fit), ip(best))
        _check(lib.lcr_cycle_results(self._h, C.byref(r)))
        return dict(T=T, iterations=it, converged=cv.astype(bool), state=st, fitness=fit,
                    best_hyp=best)

    def instances(self):
        n = self.n_pairs * self.n_total_hyp
        T = np.empty((n, 4, 4), np.float64)
        it = np.empty(n, np.int32)
        st = np.empty(n, np.int32)
        fit = np.empty(n, np.float64)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        _check(lib.lcr_cycle_instances(self._h, _dp(T), ip(it), ip(st), _dp(fit)))
        return dict(T=T, iterations=it, state=st, fitness=fit)

    def initial_poses(self):
        """(n_pairs * hypotheses, 4, 4): every instance's initial pose in the last
        run, the learned one included."""
        T = np.empty((self.n_pairs * self.n_total_hyp, 4, 4), np.float64)
        _check(lib.lcr_cycle_initial_poses(self._h, _dp(T)))
        return T


ENC_NPARAM = 174470


class Proposal:
    """Learned initial-pose proposal: Siamese point encoder (3->64->128->256,
    max-pool) + head (512->256->6 = axis-angle, translation), BF16 MFMA on the GPU.
    Inputs are pairs of normalised clouds (see `normalise_pair`)."""

    def __init__(self, weights=None):
        h = C.c_void_p()
        _check(lib.lcr_proposal_create(C.byref(h)))
        self._h = h
        if weights is not None:
            self.set_weights(weights)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib.lcr_proposal_destroy(h)
            self._h = None

    def set_weights(self, w):
        w = np.ascontiguousarray(np.asarray(w, np.float32).reshape(-1))
        _check(lib.lcr_proposal_set_weights(self._h, _fp(w), len(w)))

    def get_weights(self):
        w = np.empty(ENC_NPARAM, np.float32)
        _check(lib.lcr_proposal_get_weights(self._h, _fp(w), len(w)))
        return w

    def forward(self, src, tgt):
        s = np.ascontiguousarray(np.asarray(src, np.float32))
        t = np.ascontiguousarray(np.asarray(tgt, np.float32))
        B, n = s.shape[0], s.shape[1]
        out = np.empty((B, 6), np.float32)
        _check(lib.lcr_proposal_forward(self._h, _fp(s), _fp(t), B, n, _fp(out)))
        return out

    def loss_grad(self, src, tgt, labels):
        s = np.ascontiguousarray(np.asarray(src, np.float32))
        t = np.ascontiguousarray(np.asarray(tgt, np.float32))
        lab = np.ascontiguousarray(np.asarray(labels, np.float32))
        B, n = s.shape[0], s.shape[1]
        loss = C.c_float()
        g = np.empty(ENC_NPARAM, np.float32)
        _check(lib.lcr_proposal_loss_grad(self._h, _fp(s), _fp(t), _fp(lab), B, n, C.byref(loss),
                                          _fp(g)))
        return loss.value, g

    def dataset(self, src, tgt, labels):
        s = np.ascontiguousarray(np.asarray(src, np.float32))
        t = np.ascontiguousarray(np.asarray(tgt, np.float32))
        lab = np.ascontiguousarray(np.asarray(labels, np.float32))
        _check(lib.lcr_proposal_dataset(self._h, _fp(s), _fp(t), _fp(lab), s.shape[0], s.shape[1]))

    def synthesize(self, shapes, weights, perturb, n_points, noise=0.001, seed=0, raw=False):
        """Generate and keep resident a training set (see lcr.h).  Returns (raw clouds
        (P, 2, n, 3) or None, frames (P, 4): target centroid and RMS radius)."""
        sh = np.ascontiguousarray(np.asarray(shapes, np.float32).reshape(-1, 45))
        wt = np.ascontiguousarray(np.asarray(weights, np.float32).reshape(-1, 3))
        P = np.ascontiguousarray(np.asarray(perturb, np.float64).reshape(-1, 16))
        n = len(sh)
        out_raw = np.empty((n, 2, n_points, 3), np.float32) if raw else None
        frame = np.empty((n, 4), np.float32)
        _check(lib.lcr_proposal_synthesize(self._h, _fp(sh), _fp(wt), _dp(P), n, int(n_points),
                                           float(noise), int(seed), _fp(out_raw), _fp(frame)))
        return out_raw, frame

    def predict(self, first, count):
        out = np.empty((count, 6), np.float32)
        lab = np.empty((count, 6), np.float32)
        _check(lib.lcr_proposal_predict(self._h, int(first), int(count), _fp(out), _fp(lab)))
        return out, lab

    def train(self, steps, batch, lr=1e-3, seed=0):
        losses = np.empty(steps, np.float32)
        _check(lib.lcr_proposal_train(self._h, int(steps), int(batch), float(lr), int(seed),
                                      _fp(losses)))
        return losses


def prof_enable(on=True):
    _check(lib.lcr_prof_enable(int(bool(on))))


def prof_reset():
    _check(lib.lcr_prof_reset())


def prof_read(name):
    n = C.c_int64()
    ms = C.c_double()
    _check(lib.lcr_prof_read(name.encode(), C.byref(n), C.byref(ms)))
    return n.value, ms.value
