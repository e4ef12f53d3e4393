"""CPU oracle for the fine-registration hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker the HIP path is compared against (and the timed host
baseline).  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it.  The product library ``liblcr.so`` never links,
loads or calls it; there is no CPU fallback in the product path.

It wraps ``oracle/liblcr_oracle.so`` (plain C, built by ``oracle/Makefile``), which
restates the PCL semantics behind the reference's calls at
pcl_tools/fine_registration.cpp:105-126 and pcl_tools/normal_estimation.cpp:84-118.
Every function is pinned by tests against numpy golden vectors, closed forms or
brute force (tests/test_oracle.py).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblcr_oracle.so")

FP64 = 0
PCL32 = 1

# DefaultConvergenceCriteria::ConvergenceState
CONV_NOT_CONVERGED = 0
CONV_ITERATIONS = 1
CONV_TRANSFORM = 2
CONV_ABS_MSE = 3
CONV_REL_MSE = 4
CONV_NO_CORRESPONDENCES = 5


class OicpParams(C.Structure):
    _fields_ = [
        ("max_corr_dist", C.c_double),
        ("max_iterations", C.c_int),
        ("transformation_epsilon", C.c_double),
        ("euclidean_fitness_epsilon", C.c_double),
        ("point_to_plane", C.c_int),
        ("mode", C.c_int),
    ]


class OicpResult(C.Structure):
    _fields_ = [
        ("T", C.c_double * 16),
        ("iterations", C.c_int),
        ("converged", C.c_int),
        ("state", C.c_int),
        ("fitness", C.c_double),
        ("last_ncorr", C.c_int64),
        ("last_mse", C.c_double),
    ]


class Oconv(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_int),
        ("rotation_threshold", C.c_double),
        ("translation_threshold", C.c_double),
        ("mse_threshold_relative", C.c_double),
        ("mse_threshold_absolute", C.c_double),
        ("prev_mse", C.c_double),
        ("iterations_similar", C.c_int),
        ("max_iterations_similar", C.c_int),
        ("state", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        fp, dp, ip = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int32)
        L.oracle_nn.argtypes = [fp, C.c_int64, fp, C.c_int64, C.c_double, ip, fp]
        L.obf_nn.argtypes = [fp, C.c_int64, fp, C.c_int64, C.c_double, ip, fp]
        L.obf_knn.argtypes = [fp, C.c_int64, fp, C.c_int, ip, fp]
        L.obf_knn.restype = C.c_int
        L.okd_build.argtypes = [fp, C.c_int64, C.c_int64]
        L.okd_build.restype = C.c_void_p
        L.okd_free.argtypes = [C.c_void_p]
        L.okd_knn.argtypes = [C.c_void_p, fp, C.c_int, ip, fp]
        L.okd_knn.restype = C.c_int
        L.oracle_umeyama.argtypes = [dp, dp, C.c_int64, dp]
        L.oracle_umeyama.restype = C.c_int
        L.oracle_svd3.argtypes = [dp, dp, dp, dp]
        L.oracle_eigh3.argtypes = [dp, dp, dp]
        L.oracle_p2pl.argtypes = [dp, dp, dp, C.c_int64, dp]
        L.oracle_p2pl.restype = C.c_int
        L.oracle_p2pl_matrix.argtypes = [C.c_double] * 6 + [dp]
        L.oconv_init.argtypes = [C.POINTER(Oconv), C.c_int, C.c_double, C.c_double]
        L.oconv_step.argtypes = [C.POINTER(Oconv), C.c_int, dp, C.c_double]
        L.oconv_step.restype = C.c_int
        L.oracle_icp.argtypes = [C.POINTER(OicpParams), fp, C.c_int64, fp, fp, C.c_int64, dp,
                                 C.POINTER(OicpResult)]
        L.oracle_icp.restype = C.c_int
        L.oracle_icp_batch.argtypes = [C.POINTER(OicpParams), C.c_int, C.POINTER(fp),
                                       C.POINTER(fp), C.POINTER(C.c_int64), C.c_int, ip, ip, dp,
                                       C.c_int, C.POINTER(OicpResult)]
        L.oracle_icp_batch.restype = C.c_int
        L.oracle_fitness.argtypes = [fp, C.c_int64, fp, C.c_int64, dp, C.c_double, C.c_int]
        L.oracle_fitness.restype = C.c_double
        L.oracle_correspondences.argtypes = [fp, C.c_int64, fp, C.c_int64, dp, C.c_double, ip,
                                             fp]
        L.oracle_transform_points.argtypes = [fp, C.c_int64, dp, fp]
        L.oracle_normals.argtypes = [fp, C.c_int64, C.c_int, fp, C.c_int, fp, fp, C.c_int]
        L.oenc_forward.argtypes = [dp, fp, fp, C.c_int, C.c_int, dp, dp, ip]
        L.oenc_loss_grad.argtypes = [dp, fp, fp, C.c_int, C.c_int, dp, C.c_int, dp]
        L.oenc_loss_grad.restype = C.c_double
        u8 = C.POINTER(C.c_uint8)
        L.oracle_voxel_grid.argtypes = [fp, fp, fp, u8, C.c_int64, C.c_float, fp, fp, fp, u8]
        L.oracle_voxel_grid.restype = C.c_int64
        L.oracle_sor.argtypes = [fp, C.c_int64, C.c_int, C.c_double, u8, fp, dp]
        L.oracle_sor.restype = C.c_int64
        L.oracle_depth_to_cloud.argtypes = [C.POINTER(C.c_uint16), C.c_int, C.c_int, u8] + \
            [C.c_double] * 4 + [C.c_int] * 6 + [fp, fp, u8]
        L.oracle_depth_to_cloud.restype = C.c_int64
        L.oracle_depth_fuse.argtypes = [C.POINTER(C.c_uint16), C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.POINTER(C.c_uint16)]
        L.oracle_depth_fuse.restype = None
        L.oracle_lanczos4_weights.argtypes = [C.c_double, fp]
        L.oracle_lanczos4_weights.restype = None
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def nn(target, queries, max_dist=np.inf, brute=False):
    """1-NN of each query in target (FLANN float distance, ties -> smallest index).
    Returns (idx int32 with -1 beyond max_dist, d2 float32)."""
    t, q = _f32(target).reshape(-1, 3), _f32(queries).reshape(-1, 3)
    idx = np.empty(len(q), np.int32)
    d2 = np.empty(len(q), np.float32)
    md2 = float(max_dist) ** 2 if np.isfinite(max_dist) else float("inf")
    fn = lib().obf_nn if brute else lib().oracle_nn
    fn(_p(t, C.c_float), len(t), _p(q, C.c_float), len(q), md2, _p(idx, C.c_int32),
       _p(d2, C.c_float))
    return idx, d2


def knn(target, query, k, brute=False):
    """k nearest neighbours of one query, sorted by (d2, index)."""
    t, q = _f32(target).reshape(-1, 3), _f32(query).reshape(3)
    idx = np.empty(max(k, 1), np.int32)
    d2 = np.empty(max(k, 1), np.float32)
    if brute:
        n = lib().obf_knn(_p(t, C.c_float), len(t), _p(q, C.c_float), k, _p(idx, C.c_int32),
                          _p(d2, C.c_float))
    else:
        h = lib().okd_build(_p(t, C.c_float), len(t), 3)
        try:
            n = lib().okd_knn(h, _p(q, C.c_float), k, _p(idx, C.c_int32), _p(d2, C.c_float))
        finally:
            lib().okd_free(h)
    return idx[:n], d2[:n]


def umeyama(src, dst):
    s, d = _f64(src).reshape(-1, 3), _f64(dst).reshape(-1, 3)
    T = np.empty(16, np.float64)
    lib().oracle_umeyama(_p(s, C.c_double), _p(d, C.c_double), len(s), _p(T, C.c_double))
    return T.reshape(4, 4)


def svd3(A):
    A = _f64(A).reshape(9)
    U, s, V = np.empty(9), np.empty(3), np.empty(9)
    lib().oracle_svd3(_p(A, C.c_double), _p(U, C.c_double), _p(s, C.c_double),
                      _p(V, C.c_double))
    return U.reshape(3, 3), s, V.reshape(3, 3)


def eigh3(A):
    A = _f64(A).reshape(9)
    w, V = np.empty(3), np.empty(9)
    lib().oracle_eigh3(_p(A, C.c_double), _p(w, C.c_double), _p(V, C.c_double))
    return w, V.reshape(3, 3)


def p2pl(src, dst, nrm):
    s, d, m = _f64(src).reshape(-1, 3), _f64(dst).reshape(-1, 3), _f64(nrm).reshape(-1, 3)
    T = np.empty(16, np.float64)
    rc = lib().oracle_p2pl(_p(s, C.c_double), _p(d, C.c_double), _p(m, C.c_double), len(s),
                           _p(T, C.c_double))
    return T.reshape(4, 4), rc


def p2pl_matrix(a, b, g, tx, ty, tz):
    T = np.empty(16, np.float64)
    lib().oracle_p2pl_matrix(a, b, g, tx, ty, tz, _p(T, C.c_double))
    return T.reshape(4, 4)


class Convergence:
    """DefaultConvergenceCriteria state machine (one instance per ICP run)."""

    def __init__(self, max_iterations, transformation_epsilon, euclidean_fitness_epsilon):
        self.c = Oconv()
        lib().oconv_init(C.byref(self.c), max_iterations, transformation_epsilon,
                         euclidean_fitness_epsilon)

    def step(self, iterations, Tinc, mse):
        T = _f64(Tinc).reshape(16)
        return bool(lib().oconv_step(C.byref(self.c), iterations, _p(T, C.c_double), mse))

    @property
    def state(self):
        return self.c.state


def _params(max_corr_dist, max_iterations, transformation_epsilon, euclidean_fitness_epsilon,
            point_to_plane, mode):
    return OicpParams(float(max_corr_dist), int(max_iterations), float(transformation_epsilon),
                      float(euclidean_fitness_epsilon), int(bool(point_to_plane)), int(mode))


def _result_dict(r):
    return dict(T=np.array(r.T[:], dtype=np.float64).reshape(4, 4), iterations=r.iterations,
                converged=bool(r.converged), state=r.state, fitness=r.fitness,
                last_ncorr=r.last_ncorr, last_mse=r.last_mse)


def icp(src, tgt, tgt_nrm=None, guess=None, max_corr_dist=0.1, max_iterations=50,
        transformation_epsilon=1e-9, euclidean_fitness_epsilon=1e-3, point_to_plane=False,
        mode=FP64):
    """IterativeClosestPoint::align + getFitnessScore with the reference CLI's defaults
    (fine_registration.cpp:51-54)."""
    s, t = _f32(src).reshape(-1, 3), _f32(tgt).reshape(-1, 3)
    m = _f32(tgt_nrm).reshape(-1, 3) if tgt_nrm is not None else None
    g = _f64(guess).reshape(16) if guess is not None else None
    p = _params(max_corr_dist, max_iterations, transformation_epsilon, euclidean_fitness_epsilon,
                point_to_plane, mode)
    r = OicpResult()
    rc = lib().oracle_icp(C.byref(p), _p(s, C.c_float), len(s), _p(t, C.c_float),
                          _p(m, C.c_float) if m is not None else None, len(t),
                          _p(g, C.c_double) if g is not None else None, C.byref(r))
    out = _result_dict(r)
    out["rc"] = rc
    return out


def icp_batch(clouds, normals, pairs, guesses=None, nthreads=8, **kw):
    """Run the oracle ICP over many (src, tgt) pairs on host threads."""
    cl = [_f32(c).reshape(-1, 3) for c in clouds]
    nr = [(_f32(n).reshape(-1, 3) if n is not None else None) for n in normals] if normals \
        is not None else None
    fp = C.POINTER(C.c_float)
    pts = (fp * len(cl))(*[_p(c, C.c_float) for c in cl])
    nrm = (fp * len(cl))(*[(_p(n, C.c_float) if n is not None else fp()) for n in nr]) \
        if nr is not None else None
    ns = (C.c_int64 * len(cl))(*[len(c) for c in cl])
    pr = np.asarray(pairs, dtype=np.int32).reshape(-1, 2)
    src = np.ascontiguousarray(pr[:, 0])
    tgt = np.ascontiguousarray(pr[:, 1])
    g = _f64(guesses).reshape(-1) if guesses is not None else None
    p = _params(kw.get("max_corr_dist", 0.1), kw.get("max_iterations", 50),
                kw.get("transformation_epsilon", 1e-9), kw.get("euclidean_fitness_epsilon", 1e-3),
                kw.get("point_to_plane", False), kw.get("mode", FP64))
    res = (OicpResult * len(pr))()
    lib().oracle_icp_batch(C.byref(p), len(cl), pts, nrm, ns, len(pr), _p(src, C.c_int32),
                           _p(tgt, C.c_int32), _p(g, C.c_double) if g is not None else None,
                           nthreads, res)
    return [_result_dict(r) for r in res]


def fitness(src, tgt, T, max_range=float("inf"), mode=FP64):
    s, t = _f32(src).reshape(-1, 3), _f32(tgt).reshape(-1, 3)
    TT = _f64(T).reshape(16)
    return lib().oracle_fitness(_p(s, C.c_float), len(s), _p(t, C.c_float), len(t),
                                _p(TT, C.c_double), float(max_range), mode)


def correspondences(src, tgt, T, max_corr_dist):
    s, t = _f32(src).reshape(-1, 3), _f32(tgt).reshape(-1, 3)
    TT = _f64(T).reshape(16)
    idx = np.empty(len(s), np.int32)
    d2 = np.empty(len(s), np.float32)
    lib().oracle_correspondences(_p(s, C.c_float), len(s), _p(t, C.c_float), len(t),
                                 _p(TT, C.c_double), float(max_corr_dist), _p(idx, C.c_int32),
                                 _p(d2, C.c_float))
    return idx, d2


def transform_points(src, T):
    s = _f32(src).reshape(-1, 3)
    TT = _f64(T).reshape(16)
    out = np.empty_like(s)
    lib().oracle_transform_points(_p(s, C.c_float), len(s), _p(TT, C.c_double),
                                  _p(out, C.c_float))
    return out


def normals(xyz, k, viewpoint=(0.0, 0.0, 0.0), flip_all=False, nthreads=8):
    x = _f32(xyz).reshape(-1, 3)
    vp = np.asarray(viewpoint, dtype=np.float32)
    nrm = np.empty_like(x)
    curv = np.empty(len(x), np.float32)
    lib().oracle_normals(_p(x, C.c_float), len(x), int(k), _p(vp, C.c_float), int(bool(flip_all)),
                         _p(nrm, C.c_float), _p(curv, C.c_float), int(nthreads))
    return nrm, curv


# ---------------------------------------------------------------- learned proposal

ENC_NPARAM = 174470
ENC_LAYOUT = (("W1", (3, 64)), ("b1", (64,)), ("W2", (64, 128)), ("b2", (128,)),
              ("W3", (128, 256)), ("b3", (256,)), ("Hw1", (512, 256)), ("hb1", (256,)),
              ("Hw2", (256, 6)), ("hb2", (6,)))


def enc_forward(w, src, tgt, quantize=True):
    """FP64 forward of the Siamese encoder + head on one pair of normalised clouds.
    Returns (out6, feat512, argmax512)."""
    W = _f64(w)
    s, t = _f32(src).reshape(-1, 3), _f32(tgt).reshape(-1, 3)
    out, feat = np.empty(6), np.empty(512)
    arg = np.empty(512, np.int32)
    lib().oenc_forward(_p(W, C.c_double), _p(s, C.c_float), _p(t, C.c_float), len(s),
                       int(bool(quantize)), _p(out, C.c_double), _p(feat, C.c_double),
                       _p(arg, C.c_int32))
    return out, feat, arg


def enc_loss_grad(w, src, tgt, labels, quantize=True, grad=True):
    """Mean squared loss over a batch (src, tgt: (B, n, 3); labels (B, 6)) and its
    gradient with respect to the flat weights."""
    W = _f64(w)
    s, t = _f32(src), _f32(tgt)
    B, n = s.shape[0], s.shape[1]
    lab = _f64(labels).reshape(B, 6)
    g = np.empty(ENC_NPARAM) if grad else None
    loss = lib().oenc_loss_grad(_p(W, C.c_double), _p(s, C.c_float), _p(t, C.c_float), n, B,
                                _p(lab, C.c_double), int(bool(quantize)),
                                _p(g, C.c_double) if grad else None)
    return (loss, g) if grad else loss


def voxel_grid(xyz, leaf, nrm=None, curv=None, rgb=None):
    """pcl::VoxelGrid restated (lcr_filters_oracle.c).  Returns dict with xyz and,
    when given as input, nrm / curv / rgb of the voxel centroids."""
    x = _f32(xyz).reshape(-1, 3)
    n = len(x)
    nm = _f32(nrm).reshape(-1, 3) if nrm is not None else None
    cv = _f32(curv).reshape(-1) if curv is not None else None
    cl = np.ascontiguousarray(rgb, dtype=np.uint8).reshape(-1, 3) if rgb is not None else None
    ox = np.empty((max(n, 1), 3), np.float32)
    on = np.empty_like(ox) if nm is not None else None
    oc = np.empty(max(n, 1), np.float32) if cv is not None else None
    orgb = np.empty((max(n, 1), 3), np.uint8) if cl is not None else None
    m = lib().oracle_voxel_grid(
        _p(x, C.c_float), _p(nm, C.c_float) if nm is not None else None,
        _p(cv, C.c_float) if cv is not None else None,
        _p(cl, C.c_uint8) if cl is not None else None, n, float(leaf), _p(ox, C.c_float),
        _p(on, C.c_float) if on is not None else None, _p(oc, C.c_float) if oc is not None else None,
        _p(orgb, C.c_uint8) if orgb is not None else None)
    out = {"xyz": ox[:m]}
    if on is not None:
        out["nrm"] = on[:m]
    if oc is not None:
        out["curv"] = oc[:m]
    if orgb is not None:
        out["rgb"] = orgb[:m]
    return out


def sor(xyz, mean_k, mul):
    """pcl::StatisticalOutlierRemoval restated: (keep (n,) bool, per-point mean
    distance (n,) float32, (mean, stddev, threshold))."""
    x = _f32(xyz).reshape(-1, 3)
    keep = np.empty(len(x), np.uint8)
    dist = np.empty(len(x), np.float32)
    stats = np.empty(3)
    lib().oracle_sor(_p(x, C.c_float), len(x), int(mean_k), float(mul), _p(keep, C.c_uint8),
                     _p(dist, C.c_float), _p(stats, C.c_double))
    return keep.astype(bool), dist, tuple(stats)


def lanczos4_weights(fr):
    """The 8 normalised Lanczos-4 tap weights (float32) at fractional offset fr."""
    w = np.empty(8, np.float32)
    lib().oracle_lanczos4_weights(float(fr), _p(w, C.c_float))
    return w


def depth_fuse(frames, factor):
    """super_resolution.cpp:283-314 restated for aligned frames: (H, W) uint16."""
    f = np.ascontiguousarray(frames, dtype=np.uint16)
    k, h, w = f.shape
    out = np.empty((h, w), np.uint16)
    lib().oracle_depth_fuse(_p(f, C.c_uint16), k, w, h, int(factor), _p(out, C.c_uint16))
    return out


def depth_to_cloud(depth, intrinsics, depth_min, depth_max, crop=None, color=None):
    """The depth tool's back-projection restated: (xyz, normals, rgb or None)."""
    d = np.ascontiguousarray(depth, dtype=np.uint16)
    h, w = d.shape
    col = np.ascontiguousarray(color, dtype=np.uint8).reshape(h, w, 3) if color is not None else None
    top, bottom, left, right = crop if crop is not None else (0, h - 1, 0, w - 1)
    ox = np.empty((max(h * w, 1), 3), np.float32)
    on = np.empty_like(ox)
    orgb = np.empty((max(h * w, 1), 3), np.uint8) if col is not None else None
    fx, fy, cx, cy = intrinsics
    m = lib().oracle_depth_to_cloud(
        _p(d, C.c_uint16), w, h, _p(col, C.c_uint8) if col is not None else None, fx, fy, cx, cy,
        int(depth_min), int(depth_max), int(top), int(bottom), int(left), int(right),
        _p(ox, C.c_float), _p(on, C.c_float), _p(orgb, C.c_uint8) if orgb is not None else None)
    return ox[:m], on[:m], (orgb[:m] if orgb is not None else None)
