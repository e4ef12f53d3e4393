"""Seeded synthetic inputs for the fine-registration workloads (numpy only).

The reference ships no data (its one dataset is a remote download, README.md:163),
so every workload is synthesised here:

* ``pair_ellipsoids``  -- 20k-point resamples of a closed surface (the boundary of a
  union of 3 ellipsoids), one moved by 10 deg about a random axis + 2 cm.
* ``render_depth`` / ``depth_cloud`` -- a depth map of the same kind of object seen
  by a Kinect-style camera, back-projected with the capture tools' convention
  (x = z(j-cx)/fx, y = -z(i-cy)/fy, z_out = -z; capture/depth_capture/depth_capture.cpp:297-305)
  and voxel-downsampled with VoxelGrid centroid semantics
  (pcl_tools/cloud_downsampling.cpp:73-75).
* ``turntable_cycle`` -- N captures at fixed yaw steps about the turntable axis (the
  camera Y axis through the object centre, rotate_align/rotate_align.cpp:232), each
  with a random half-space occlusion crop; pairs (i -> i-1) as alignment.sh pairs
  captures (scripts/alignment.sh:106-113), plus the closing pair.

All randomness comes from ``numpy.random.default_rng(seed)`` (PCG64), so a seed gives
the same clouds on any machine with this numpy.  Units are metres.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def rot_axis_angle(axis, angle):
    """Rotation matrix for a unit axis and an angle in radians (Rodrigues)."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * (K @ K)


def random_unit(rng, n=None):
    v = rng.normal(size=(3,) if n is None else (n, 3))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def rigid(R, t):
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = t
    return T


def apply(T, pts):
    pts = np.asarray(pts, dtype=np.float64)
    return pts @ T[:3, :3].T + T[:3, 3]


@dataclass
class Shape:
    """Union of ellipsoids: centre c_i, rotation R_i, semi-axes a_i."""
    centres: np.ndarray
    rots: np.ndarray
    axes: np.ndarray

    def inside(self, x, skip=None):
        """Boolean mask: strictly inside any ellipsoid other than `skip`."""
        m = np.zeros(len(x), dtype=bool)
        for i in range(len(self.axes)):
            if skip is not None and i == skip:
                continue
            local = (x - self.centres[i]) @ self.rots[i] / self.axes[i]
            m |= np.einsum("ij,ij->i", local, local) < 1.0
        return m


def make_shape(rng, n_ell=3, axis_range=(0.05, 0.15), centre_range=0.05, scale=1.0):
    axes = rng.uniform(*axis_range, size=(n_ell, 3)) * scale
    rots = np.stack([random_rotation(rng) for _ in range(n_ell)])
    centres = rng.uniform(-centre_range, centre_range, size=(n_ell, 3)) * scale
    return Shape(centres, rots, axes)


def _area_factor(u, a):
    # dA_ellipsoid / dA_sphere at unit-sphere point u for axis-aligned semi-axes a
    return np.prod(a) * np.sqrt(np.sum((u / a) ** 2, axis=1))


def sample_surface(shape, n, rng):
    """Area-uniform samples on the boundary of the union of ellipsoids."""
    k = len(shape.axes)
    probe = random_unit(rng, 4096)
    areas = np.array([4 * np.pi * _area_factor(probe, shape.axes[i]).mean() for i in range(k)])
    w = areas / areas.sum()
    out = []
    have = 0
    while have < n:
        m = max(2 * (n - have), 1024)
        which = rng.choice(k, size=m, p=w)
        u = random_unit(rng, m)
        keep = np.zeros(m, dtype=bool)
        pts = np.empty((m, 3))
        acc = rng.uniform(size=m)
        for i in range(k):
            sel = which == i
            a = shape.axes[i]
            ui = u[sel]
            prob = _area_factor(ui, a) / (np.prod(a) / a.min())
            ok = acc[sel] < prob
            p = (ui * a) @ shape.rots[i].T + shape.centres[i]
            ok &= ~shape.inside(p, skip=i)
            pts[sel] = p
            keep[sel] = ok
        got = pts[keep]
        out.append(got)
        have += len(got)
    return np.concatenate(out)[:n]


@dataclass
class Pair:
    src: np.ndarray          # (Ns, 3) float32
    tgt: np.ndarray          # (Nt, 3) float32
    T_gt: np.ndarray         # 4x4: maps src onto tgt
    src_nrm: np.ndarray = None
    tgt_nrm: np.ndarray = None
    meta: dict = field(default_factory=dict)


def pair_ellipsoids(seed, n_src=20000, n_tgt=20000, angle_deg=10.0, trans=0.02, noise=0.001,
                    shape=None):
    """Two independent area-uniform resamples of one ellipsoid-union surface; the
    target is moved by `angle_deg` about a uniform random axis plus `trans` metres in
    a uniform random direction; both get N(0, noise^2) per coordinate."""
    rng = np.random.default_rng(seed)
    if shape is None:
        shape = make_shape(rng)
    src = sample_surface(shape, n_src, rng)
    tgt0 = sample_surface(shape, n_tgt, rng)
    R = rot_axis_angle(random_unit(rng), np.deg2rad(angle_deg))
    t = trans * random_unit(rng)
    T = rigid(R, t)
    tgt = apply(T, tgt0)
    src = src + rng.normal(scale=noise, size=src.shape)
    tgt = tgt + rng.normal(scale=noise, size=tgt.shape)
    return Pair(src.astype(np.float32), tgt.astype(np.float32), T,
                meta=dict(seed=seed, kind="ellipsoid_pair"))


# --------------------------------------------------------------------------- depth

@dataclass
class Camera:
    width: int = 640
    height: int = 480
    fx: float = 525.0
    fy: float = 525.0
    cx: float = 319.5
    cy: float = 239.5


KINECT_V1 = Camera()
# capture/super_resolution_kv2/super_resolution_kv2.cpp:100
KINECT_V2 = Camera(512, 424, 365.531799, 365.531799, 256.136810, 206.013901)


def render_depth(shape, T_obj, cam=KINECT_V1):
    """Depth (positive metres along -Z, 0 = no hit) of the union of ellipsoids placed
    by T_obj (object -> camera frame), by closed-form ray/ellipsoid intersection."""
    j, i = np.meshgrid(np.arange(cam.width), np.arange(cam.height))
    d = np.stack([(j - cam.cx) / cam.fx, -(i - cam.cy) / cam.fy, -np.ones_like(j, float)],
                 axis=-1).reshape(-1, 3)
    R, t = T_obj[:3, :3], T_obj[:3, 3]
    best = np.full(len(d), np.inf)
    for e in range(len(shape.axes)):
        C = R @ shape.centres[e] + t
        M = R @ shape.rots[e]                      # columns: ellipsoid axes in camera
        A = (M / 1.0).T / shape.axes[e][:, None]   # A x = local unit-sphere coordinates
        Ad = d @ A.T
        AC = A @ C
        a = np.einsum("ij,ij->i", Ad, Ad)
        b = -2.0 * Ad @ AC
        c = AC @ AC - 1.0
        disc = b * b - 4 * a * c
        ok = disc > 0
        s = np.full(len(d), np.inf)
        sq = np.sqrt(np.where(ok, disc, 0.0))
        s0 = (-b - sq) / (2 * a)
        s[ok & (s0 > 0)] = s0[ok & (s0 > 0)]
        best = np.minimum(best, s)
    depth = np.where(np.isfinite(best), best, 0.0)   # d_z = -1, so depth = s
    return depth.reshape(cam.height, cam.width)


def backproject(depth, cam=KINECT_V1):
    """capture/depth_capture/depth_capture.cpp:297-305: x = z(j-cx)/fx, y = -z(i-cy)/fy,
    z_out = -z, for pixels with z != 0."""
    i, j = np.nonzero(depth > 0)
    z = depth[i, j]
    return np.stack([z * (j - cam.cx) / cam.fx, -z * (i - cam.cy) / cam.fy, -z], axis=1)


def voxel_downsample(pts, leaf):
    """pcl::VoxelGrid centroid semantics (cloud_downsampling.cpp:73-75): voxel index
    floor(p / leaf) - min index, one centroid per occupied voxel, output ordered by
    voxel index (x fastest)."""
    if len(pts) == 0:
        return pts
    ijk = np.floor(pts / leaf).astype(np.int64)
    ijk -= ijk.min(axis=0)
    dims = ijk.max(axis=0) + 1
    key = ijk[:, 0] + dims[0] * (ijk[:, 1] + dims[1] * ijk[:, 2])
    uniq, inv = np.unique(key, return_inverse=True)
    cnt = np.bincount(inv).astype(np.float64)
    out = np.stack([np.bincount(inv, weights=pts[:, a]) / cnt for a in range(3)], axis=1)
    return out


def turntable_prior(yaw_step_deg, centre):
    """The rig's prior for a pair (capture i onto capture i-1): a rotation by
    -yaw_step about the vertical axis through the turntable centre, as
    rotate_align/rotate_align.cpp:224-235 applies per capture."""
    P = rigid(np.eye(3), np.asarray(centre, dtype=np.float64))
    return P @ rigid(rot_axis_angle([0, 1, 0], np.deg2rad(-yaw_step_deg)), np.zeros(3)) @ np.linalg.inv(P)


def turntable_pose(yaw, centre, jitter=np.zeros(3)):
    """Object -> camera: rotate by `yaw` about the vertical (camera Y) axis through
    the object centre, then place at `centre` (+ jitter)."""
    return rigid(rot_axis_angle([0, 1, 0], yaw), np.asarray(centre) + jitter)


def depth_cloud(shape, T_obj, cam=KINECT_V1, noise=lambda z: 0.0012 * z * z, leaf=0.002, rng=None):
    depth = render_depth(shape, T_obj, cam)
    if rng is not None:
        m = depth > 0
        depth[m] = depth[m] + rng.normal(size=m.sum()) * noise(depth[m])
    pts = backproject(depth, cam)
    if leaf:
        pts = voxel_downsample(pts, leaf)
    return pts


def occlusion_crop(pts, frac, rng):
    """Remove `frac` of the points with a random half-space through the cloud."""
    if frac <= 0:
        return pts
    nrm = random_unit(rng)
    proj = pts @ nrm
    thr = np.quantile(proj, frac)
    return pts[proj >= thr]


def pair_depth(seed, yaw_step_deg=10.0, jitter=0.005, centre=(0.0, 0.0, -0.9), leaf=0.002,
               cam=KINECT_V1, scale=1.0):
    """Two depth captures of one object differing by one turntable yaw step plus a
    random translation jitter (metres); source = the later capture."""
    rng = np.random.default_rng(seed)
    shape = make_shape(rng, scale=scale)
    T0 = turntable_pose(0.0, centre)
    T1 = turntable_pose(np.deg2rad(yaw_step_deg), centre, jitter * random_unit(rng))
    tgt = depth_cloud(shape, T0, cam, leaf=leaf, rng=rng)
    src = depth_cloud(shape, T1, cam, leaf=leaf, rng=rng)
    T_gt = T0 @ np.linalg.inv(T1)
    return Pair(src.astype(np.float32), tgt.astype(np.float32), T_gt,
                meta=dict(seed=seed, kind="depth_pair"))


@dataclass
class Cycle:
    clouds: list               # list of (N_i, 3) float32
    pairs: np.ndarray          # (P, 2) int32: (source cloud, target cloud)
    T_gt: np.ndarray           # (P, 4, 4): maps source onto target
    poses: np.ndarray          # (N, 4, 4): object -> camera per capture
    meta: dict = field(default_factory=dict)


def tilt_prior(tilt_deg, centre):
    """Prior for a top (tilt > 0) or bottom (tilt < 0) view onto capture 0: the
    inverse of the tilt about the horizontal camera axis through the centre."""
    P = rigid(np.eye(3), np.asarray(centre, dtype=np.float64))
    return P @ rigid(rot_axis_angle([1, 0, 0], np.deg2rad(-tilt_deg)), np.zeros(3)) @ np.linalg.inv(P)


def turntable_cycle(seed, n_captures=36, yaw_step_deg=10.0, crop=(0.05, 0.60), jitter=0.005,
                    centre=(0.0, 0.0, -0.9), leaf=0.002, cam=KINECT_V1, scale=1.0,
                    noise=lambda z: 0.0012 * z * z, closing_pair=True, top_bottom=False,
                    tilt_deg=45.0):
    """A turntable capture cycle: capture i at yaw i*step (+ translation jitter), each
    cropped by a random half-space removing a U[crop] fraction.  Pairs are (i, i-1)
    for i = 1..N-1 (alignment.sh's order) plus (0, N-1) when closing_pair.  With
    top_bottom, two more captures see the object tilted by +-tilt_deg about the
    horizontal axis; each is paired onto capture 0 (centroid_align.cpp:104-130 aligns
    the top and bottom views onto the first capture)."""
    rng = np.random.default_rng(seed)
    shape = make_shape(rng, scale=scale)
    poses, clouds = [], []
    for i in range(n_captures):
        jit = jitter * random_unit(rng) if i > 0 else np.zeros(3)
        T = turntable_pose(np.deg2rad(yaw_step_deg * i), centre, jit)
        pts = depth_cloud(shape, T, cam, noise=noise, leaf=leaf, rng=rng)
        if crop is not None:
            pts = occlusion_crop(pts, rng.uniform(*crop), rng)
        poses.append(T)
        clouds.append(pts.astype(np.float32))
    pairs = [(i, i - 1) for i in range(1, n_captures)]
    if closing_pair:
        pairs.append((0, n_captures - 1))
    prior = turntable_prior(yaw_step_deg, centre)
    priors = [prior] * len(pairs)
    if top_bottom:
        for tilt in (tilt_deg, -tilt_deg):
            jit = jitter * random_unit(rng)
            T = rigid(rot_axis_angle([1, 0, 0], np.deg2rad(tilt)), np.asarray(centre) + jit)
            pts = depth_cloud(shape, T, cam, noise=noise, leaf=leaf, rng=rng)
            if crop is not None:
                pts = occlusion_crop(pts, rng.uniform(*crop), rng)
            poses.append(T)
            clouds.append(pts.astype(np.float32))
            pairs.append((len(clouds) - 1, 0))
            priors.append(tilt_prior(tilt, centre))
    pairs = np.asarray(pairs, dtype=np.int32)
    T_gt = np.stack([poses[t] @ np.linalg.inv(poses[s]) for s, t in pairs])
    return Cycle(clouds, pairs, T_gt, np.stack(poses),
                 meta=dict(seed=seed, prior=np.stack(priors)))


def hypotheses(cycle, n_hyp, max_angle_deg, max_trans, seed):
    """(P * n_hyp, 4, 4) initial poses, hypothesis-major within a pair: the pair's
    prior, then n_hyp - 1 copies moved by a random rotation of up to max_angle_deg
    about a uniform axis through the prior-mapped source centroid, and a
    translation of up to max_trans."""
    rng = np.random.default_rng(seed)
    out = []
    for (s, _), P in zip(cycle.pairs, cycle.meta["prior"]):
        c = P[:3, :3] @ cycle.clouds[s].astype(np.float64).mean(axis=0) + P[:3, 3]
        C = rigid(np.eye(3), c)
        out.append(P)
        for _ in range(n_hyp - 1):
            out.append(C @ perturb(rng, max_angle_deg, max_trans) @ np.linalg.inv(C) @ P)
    return np.stack(out)


def perturb(rng, max_angle_deg, max_trans):
    """A random rigid increment: angle U[0, max] about a uniform axis, translation
    with uniform direction and length U[0, max_trans]."""
    R = rot_axis_angle(random_unit(rng), np.deg2rad(rng.uniform(0, max_angle_deg)))
    t = random_unit(rng) * rng.uniform(0, max_trans)
    return rigid(R, t)


# --------------------------------------------------------------------------- metrics

def rotation_error(Ra, Rb):
    """Geodesic angle of Ra^T Rb via 2 asin(||Ra - Rb||_F / (2 sqrt 2)), stable for
    small angles."""
    f = np.linalg.norm(np.asarray(Ra)[:3, :3] - np.asarray(Rb)[:3, :3])
    return 2.0 * np.arcsin(min(1.0, f / (2.0 * np.sqrt(2.0))))


def transform_error(Ta, Tb):
    """(rotation error in radians, translation error in metres)."""
    Ta, Tb = np.asarray(Ta), np.asarray(Tb)
    return rotation_error(Ta, Tb), float(np.linalg.norm(Ta[:3, 3] - Tb[:3, 3]))
