"""The oracle's VoxelGrid, StatisticalOutlierRemoval and depth back-projection
(oracle/lcr_filters_oracle.c) pinned by closed forms, brute force and planted
structure (CPU only)."""
import numpy as np

import oracle


def lattice_points(rng, leaf, n_vox=60, per=(1, 6)):
    """Points near voxel centres (within 0.3 leaf), so float and exact floor agree."""
    vox = rng.integers(-20, 20, size=(n_vox, 3))
    vox = np.unique(vox, axis=0)
    pts, owner = [], []
    for v, c in enumerate(vox):
        k = rng.integers(per[0], per[1] + 1)
        pts.append((c + 0.5 + rng.uniform(-0.3, 0.3, size=(k, 3))) * leaf)
        owner += [v] * k
    pts = np.concatenate(pts)
    owner = np.array(owner)
    perm = rng.permutation(len(pts))
    return pts[perm].astype(np.float32), vox, owner[perm]


def test_voxel_grid_centroids_and_order():
    rng = np.random.default_rng(0)
    leaf = 0.01
    pts, vox, owner = lattice_points(rng, leaf)
    nrm = rng.normal(size=pts.shape).astype(np.float32)
    curv = rng.uniform(0, 1, size=len(pts)).astype(np.float32)
    rgb = rng.integers(0, 256, size=pts.shape).astype(np.uint8)
    out = oracle.voxel_grid(pts, leaf, nrm=nrm, curv=curv, rgb=rgb)
    # expected: one centroid per voxel, in increasing index x + y*dx + z*dx*dy of the
    # voxel coordinates relative to the minimum
    mn = vox.min(axis=0)
    d = vox.max(axis=0) - mn + 1
    key = (vox[:, 0] - mn[0]) + (vox[:, 1] - mn[1]) * d[0] + (vox[:, 2] - mn[2]) * d[0] * d[1]
    order = np.argsort(key)
    assert len(out["xyz"]) == len(vox)
    for o, v in enumerate(order):
        sel = owner == v
        np.testing.assert_allclose(out["xyz"][o], pts[sel].astype(np.float64).mean(0), atol=1e-7)
        m = nrm[sel].astype(np.float64).sum(0)
        np.testing.assert_allclose(out["nrm"][o], m / np.linalg.norm(m), atol=2e-6)
        np.testing.assert_allclose(out["curv"][o], curv[sel].mean(), atol=1e-6)
        np.testing.assert_array_equal(out["rgb"][o], np.floor(rgb[sel].mean(0) + 1e-9).astype(np.uint8))


def test_voxel_grid_edges():
    assert len(oracle.voxel_grid(np.full((5, 3), np.nan, np.float32), 0.1)["xyz"]) == 0
    # a lattice of more than INT32_MAX voxels returns the cloud unchanged (PCL)
    p = np.array([[0, 0, 0], [1, 1, 1], [0.5, 0.2, 0.1]], np.float32)
    np.testing.assert_array_equal(oracle.voxel_grid(p, 1e-4)["xyz"], p)
    # NaN points are dropped, the rest voxelised
    q = np.array([[0.01, 0.01, 0.01], [np.nan, 0, 0], [0.015, 0.012, 0.018]], np.float32)
    out = oracle.voxel_grid(q, 0.1)["xyz"]
    np.testing.assert_allclose(out, [[0.0125, 0.011, 0.014]], atol=1e-8)


def test_sor_mean_distances_match_brute_force_and_remove_planted_outliers():
    rng = np.random.default_rng(1)
    plane = np.c_[rng.uniform(0, 1, size=(1500, 2)), rng.normal(scale=1e-3, size=1500)]
    outl = rng.uniform(0, 1, size=(15, 3)) + [0, 0, 0.5]
    pts = np.concatenate([plane, outl]).astype(np.float32)
    pts[7] = np.nan
    k = 20
    keep, dist, (mean, sd, thr) = oracle.sor(pts, k, 3.0)
    fin = np.isfinite(pts).all(1)
    P = pts[fin].astype(np.float64)
    D = np.sqrt(((P[:, None, :] - P[None, :, :]) ** 2).sum(-1))
    D.sort(axis=1)
    ref = D[:, 1:k + 1].mean(1)
    np.testing.assert_allclose(dist[fin], ref, rtol=2e-6)
    assert dist[7] == 0 and keep[7]              # non-finite: value 0, kept
    np.testing.assert_allclose(mean, ref.mean(), rtol=1e-6)
    np.testing.assert_allclose(sd, ref.std(ddof=1), rtol=1e-5)
    assert not keep[1500:].any()                 # every planted outlier removed
    assert keep[:1500][fin[:1500]].mean() > 0.98


def test_depth_to_cloud_closed_forms():
    h, w = 12, 16
    i, j = np.mgrid[0:h, 0:w]
    depth = (1000 + 2 * i + 3 * j).astype(np.uint16)
    fx, fy, cx, cy = 500.0, 400.0, 8.0, 6.0
    xyz, nrm, _ = oracle.depth_to_cloud(depth, (fx, fy, cx, cy), 0, 65535)
    assert len(xyz) == h * w
    xyz = xyz.reshape(h, w, 3)
    nrm = nrm.reshape(h, w, 3)
    # interior: right-left = 2 * 2, down-up = 2 * 3 -> n = (-2, -3, 1) / |.|
    n = np.array([-2.0, -3.0, 1.0]) / np.sqrt(14.0)
    np.testing.assert_allclose(nrm[1:-1, 1:-1], np.broadcast_to(n, (h - 2, w - 2, 3)), atol=1e-7)
    # border pixels: no neighbours -> (0, 0, 1)
    np.testing.assert_array_equal(nrm[0, 0], [0, 0, 1])
    # the principal point back-projects onto the optical axis
    np.testing.assert_array_equal(xyz[6, 8, :2], [0, 0])
    assert xyz[6, 8, 2] == -float(depth[6, 8])
    z = float(depth[3, 11])
    np.testing.assert_allclose(xyz[3, 11], [z * 3 / fx, -z * -3 / fy, -z], rtol=1e-7)
    # a step of more than 20 mm drops that neighbour
    d2 = np.full((5, 5), 1000, np.uint16)
    d2[3:, :] = 1030
    _, n2, _ = oracle.depth_to_cloud(d2, (500, 500, 2, 2), 0, 65535)
    n2 = n2.reshape(5, 5, 3)
    # pixel (2, 2): "right" = (3, 2) differs by 30 -> 0; left = 1000 -> dzdx = -500
    v = np.array([500.0, 0.0, 1.0])
    np.testing.assert_allclose(n2[2, 2], v / np.linalg.norm(v), atol=1e-7)
    # range and crop filters
    xyz_c, _, _ = oracle.depth_to_cloud(depth, (fx, fy, cx, cy), 1010, 1020, crop=(2, 9, 3, 12))
    zz = depth[2:10, 3:13]
    assert len(xyz_c) == int(((zz >= 1010) & (zz <= 1020)).sum())
