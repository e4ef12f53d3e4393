"""The oracle's VoxelGrid, StatisticalOutlierRemoval, depth back-projection and
burst depth fusion (oracle/lcr_filters_oracle.c) pinned by closed forms, brute
force, numpy restatements and planted structure (CPU only)."""
import numpy as np
import pytest

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


def test_lanczos4_weights_are_the_windowed_sinc():
    """Tap weights against numpy's sinc: sinc(d) sinc(d / 4), d = fr + 3 - i,
    normalised; fr = 0 is the unit tap; fr = 0.5 is symmetric."""
    for fr in (0.5, 0.25, 0.1, 0.9, 1e-3):
        d = fr + 3.0 - np.arange(8)
        c = np.sinc(d) * np.sinc(d / 4.0)
        ref = (c / c.sum()).astype(np.float32)
        w = oracle.lanczos4_weights(fr)
        np.testing.assert_allclose(w, ref, rtol=2e-7, atol=1e-9)
        assert abs(float(w.astype(np.float64).sum()) - 1.0) < 1e-6
    np.testing.assert_array_equal(oracle.lanczos4_weights(0.0), [0, 0, 0, 1, 0, 0, 0, 0])
    w = oracle.lanczos4_weights(0.5)
    np.testing.assert_array_equal(w, w[::-1])


def fuse_numpy(frames, factor):
    """super_resolution.cpp:283-314 written with numpy: ZE mean (double sums of the
    positive values), the nearest-upsampled image by np.repeat, then each output
    pixel as the separable 8-tap sum over the clamped upsampled indices, horizontal
    first, float32 throughout, then round half to even, clamp, NaN -> 0."""
    f = frames.astype(np.float32)
    pos = f > 0
    cnt = pos.sum(0)
    s = np.where(pos, f.astype(np.float64), 0.0).sum(0)
    with np.errstate(invalid="ignore", divide="ignore"):
        mean = np.where(cnt > 0, (s / np.maximum(cnt, 1)).astype(np.float32), np.float32(np.nan))
    up = np.repeat(np.repeat(mean, factor, axis=0), factor, axis=1).astype(np.float32)
    H, W = up.shape
    h, w = mean.shape
    p = (factor - 1) / 2.0
    base = int(np.floor(p))
    wt = oracle.lanczos4_weights(p - base)
    ys = factor * np.arange(h) + base
    xs = factor * np.arange(w) + base
    acc = np.zeros((h, w), np.float32)
    for ky in range(8):
        rows = np.clip(ys - 3 + ky, 0, H - 1)
        hs = np.zeros((h, w), np.float32)
        for kx in range(8):
            cols = np.clip(xs - 3 + kx, 0, W - 1)
            hs = (hs + wt[kx] * up[rows][:, cols]).astype(np.float32)
        acc = (acc + wt[ky] * hs).astype(np.float32)
    out = np.where(np.isnan(acc), 0.0, np.clip(np.rint(acc), 0, 65535))
    return out.astype(np.uint16)


def test_depth_fuse_closed_forms():
    rng = np.random.default_rng(12)
    # a constant burst stays constant; zeros are eliminated from the mean
    f = np.full((6, 20, 24), 1234, np.uint16)
    f[rng.random(f.shape) < 0.3] = 0
    f[:, 0, 0] = 1234
    np.testing.assert_array_equal(oracle.depth_fuse(f, 2), np.full((20, 24), 1234, np.uint16))
    # the mean of the positive values: frames 1000 and 1003 where present -> 1001.5,
    # rounded half to even -> 1002 with the unit taps of factor 1; 1000 and 1004 ->
    # 1002 with the normalised float taps of factor 2 (they sum to 1 within 1e-7)
    g = np.zeros((4, 16, 16), np.uint16)
    g[0], g[2] = 1000, 1003
    np.testing.assert_array_equal(oracle.depth_fuse(g, 1), np.full((16, 16), 1002, np.uint16))
    g[2] = 1004
    np.testing.assert_array_equal(oracle.depth_fuse(g, 2), np.full((16, 16), 1002, np.uint16))
    # a pixel without any depth poisons the taps that reach it (NaN -> 0)
    h_ = np.full((3, 16, 16), 900, np.uint16)
    h_[:, 8, 8] = 0
    out = oracle.depth_fuse(h_, 2)
    assert out[8, 8] == 0 and out[0, 0] == 900 and out[15, 15] == 900
    hole = np.argwhere(out == 0)
    assert hole[:, 0].min() >= 6 and hole[:, 0].max() <= 10


@pytest.mark.parametrize("factor", [1, 2, 3, 4])
def test_depth_fuse_matches_numpy(factor):
    rng = np.random.default_rng(factor)
    base = 800 + 200 * np.sin(np.linspace(0, 3, 40))[:, None] * np.cos(np.linspace(0, 2, 56))[None, :]
    f = (base[None] + rng.normal(scale=4.0, size=(8, 40, 56))).round().astype(np.uint16)
    f[rng.random(f.shape) < 0.2] = 0
    f[:, 30:33, 10:12] = 0
    np.testing.assert_array_equal(oracle.depth_fuse(f, factor), fuse_numpy(f, factor))
