"""Writes the golden fixtures in tests/golden/*.npz with numpy alone.

The reference holds no golden vectors, known-answer tests or fixtures for this
path (it has no tests at all), and PCL is not installed, so these vectors come
from independent numpy/LAPACK formulations of the same quantities:

* nn.npz       -- exact 1-NN by brute force in float32 with FLANN's distance form
                  ((0 + dx^2) + dy^2) + dz^2, ties broken by the smallest index.
* knn.npz      -- k nearest neighbours by the same distances, lexicographic sort.
* umeyama.npz  -- rigid fits via numpy.linalg.svd (Umeyama 1991), including a
                  reflection case and noiseless known motions.
* eigh.npz     -- numpy.linalg.eigh of random symmetric 3x3 matrices.
* p2pl.npz     -- point-to-plane linear least squares solved by numpy.linalg.lstsq
                  on the stacked rows (not the normal equations), then the
                  Rz(g) Ry(b) Rx(a) matrix.
* normals.npz  -- kNN + covariance + eigh normals with viewpoint flip.

Run: python tests/golden/make_golden.py   (deterministic; seeds fixed below)
"""
import os

import numpy as np

OUT = os.path.dirname(os.path.abspath(__file__))


def flann_d2(q, P):
    q = q.astype(np.float32)
    P = P.astype(np.float32)
    dx = (q[0] - P[:, 0]).astype(np.float32)
    dy = (q[1] - P[:, 1]).astype(np.float32)
    dz = (q[2] - P[:, 2]).astype(np.float32)
    r = (dx * dx).astype(np.float32)
    r = (r + (dy * dy).astype(np.float32)).astype(np.float32)
    r = (r + (dz * dz).astype(np.float32)).astype(np.float32)
    return r


def nn_fixture(rng):
    tgt = rng.uniform(-0.1, 0.1, size=(600, 3)).astype(np.float32)
    tgt[100:110] = tgt[0:10]               # duplicated points: exact ties
    tgt[200] = [np.nan, 0, 0]              # a non-finite point (excluded)
    q = rng.uniform(-0.12, 0.12, size=(300, 3)).astype(np.float32)
    q[:10] = tgt[:10]                      # queries on duplicated points (d2 = 0 ties)
    q[10] = [np.inf, 0, 0]                 # a non-finite query (no neighbour)
    idx = np.full(len(q), -1, np.int32)
    d2 = np.full(len(q), np.inf, np.float32)
    finite_t = np.all(np.isfinite(tgt), axis=1)
    for i, qi in enumerate(q):
        if not np.all(np.isfinite(qi)):
            continue
        d = flann_d2(qi, tgt)
        d[~finite_t] = np.inf
        order = np.lexsort((np.arange(len(tgt)), d))
        idx[i] = order[0]
        d2[i] = d[order[0]]
    np.savez(os.path.join(OUT, "nn.npz"), tgt=tgt, q=q, idx=idx, d2=d2)


def knn_fixture(rng):
    pts = rng.normal(scale=0.05, size=(400, 3)).astype(np.float32)
    pts[50:55] = pts[0:5]
    q = pts[:40].copy()
    k = 12
    idx = np.empty((len(q), k), np.int32)
    d2 = np.empty((len(q), k), np.float32)
    for i, qi in enumerate(q):
        d = flann_d2(qi, pts)
        order = np.lexsort((np.arange(len(pts)), d))[:k]
        idx[i] = order
        d2[i] = d[order]
    np.savez(os.path.join(OUT, "knn.npz"), pts=pts, q=q, k=k, idx=idx, d2=d2)


def umeyama_np(src, dst):
    ms, md = src.mean(0), dst.mean(0)
    S = (dst - md).T @ (src - ms) / len(src)
    U, s, Vt = np.linalg.svd(S)
    D = np.eye(3)
    if np.linalg.det(U) * np.linalg.det(Vt) < 0:
        D[2, 2] = -1
    R = U @ D @ Vt
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = md - R @ ms
    return T


def rand_rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def umeyama_fixture(rng):
    cases_src, cases_dst, cases_T, cases_n, known = [], [], [], [], []
    for c in range(8):
        n = int(rng.integers(3, 200))
        src = rng.normal(scale=0.1, size=(n, 3))
        R = rand_rot(rng)
        t = rng.normal(scale=0.05, size=3)
        noise = 0.0 if c < 3 else 0.002
        dst = src @ R.T + t + rng.normal(scale=noise, size=(n, 3)) if noise else src @ R.T + t
        if c == 7:
            dst = dst * np.array([1, 1, -1])   # a reflected target: det correction
        T = umeyama_np(src, dst)
        Tk = np.eye(4)
        Tk[:3, :3] = R
        Tk[:3, 3] = t
        cases_src.append(src)
        cases_dst.append(dst)
        cases_T.append(T)
        cases_n.append(n)
        known.append(Tk if c < 3 else np.full((4, 4), np.nan))
    np.savez(os.path.join(OUT, "umeyama.npz"), src=np.concatenate(cases_src),
             dst=np.concatenate(cases_dst), n=np.array(cases_n), T=np.stack(cases_T),
             known=np.stack(known))


def eigh_fixture(rng):
    A = []
    for c in range(20):
        M = rng.normal(size=(3, 3))
        S = M @ M.T if c % 2 == 0 else M + M.T
        if c == 4:
            S = np.diag([1.0, 1.0, 2.0])       # repeated eigenvalue
        if c == 6:
            S = np.outer([1.0, 2.0, 3.0], [1.0, 2.0, 3.0]) * 1e-6  # rank 1
        A.append(S)
    A = np.stack(A)
    w, V = np.linalg.eigh(A)
    np.savez(os.path.join(OUT, "eigh.npz"), A=A, w=w, V=V)


def p2pl_matrix(a, b, g, tx, ty, tz):
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rz = np.array([[np.cos(g), -np.sin(g), 0], [np.sin(g), np.cos(g), 0], [0, 0, 1]])
    T = np.eye(4)
    T[:3, :3] = Rz @ Ry @ Rx
    T[:3, 3] = [tx, ty, tz]
    return T


def p2pl_fixture(rng):
    srcs, dsts, nrms, Ts, ns = [], [], [], [], []
    for c in range(6):
        n = int(rng.integers(10, 300))
        dst = rng.normal(scale=0.1, size=(n, 3)) + np.array([0, 0, -0.9])
        nrm = rng.normal(size=(n, 3))
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        src = dst + rng.normal(scale=0.01, size=(n, 3))
        # rows [s x n ; n], residual n.(d - s): lstsq on the stacked system
        A = np.concatenate([np.cross(src, nrm), nrm], axis=1)
        bvec = np.einsum("ij,ij->i", nrm, dst - src)
        x, *_ = np.linalg.lstsq(A, bvec, rcond=None)
        srcs.append(src)
        dsts.append(dst)
        nrms.append(nrm)
        Ts.append(p2pl_matrix(*x))
        ns.append(n)
    np.savez(os.path.join(OUT, "p2pl.npz"), src=np.concatenate(srcs), dst=np.concatenate(dsts),
             nrm=np.concatenate(nrms), n=np.array(ns), T=np.stack(Ts))


def normals_fixture(rng):
    # a noisy patch of a sphere of radius 0.1 around (0, 0, -0.5), seen from the origin
    u = rng.normal(size=(500, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    u = u[u[:, 2] > 0.2]
    pts = (0.1 * u + np.array([0, 0, -0.5]) + rng.normal(scale=3e-4, size=u.shape)).astype(np.float32)
    k = 10
    vp = np.zeros(3)
    nrm = np.empty((len(pts), 3))
    curv = np.empty(len(pts))
    gap = np.empty(len(pts))
    for i, p in enumerate(pts):
        d = flann_d2(p, pts)
        order = np.lexsort((np.arange(len(pts)), d))
        sel = order[:k]
        gap[i] = float(d[order[k]]) - float(d[order[k - 1]])
        X = pts[sel].astype(np.float64)
        C = np.cov(X.T, bias=True)
        w, V = np.linalg.eigh(C)
        n = V[:, 0]
        if np.dot(vp - p.astype(np.float64), n) < 0:
            n = -n
        nrm[i] = n
        curv[i] = abs(w[0]) / w.sum()
    np.savez(os.path.join(OUT, "normals.npz"), pts=pts, k=k, vp=vp, nrm=nrm, curv=curv, gap=gap)


def main():
    rng = np.random.default_rng(20241202)
    nn_fixture(rng)
    knn_fixture(rng)
    umeyama_fixture(rng)
    eigh_fixture(rng)
    p2pl_fixture(rng)
    normals_fixture(rng)
    print("wrote fixtures to", OUT)


if __name__ == "__main__":
    main()
