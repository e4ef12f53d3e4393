"""Training and evaluation of the learned initial-pose proposal.

Training pairs follow the learned-proposal configuration: each pair has its own
random shape (a union of 3 ellipsoids, semi-axes 0.05-0.15 m, as the ellipsoid
pair generator in synth.py), two independent 8192-point area-uniform samples with
1 mm noise, and the source moved by an initial error of U[0, 45 deg] about a
uniform axis plus a translation of length U[0, 0.1 m] in a uniform direction.  The
label is the inverse motion; `label_check` runs plain point-to-plane ICP from each
label on its own pair and measures how far the converged plain-ICP transform (the
label the configuration names) lies from it.  Everything except the random shape and
motion parameters is generated, trained and evaluated on the GPU.

Usage (on a GPU):  python -m lowcost3dreconstruction_amd.proposal_train --pairs 200000
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import synth

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "models", "proposal_c4.npy")

LAYOUT = (("W1", (3, 64)), ("b1", (64,)), ("W2", (64, 128)), ("b2", (128,)),
          ("W3", (128, 256)), ("b3", (256,)), ("Hw1", (512, 256)), ("hb1", (256,)),
          ("Hw2", (256, 6)), ("hb2", (6,)))
NPARAM = 174470


def init_weights(seed=0):
    """He-normal weights, zero biases, a small last layer (near-identity start)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(NPARAM, np.float32)
    off = 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        if name.startswith(("W", "Hw")):
            scale = np.sqrt(2.0 / shape[0]) * (0.01 if name == "Hw2" else 1.0)
            w[off:off + n] = rng.normal(scale=scale, size=n)
        off += n
    return w


def _thomsen_area(a):
    p = 1.6075
    ap = a ** p
    return 4 * np.pi * ((ap[:, 0] * ap[:, 1] + ap[:, 0] * ap[:, 2] + ap[:, 1] * ap[:, 2]) / 3) ** (1 / p)


def random_shapes(rng, n, scale=1.0):
    """(n, 45) float32 ellipsoid unions (c3, R9, a3 per ellipsoid) and (n, 3) area
    shares (Knud Thomsen's surface-area approximation)."""
    axes = rng.uniform(0.05, 0.15, size=(n, 3, 3)) * scale
    centres = rng.uniform(-0.05, 0.05, size=(n, 3, 3)) * scale
    q = rng.normal(size=(n, 3, 4))
    q /= np.linalg.norm(q, axis=-1, keepdims=True)
    w_, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w_), 2 * (x * z + y * w_),
                  2 * (x * y + z * w_), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w_),
                  2 * (x * z - y * w_), 2 * (y * z + x * w_), 1 - 2 * (x * x + y * y)], axis=-1)
    shapes = np.concatenate([centres, R, axes], axis=-1).reshape(n, 45).astype(np.float32)
    areas = np.stack([_thomsen_area(axes[:, e]) for e in range(3)], axis=1)
    return shapes, (areas / areas.sum(1, keepdims=True)).astype(np.float32)


def random_motions(rng, n, max_angle_deg=45.0, max_trans=0.1, min_angle_deg=0.0):
    ax = rng.normal(size=(n, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    ang = np.deg2rad(rng.uniform(min_angle_deg, max_angle_deg, size=n))
    td = rng.normal(size=(n, 3))
    td /= np.linalg.norm(td, axis=1, keepdims=True)
    tl = rng.uniform(0, max_trans, size=n)
    M = np.zeros((n, 4, 4))
    for i in range(n):
        M[i] = synth.rigid(synth.rot_axis_angle(ax[i], ang[i]), td[i] * tl[i])
    return M


def train(lcr, pairs=200_000, npts=8192, epochs=4, batch=256, lr=1e-3, lr_final=1e-4, seed=0,
          weights=None, log=print, time_budget=None):
    """Synthesise `pairs` training pairs on the device and train with Adam; the rate
    decays geometrically from lr to lr_final over 20 segments.  With time_budget
    (seconds of training), the steps of the remaining segments are scaled down
    after the first segment if the projected time exceeds it.  Returns (proposal,
    losses)."""
    rng = np.random.default_rng(seed)
    shapes, shares = random_shapes(rng, pairs)
    motions = random_motions(rng, pairs)
    prop = lcr.Proposal(init_weights(seed) if weights is None else weights)
    t0 = time.time()
    prop.synthesize(shapes, shares, motions, npts, noise=0.001, seed=seed)
    log(f"synthesised {pairs} pairs x {npts} points in {time.time() - t0:.1f} s")
    steps = max(1, epochs * pairs // batch)
    segments = 20
    losses = []
    t0 = time.time()
    scale = 1.0
    for k in range(segments):
        n = steps // segments + (1 if k < steps % segments else 0)
        n = max(1, int(n * scale)) if k > 0 else n
        if n == 0:
            continue
        rate = lr * (lr_final / lr) ** (k / max(segments - 1, 1))
        losses.append(prop.train(n, batch, lr=rate, seed=seed * 1000 + k))
        log(f"segment {k}: {n} steps, lr {rate:.2e}, loss {losses[-1][-50:].mean():.4f}, "
            f"{time.time() - t0:.1f} s")
        if k == 0 and time_budget:
            per_step = (time.time() - t0) / n
            projected = per_step * steps
            if projected > time_budget:
                scale = time_budget / projected
                log(f"projected {projected:.0f} s > budget {time_budget:.0f} s: "
                    f"later segments scaled by {scale:.3f}")
    return prop, np.concatenate(losses)


def make_eval_pairs(lcr, prop, n, npts=8192, seed=1234, min_angle_deg=0.0, max_angle_deg=45.0):
    """Held-out pairs: returns (source clouds, target clouds, ground-truth transforms
    mapping source onto target)."""
    rng = np.random.default_rng(seed)
    shapes, shares = random_shapes(rng, n)
    motions = random_motions(rng, n, max_angle_deg=max_angle_deg, min_angle_deg=min_angle_deg)
    raw, _ = prop.synthesize(shapes, shares, motions, npts, noise=0.001, seed=seed, raw=True)
    T_gt = np.stack([np.linalg.inv(M) for M in motions])
    return raw[:, 0], raw[:, 1], T_gt


def label_check(lcr, prop, n=64, npts=8192, seed=777, max_corr_dist=0.01):
    """The training labels are the inverse synthetic motions; the learned-proposal
    configuration asks for converged plain-ICP transforms.  On `n` pairs drawn as
    the training pairs are, plain point-to-plane ICP (30-NN normals on the device)
    is run from each label on its own pair: the transform it converges to is the
    plain-ICP label.  Returns per pair (rotation difference in degrees, translation
    difference in metres) between the two labels, in the clouds' raw frame, the ICP
    iteration counts, and the network's own error against the label (degrees,
    metres; with the weights `prop` holds)."""
    rng = np.random.default_rng(seed)
    shapes, shares = random_shapes(rng, n)
    motions = random_motions(rng, n)
    raw, frame = prop.synthesize(shapes, shares, motions, npts, noise=0.001, seed=seed, raw=True)
    pred, lab = prop.predict(0, n)

    def rotvec(v):
        a = float(np.linalg.norm(v))
        return synth.rot_axis_angle(v, a) if a > 0 else np.eye(3)

    net_rot = np.array([np.rad2deg(synth.rotation_error(rotvec(p[:3]), rotvec(q[:3])))
                        for p, q in zip(pred.astype(np.float64), lab.astype(np.float64))])
    net_tr = np.linalg.norm(pred[:, 3:] - lab[:, 3:], axis=1) * frame[:, 3]
    T_lab = np.stack([np.linalg.inv(M) for M in motions])
    clouds = [c for i in range(n) for c in (raw[i, 0], raw[i, 1])]
    pairs = [(2 * i, 2 * i + 1) for i in range(n)]
    c = lcr.Cycle(clouds, pairs, guesses=T_lab, point_to_plane=True, normals_k=30,
                  max_corr_dist=max_corr_dist, max_iterations=60)
    c.run()
    r = c.results()
    err = np.array([synth.transform_error(r["T"][i], T_lab[i]) for i in range(n)])
    return np.rad2deg(err[:, 0]), err[:, 1], r["iterations"], net_rot, net_tr


def ablation(lcr, prop, src, tgt, T_gt, max_corr_dist=0.1, max_iterations=60,
             ok_rot_deg=1.0, ok_trans=0.005, hyp_gate=False):
    """Plain point-to-plane ICP from the identity vs the same ICP with the learned
    hypothesis added (both refined and the best by final fitness kept; with hyp_gate
    the learned pose is refined only when it has more inliers at its initial pose
    than the identity).  Returns per-pair success, errors and iterations."""
    n = len(src)
    clouds = [c for pair in zip(src, tgt) for c in pair]
    pairs = [(2 * i, 2 * i + 1) for i in range(n)]
    kw = dict(max_corr_dist=max_corr_dist, max_iterations=max_iterations,
              transformation_epsilon=1e-9, euclidean_fitness_epsilon=1e-3, point_to_plane=True,
              normals_k=30)
    out = {}
    for name, p in (("plain", None), ("learned", prop)):
        c = lcr.Cycle(clouds, pairs, proposal=p, hyp_gate=hyp_gate and p is not None, **kw)
        c.run()
        r = c.results()
        err = np.array([synth.transform_error(r["T"][i], T_gt[i]) for i in range(n)])
        ok = (np.rad2deg(err[:, 0]) < ok_rot_deg) & (err[:, 1] < ok_trans)
        out[name] = dict(success=float(ok.mean()), ok=ok, rot_deg=np.rad2deg(err[:, 0]),
                         trans=err[:, 1], iterations=r["iterations"], fitness=r["fitness"],
                         best_hyp=r["best_hyp"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=200_000)
    ap.add_argument("--points", type=int, default=8192)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--eval-pairs", type=int, default=256)
    ap.add_argument("--time-budget", type=float, default=0.0,
                    help="seconds of training; later segments shrink to fit (0 = no limit)")
    args = ap.parse_args()
    import lowcost3dreconstruction_amd as lcr
    lcr.init(0)
    prop, losses = train(lcr, args.pairs, args.points, args.epochs, args.batch, args.lr,
                         seed=args.seed, log=lambda *a: print(*a, file=sys.stderr, flush=True),
                         time_budget=args.time_budget or None)
    w = prop.get_weights()
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    np.save(args.out, w)
    src, tgt, T_gt = make_eval_pairs(lcr, prop, args.eval_pairs, args.points, seed=args.seed + 777)
    res = ablation(lcr, prop, src, tgt, T_gt)
    summary = {"pairs": args.pairs, "points": args.points, "epochs": args.epochs,
               "steps": int(len(losses)),
               "final_loss": float(losses[-100:].mean()), "first_loss": float(losses[:100].mean()),
               "success_plain": res["plain"]["success"], "success_learned": res["learned"]["success"],
               "iters_plain": float(res["plain"]["iterations"].mean()),
               "iters_learned": float(res["learned"]["iterations"].mean()),
               "learned_chosen": float((res["learned"]["best_hyp"] == 1).mean())}
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
