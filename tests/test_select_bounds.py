"""The bounds the normals selection trims its rows to (csrc/normals.hip): a point
whose histogram key floor(d2 * scale) is at most B must lie below the float32
bound the kernels compute from B, or the trimmed rows could drop one of the k
nearest.  numpy's float32 arithmetic rounds as the device code does (round to
nearest, no contraction: the kernels build with -ffp-contract=off), so the checks
restate the kernels' expressions operation by operation over random values and
values at the bucket edges.

- lane pass (lane_query): scale = NB / D2, key min(int(d2 * scale), NB - 1), bound
  min(D2, D2 * ((B + 1) / NB) * 1.0001) for B < NB - 1;
- wavefront pass (k_normal_wave): the same with 256 buckets over R2, and with the
  split of bucket B into 256 sub-buckets, sub-key min(int((x - B) * 256), 255),
  bound min(D2b, R2 * ((B + (SB + 1) / 256) / 256) * 1.0001);
- a carried bound U = Ub * 1.0001 stays above the point's own d2 Ub."""
import numpy as np

f32 = np.float32
NB_LANE, NB_WAVE = 32, 256


def samples(rng, n):
    """(R2, d2) pairs: R2 over many magnitudes, d2 in [0, R2) uniformly and at the
    float neighbours of every bucket edge k * R2 / nb."""
    R2 = (10.0 ** rng.uniform(-8, 3, n)).astype(f32)
    d_uni = (R2 * rng.uniform(0, 1, n).astype(f32)).astype(f32)
    out = [(R2, d_uni)]
    for nb in (NB_LANE, NB_WAVE):
        k = rng.integers(1, nb, n).astype(f32)
        edge = (R2 * (k / f32(nb))).astype(f32)
        for direction in (-np.inf, np.inf):
            d = edge
            for _ in range(3):
                d = np.nextafter(d, f32(direction)).astype(f32)
                out.append((R2, d))
    R2 = np.concatenate([o[0] for o in out])
    d = np.concatenate([o[1] for o in out])
    keep = (d >= 0) & (d < R2)
    return R2[keep], d[keep]


def lane_key_and_bound(R2, d, nb, widen=f32(1.0001)):
    scale = (f32(nb) / R2).astype(f32)
    x = (d * scale).astype(f32)
    b = np.minimum(x.astype(np.int64), nb - 1)
    ratio = ((b + 1).astype(f32) / f32(nb)).astype(f32)
    bound = np.minimum(R2, ((R2 * ratio).astype(f32) * widen).astype(f32))
    return x, b, bound


def test_bucket_bounds_hold_the_bucket():
    rng = np.random.default_rng(0)
    R2, d = samples(rng, 200_000)
    for nb in (NB_LANE, NB_WAVE):
        x, b, bound = lane_key_and_bound(R2, d, nb)
        inner = b < nb - 1  # the last bucket keeps the whole range
        assert np.all(d[inner] < bound[inner])
        # the bound is tight enough to matter: on average it trims to (B + 1) / nb
        assert np.median(bound[inner] / R2[inner]) < 0.6


def test_edge_samples_reach_the_rounding():
    """Without the 1e-4 widening the bound misses a point at a bucket edge: the
    edge samples reach the cases the widening is for."""
    rng = np.random.default_rng(0)
    R2, d = samples(rng, 200_000)
    missed = 0
    for nb in (NB_LANE, NB_WAVE):
        x, b, bound = lane_key_and_bound(R2, d, nb, widen=f32(1.0))
        inner = b < nb - 1
        missed += int(np.sum(d[inner] >= bound[inner]))
    assert missed > 0


def test_sub_bucket_bounds_hold_the_sub_bucket():
    rng = np.random.default_rng(1)
    R2, d = samples(rng, 200_000)
    x, b, D2b = lane_key_and_bound(R2, d, NB_WAVE)
    inner = b < NB_WAVE - 1
    R2, d, x, b, D2b = R2[inner], d[inner], x[inner], b[inner], D2b[inner]
    bf = b.astype(f32)
    frac = (x - bf).astype(f32)  # exact: x lies in [B, B + 1)
    assert np.all(frac.astype(np.float64) == x.astype(np.float64) - b.astype(np.float64))
    sb = np.minimum((frac * f32(256)).astype(f32).astype(np.int64), 255)
    inner_sb = ((sb + 1).astype(f32) * f32(1.0 / 256.0)).astype(f32)
    v = ((bf + inner_sb).astype(f32) / f32(NB_WAVE)).astype(f32)
    D2s = np.minimum(D2b, ((R2 * v).astype(f32) * f32(1.0001)).astype(f32))
    assert np.all(d < D2s)


def test_carried_bound_stays_above_its_point():
    rng = np.random.default_rng(2)
    Ub = (10.0 ** rng.uniform(-29, 3, 500_000)).astype(f32)
    assert np.all((Ub * f32(1.0001)).astype(f32) > Ub)
