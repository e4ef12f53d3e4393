"""Pins the FP64 encoder oracle (oracle/lcr_encoder_oracle.c): its analytic
backward against central finite differences of its own FP64 loss, its layer
shapes and its max-pool tie rule.  CPU only."""
import numpy as np

import oracle


def init_weights(rng):
    w = np.empty(oracle.ENC_NPARAM)
    off = 0
    for name, shape in oracle.ENC_LAYOUT:
        n = int(np.prod(shape))
        if name.startswith(("W", "Hw")):
            w[off:off + n] = rng.normal(scale=np.sqrt(2.0 / shape[0]), size=n)
        else:
            w[off:off + n] = rng.normal(scale=0.05, size=n)
        off += n
    assert off == oracle.ENC_NPARAM == 174470
    return w


def test_parameter_count_matches_layout():
    assert sum(int(np.prod(s)) for _, s in oracle.ENC_LAYOUT) == 174470


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = init_weights(rng)
    B, n = 2, 48
    src = rng.normal(size=(B, n, 3)).astype(np.float32)
    tgt = rng.normal(size=(B, n, 3)).astype(np.float32)
    lab = rng.normal(size=(B, 6))
    loss, g = oracle.enc_loss_grad(w, src, tgt, lab, quantize=False)
    # probe coordinates from every layer, including the largest-gradient entries
    off, probes = 0, []
    for name, shape in oracle.ENC_LAYOUT:
        m = int(np.prod(shape))
        blk = np.abs(g[off:off + m])
        probes += list(off + np.argsort(blk)[-3:]) + list(off + rng.integers(0, m, 3))
        off += m
    eps = 1e-6
    for i in probes:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        fd = (oracle.enc_loss_grad(wp, src, tgt, lab, quantize=False, grad=False) -
              oracle.enc_loss_grad(wm, src, tgt, lab, quantize=False, grad=False)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, g[i])


def test_maxpool_ties_pick_lowest_index_and_relu_zero():
    rng = np.random.default_rng(1)
    w = init_weights(rng)
    pts = rng.normal(size=(10, 3)).astype(np.float32)
    dup = np.concatenate([pts, pts])  # every maximum appears twice
    _, _, arg = oracle.enc_forward(w, dup, dup, quantize=True)
    assert np.all(arg < 10)
    # a channel with a negative bias and zero weights is relu(<0) = 0 everywhere
    w2 = w.copy()
    s = 0
    for name, shape in oracle.ENC_LAYOUT:
        m = int(np.prod(shape))
        if name == "W3":
            w2[s:s + m].reshape(shape)[:, 7] = 0.0
        if name == "b3":
            w2[s + 7] = -1.0
        s += m
    _, feat, arg = oracle.enc_forward(w2, pts, pts)
    assert feat[7] == 0.0 and arg[7] == 0


def test_forward_matches_torch_fixture():
    """The forward pass against torch's float64 linear / relu / max on seeded
    pairs (tests/golden/make_encoder_golden.py), plain and with the device's
    bfloat16 rounding points: outputs and features to 1e-12 relative (the FP64 sums
    run in another order), argmax equal.  A transposed weight block, a dropped bias
    or a rounding at the wrong point fails this."""
    import importlib.util
    import os

    from conftest import GOLDEN
    spec = importlib.util.spec_from_file_location(
        "make_encoder_golden", os.path.join(GOLDEN, "make_encoder_golden.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    g = np.load(os.path.join(GOLDEN, "encoder.npz"), allow_pickle=False)
    w = gen.weights(int(g["seed"]))
    for q in (0, 1):
        for i in range(len(g["src"])):
            out, feat, arg = oracle.enc_forward(w, g["src"][i], g["tgt"][i], quantize=bool(q))
            np.testing.assert_allclose(out, g[f"out_q{q}"][i], rtol=1e-12, atol=1e-13)
            np.testing.assert_allclose(feat, g[f"feat_q{q}"][i], rtol=1e-12, atol=1e-13)
            np.testing.assert_array_equal(arg, g[f"arg_q{q}"][i])


def test_quantized_forward_close_to_fp64():
    rng = np.random.default_rng(2)
    w = init_weights(rng)
    src = rng.normal(size=(256, 3)).astype(np.float32)
    tgt = rng.normal(size=(256, 3)).astype(np.float32)
    a, _, _ = oracle.enc_forward(w, src, tgt, quantize=False)
    b, _, _ = oracle.enc_forward(w, src, tgt, quantize=True)
    assert np.max(np.abs(a - b)) < 0.05 * max(1.0, np.max(np.abs(a)))
