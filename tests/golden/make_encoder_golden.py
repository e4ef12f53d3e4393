"""Writes tests/golden/encoder.npz: the learned proposal's forward pass computed by
torch (CPU, float64, torch.nn.functional.linear / relu / max) -- an implementation
independent of the C oracle -- for a few seeded pairs, in the plain FP64 form and
with the device's bfloat16 rounding points (h1, h2 and the weights W2, W3 rounded
to bfloat16, round to nearest even from float32, as torch's bfloat16 conversion
does).  The weights are regenerated from the stored seed by `weights(seed)`;
the fixture holds the clouds and the expected outputs, features and argmax.

Run once in the build container:  python tests/golden/make_encoder_golden.py
"""
import os

import numpy as np

LAYOUT = (("W1", (3, 64)), ("b1", (64,)), ("W2", (64, 128)), ("b2", (128,)),
          ("W3", (128, 256)), ("b3", (256,)), ("Hw1", (512, 256)), ("hb1", (256,)),
          ("Hw2", (256, 6)), ("hb2", (6,)))
NPARAM = 174470
SEED = 20261203
PAIRS, POINTS = 3, 96


def weights(seed):
    """Flat float64 weights: uniform He-scaled matrices, uniform biases in
    [-0.05, 0.05] (uniform doubles from the generator's bit stream, which numpy
    keeps stable)."""
    rng = np.random.default_rng(seed)
    w = np.empty(NPARAM)
    off = 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        a = np.sqrt(6.0 / shape[0]) if name.startswith(("W", "Hw")) else 0.05
        w[off:off + n] = rng.uniform(-a, a, size=n)
        off += n
    return w


def clouds(seed):
    rng = np.random.default_rng(seed + 1)
    src = rng.normal(size=(PAIRS, POINTS, 3)).astype(np.float32)
    tgt = rng.normal(size=(PAIRS, POINTS, 3)).astype(np.float32)
    return src, tgt


def torch_forward(w, src, tgt, quantize):
    import torch
    import torch.nn.functional as F
    t = torch.from_numpy(w)
    parts, off = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        parts[name] = t[off:off + n].reshape(shape)
        off += n

    def bf16(x):
        return x.to(torch.float32).to(torch.bfloat16).to(torch.float64) if quantize else x

    def encode(x):
        x = torch.from_numpy(x.astype(np.float64))
        h1 = bf16(F.relu(F.linear(x, parts["W1"].T, parts["b1"])))
        h2 = bf16(F.relu(F.linear(h1, bf16(parts["W2"]).T, parts["b2"])))
        a3 = F.relu(F.linear(h2, bf16(parts["W3"]).T, parts["b3"]))
        f, arg = torch.max(a3, dim=0)  # first index of the maximum
        return f, arg

    outs, feats, args = [], [], []
    for s, q in zip(src, tgt):
        fs, as_ = encode(s)
        ft, at = encode(q)
        g = torch.cat([fs, ft])
        z = F.relu(F.linear(g, parts["Hw1"].T, parts["hb1"]))
        outs.append(F.linear(z, parts["Hw2"].T, parts["hb2"]).numpy())
        feats.append(g.numpy())
        args.append(torch.cat([as_, at]).numpy().astype(np.int32))
    return np.stack(outs), np.stack(feats), np.stack(args)


def main():
    w = weights(SEED)
    src, tgt = clouds(SEED)
    out = {"seed": np.int64(SEED), "src": src, "tgt": tgt}
    for q in (0, 1):
        o, f, a = torch_forward(w, src, tgt, bool(q))
        out[f"out_q{q}"], out[f"feat_q{q}"], out[f"arg_q{q}"] = o, f, a
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "encoder.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
