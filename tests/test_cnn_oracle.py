"""Pins the CPU oracle of the learned-guess network (oracle/cnn_oracle.py): its convolution
against a direct loop, its BF16 rounding against known values, and its backward pass against
central finite differences of its own forward pass (exact mode, float64)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import cnn_oracle as N  # noqa: E402


def test_bf16_rounding_known_values():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties go to even (1.0); just above rounds up
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40])
    exp = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 2.9999732e38, 0.0])
    got = N.bf16(vals)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -7 and got[3] == 1.0 + 2 ** -6
    assert got[4] == -2.5
    assert abs(got[5] - exp[5]) / exp[5] < 2 ** -8
    assert abs(got[6]) < 1e-39


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    H, W, K, Nn = 5, 7, 3, 4
    x = rng.normal(size=(H, W, K))
    Wt = rng.normal(size=(9, Nn, K))
    b = rng.normal(size=Nn)
    out = N.conv_fwd(x, Wt, b)
    ref = np.zeros((H, W, Nn))
    for y in range(H):
        for xx in range(W):
            for n in range(Nn):
                s = b[n]
                for t in range(9):
                    yy, xs = y + t // 3 - 1, xx + t % 3 - 1
                    if 0 <= yy < H and 0 <= xs < W:
                        s += np.dot(x[yy, xs], Wt[t, n])
                ref[y, xx, n] = s
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    H, W = 8, 12
    flat = N.init_params(seed=2)
    x4 = rng.normal(size=(H, W, 4))
    x4[..., 3] = rng.uniform(0, 0.3, size=(H, W))
    t = rng.normal(size=(H, W)) * 0.3
    m = (rng.uniform(size=(H, W)) > 0.2).astype(np.uint8)
    loss, _, g = N.loss_and_grads(flat, x4, t, m, use_bf16=False)
    assert loss > 0
    sizes = [9 * n * k + n for _, k, n in N.LAYERS] + [N.C1, 1]
    starts = np.concatenate([[0], np.cumsum(sizes)])
    picks = []
    for li, (name, k, n) in enumerate(N.LAYERS):
        s0 = starts[li]
        kr = 4 if name == "c1a" else k
        for _ in range(3):     # weights on real input channels
            tap, nn, kk = rng.integers(9), rng.integers(n), rng.integers(kr)
            picks.append(s0 + (tap * n + nn) * k + kk)
        picks.append(s0 + 9 * n * k + rng.integers(n))   # a bias
    picks += [starts[-3] + 5, starts[-2]]                 # head weight, head bias
    for i in picks:
        eps = 1e-6 * max(1.0, abs(flat[i]))
        fp = flat.copy()
        fp[i] += eps
        fm = flat.copy()
        fm[i] -= eps
        lp = N.loss_and_grads(fp, x4, t, m, use_bf16=False)[0]
        lm = N.loss_and_grads(fm, x4, t, m, use_bf16=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (i, fd, g[i])


def test_parameter_count_and_packing_round_trip():
    flat = N.init_params(seed=3)
    assert len(flat) == N.num_params() == 479905
    assert np.array_equal(N.pack(N.unpack(flat)), flat)


def test_device_activation_layout_round_trip():
    """acts_from_device reads the device's concatenated tensors (u1 = up(d2) ++ e1,
    u2 = up(e3) ++ e2) back into per-layer (input, output) pairs: the backward from them equals
    the backward from the forward pass itself."""
    rng = np.random.default_rng(4)
    H, W = 8, 12
    flat = N.init_params(seed=5)
    x4 = rng.normal(size=(H, W, 4))
    y, acts = N.forward(flat, x4)
    A = {k: v for k, v in acts.items()}
    dev = {"x0": A["c1a"][0], "e1a": A["c1a"][1], "u1": A["c5a"][0], "p1": A["c2a"][0], "e2a": A["c2a"][1],
           "u2": A["c4a"][0], "p2": A["c3a"][0], "e3a": A["c3a"][1], "e3": A["c3b"][1], "d2a": A["c4a"][1],
           "d2": A["c4b"][1], "d1a": A["c5a"][1], "d1": A["c5b"][1]}
    back = N.acts_from_device({k: v.ravel() for k, v in dev.items()}, H, W)
    for name in acts:
        assert np.array_equal(back[name][0], acts[name][0]) and np.array_equal(back[name][1], acts[name][1])
    dy = rng.normal(size=(H, W))
    assert np.array_equal(N.backward(flat, back, dy), N.backward(flat, acts, dy))
