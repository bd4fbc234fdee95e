"""Learned-guess network on the GPU: matrix-core operand layout, forward output, loss and every
parameter gradient against the CPU oracle (oracle/cnn_oracle.py, same BF16 rounding points),
and training by the device kernels."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))


@pytest.fixture(scope="module")
def ctx():
    from starfish_amd.plugin import Context
    c = Context(0)
    yield c
    c.close()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.gpu
def test_mfma_operand_layout(ctx):
    import cnn_oracle as N
    from starfish_amd.learned import mfma_probe
    rng = np.random.default_rng(0)
    a = rng.normal(size=(16, 32))
    b = rng.normal(size=(32, 16))
    d = mfma_probe(ctx, a, b)
    ref = N.bf16(a) @ N.bf16(b)
    assert np.abs(d - ref).max() <= 1e-5 * np.abs(ref).max()


def _synthetic_sample(rng, H, W):
    x4 = np.zeros((H, W, 4))
    x4[..., 0] = rng.uniform(0, 1, (H, W))
    x4[..., 1] = (rng.uniform(size=(H, W)) > 0.9)
    x4[..., 2] = np.where(x4[..., 1] > 0, rng.uniform(-1, 0, (H, W)), 0)
    x4[..., 3] = rng.uniform(0, 0.3, (H, W))
    t = rng.normal(size=(H, W)) * 0.2
    m = (x4[..., 1] == 0).astype(np.uint8)
    return x4, t, m


@pytest.mark.gpu
@pytest.mark.parametrize("ni,nj", [(37, 21), (70, 33), (1024, 100)])
def test_layers_output_and_gradients_match_oracle(ctx, ni, nj):
    """Every convolution, pooling, upsampling and the head are checked against the oracle on the
    device's own input activations.  Differences then come only from FP32 vs FP64 accumulation
    order: an output whose FP32 and FP64 sums straddle a BF16 rounding boundary rounds one BF16
    step (2^-8 relative) apart.  So each differing value must be within one step, at most 1e-3 of
    the values may differ, and norm-wise the error stays below 3e-4 (a few such values in these
    small tensors give about 1e-4).  The backward pass is checked the same way from the device's forward
    activations (1e-2: a rounding difference of one BF16 gradient propagates through the
    remaining layers).  End to end the output stays within 2e-2 of the oracle's own forward
    pass: BF16 rounding flips amplify with depth, while a wrong index, tap or operand gives
    errors of order one.  At 1024 x 100 the weight-gradient workgroups take runs of three chunks
    that cross from one column band to the next, so the kernel's row ring is reused and refilled
    mid-run; on the small grids every run is a single chunk."""
    import cnn_oracle as N
    from starfish_amd.learned import LearnedGuess
    rng = np.random.default_rng(ni)
    net = LearnedGuess(ctx, ni=ni, nj=nj, max_batch=1)
    try:
        H, W = net.H, net.W
        assert (H, W) == ((nj + 3) // 4 * 4, (ni + 3) // 4 * 4)
        assert net.nparams == N.num_params()
        flat = N.init_params(seed=ni).astype(np.float32)
        net.set_params(flat)
        assert np.array_equal(net.params(), flat)
        f64 = flat.astype(np.float64)
        x4, t, m = _synthetic_sample(rng, H, W)
        net.add_sample(x4, t, m)
        sx4, st, sm = net.sample(0)
        assert np.array_equal(sx4, N.bf16(x4).astype(np.float32))
        loss, y, g = net.loss_and_grads(0)
        dev = {name: net.activation(k) for k, name in enumerate(N.DEVICE_TENSORS)}
        acts = N.acts_from_device(dev, H, W)
        # input assembly
        x0 = acts["c1a"][0]
        assert np.array_equal(x0[..., :4], sx4) and not x0[..., 4:].any()
        # each convolution from the device's input
        for name, _, _ in N.LAYERS:
            x, a = acts[name]
            ref = N.conv_layer(f64, name, x)
            e = _rel(a, ref)
            print("%-4s forward rel %.2e, BF16 values differing %.2e" % (name, e, np.mean(a != ref)), flush=True)
            step = 2.0 ** -7 * np.maximum(np.abs(a), np.abs(ref)) + 1e-6 * np.abs(ref).max()
            assert np.all(np.abs(a - ref) <= step), name
            assert np.mean(a != ref) < 1e-3, name
            assert e < 3e-4, (name, e)
        # pooling, upsampling and the concatenations
        e1, e2 = acts["c1b"][1], acts["c2b"][1]
        assert np.array_equal(acts["c2a"][0], N.bf16(N.pool(e1)))
        assert np.array_equal(acts["c3a"][0], N.bf16(N.pool(e2)))
        assert np.array_equal(acts["c4a"][0], np.concatenate([N.up(acts["c3b"][1]), e2], -1))
        assert np.array_equal(acts["c5a"][0], np.concatenate([N.up(acts["c4b"][1]), e1], -1))
        # head
        p = N.unpack(f64)
        yref = x0[..., 3] + acts["c5b"][1] @ p["head.w"] + p["head.b"][0]
        assert _rel(y, yref) < 1e-5
        assert np.array_equal(dev["y"].reshape(H, W), y)
        # loss and backward from the device's activations
        lo, dy = N.loss_grad_y(y.astype(np.float64), st.astype(np.float64), sm)
        assert abs(loss - lo) <= 1e-5 * lo
        go = N.backward(f64, acts, dy)
        off = 0
        for (name, k, nn) in N.LAYERS:
            for part, cnt in (("W", 9 * nn * k), ("b", nn)):
                e = _rel(g[off:off + cnt], go[off:off + cnt])
                print("%-4s %s grad rel err %.2e" % (name, part, e), flush=True)
                assert e < 1e-2, (name, part, e)
                off += cnt
        for part, cnt in (("head.w", N.C1), ("head.b", 1)):
            e = _rel(g[off:off + cnt], go[off:off + cnt])
            assert e < 1e-4, (part, e)
            off += cnt
        assert off == len(g)
        # end to end against the oracle's own forward pass
        yo, _ = N.forward(f64, sx4.astype(np.float64))
        print("end-to-end output rel err %.2e" % _rel(y, yo), flush=True)
        assert _rel(y, yo) < 2e-2
    finally:
        net.close()


@pytest.mark.gpu
def test_row_tile_walk_matches_oracle(ctx):
    """On a 1024 x 2048 grid a 32-channel convolution has 4096 tiles of 4 rows x 128 pixels, so each
    workgroup walks four row tiles (the next tile's operands loaded during the current one's
    MFMAs, the outputs stored and the accumulators cleared between tiles).  c5b's output is
    checked against the oracle on the device's own input, with the criteria of the layer test."""
    import cnn_oracle as N
    from starfish_amd.learned import LearnedGuess
    rng = np.random.default_rng(5)
    net = LearnedGuess(ctx, ni=1024, nj=2048, max_batch=1)
    try:
        H, W = net.H, net.W
        flat = N.init_params(seed=5).astype(np.float32)
        net.set_params(flat)
        x4, t, m = _synthetic_sample(rng, H, W)
        net.add_sample(x4, t, m)
        net.loss_and_grads(0)
        k_in, k_out = N.DEVICE_TENSORS.index("d1a"), N.DEVICE_TENSORS.index("d1")
        x = np.asarray(net.activation(k_in), np.float64).reshape(H, W, N.C1)
        a = np.asarray(net.activation(k_out), np.float64).reshape(H, W, N.C1)
        ref = N.conv_layer(flat.astype(np.float64), "c5b", x)
        e = _rel(a, ref)
        print("c5b at %dx%d: forward rel %.2e, BF16 values differing %.2e" % (W, H, e, np.mean(a != ref)), flush=True)
        step = 2.0 ** -7 * np.maximum(np.abs(a), np.abs(ref)) + 1e-6 * np.abs(ref).max()
        assert np.all(np.abs(a - ref) <= step)
        assert np.mean(a != ref) < 1e-3
        assert e < 3e-4
    finally:
        net.close()


@pytest.mark.gpu
def test_training_reduces_the_loss(ctx):
    """A smooth target that depends on the input through a local stencil: Adam on the device
    lowers the masked MSE of the initial network by at least 10x within 300 steps."""
    from starfish_amd.learned import LearnedGuess
    rng = np.random.default_rng(7)
    net = LearnedGuess(ctx, ni=64, nj=48, max_batch=4, seed=3)
    try:
        H, W = net.H, net.W
        yy, xx = np.mgrid[0:H, 0:W]
        for s in range(16):
            x4 = np.zeros((H, W, 4))
            cx, cy, r = rng.uniform(10, W - 10), rng.uniform(10, H - 10), rng.uniform(3, 8)
            inside = (xx - cx) ** 2 + (yy - cy) ** 2 < r * r
            v = rng.uniform(-1, 0)
            x4[..., 1] = inside
            x4[..., 2] = np.where(inside, v, 0)
            x4[..., 0] = np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (4 * r * r))
            t = v * np.exp(-np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * r))
            net.add_sample(x4, t, (~inside).astype(np.uint8))
        initial = np.mean([net.loss_and_grads(k)[0] for k in range(16)])
        last = net.train(300, batch=4, lr=1e-3)
        print("loss initial %.3e after 300 steps %.3e" % (initial, last), flush=True)
        assert np.isfinite(last) and last < initial / 10
    finally:
        net.close()
