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


ACT_NAMES = ["x0", "e1a", "u1", "p1", "e2a", "u2", "p2", "e3a", "e3", "d2a", "d2", "d1a", "d1", "y"]


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
@pytest.mark.parametrize("ni,nj", [(37, 21), (70, 33)])
def test_loss_output_and_gradients_match_oracle(ctx, ni, nj):
    """Norm-wise relative error per tensor: output 2e-3, gradients 2e-2 (FP32 vs FP64 accumulation
    flips an occasional BF16 rounding of an activation; a wrong index or sign gives O(1))."""
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
        x4, t, m = _synthetic_sample(rng, H, W)
        net.add_sample(x4, t, m)
        sx4, st, sm = net.sample(0)
        assert np.array_equal(sx4, N.bf16(x4).astype(np.float32))
        loss, y, g = net.loss_and_grads(0)
        acts = {}
        lo, yo, go = N.loss_and_grads(flat.astype(np.float64), sx4.astype(np.float64), st.astype(np.float64), sm,
                                      acts_out=acts)
        for k, name in enumerate(ACT_NAMES):
            a = net.activation(k)
            r = acts[name].reshape(-1)
            d = np.abs(a - r).reshape(acts[name].shape)
            rows = d.reshape(d.shape[0], -1).max(1)
            print("%-4s rel %.2e  worst row %d of %d (%.2e)" % (name, _rel(a, r), rows.argmax(), len(rows),
                                                              rows.max()), flush=True)
        print("loss", loss, lo, "y err", _rel(y, yo), flush=True)
        assert _rel(y, yo) < 2e-3
        assert abs(loss - lo) <= 5e-3 * lo
        sizes = [("%s.W" % n, 9 * nn * k) for n, k, nn in N.LAYERS]
        off = 0
        for (name, k, nn) in N.LAYERS:
            for part, cnt in (("W", 9 * nn * k), ("b", nn)):
                e = _rel(g[off:off + cnt], go[off:off + cnt])
                print(name, part, "grad rel err %.2e" % e, flush=True)
                assert e < 2e-2, (name, part, e)
                off += cnt
        for part, cnt in (("head.w", N.C1), ("head.b", 1)):
            e = _rel(g[off:off + cnt], go[off:off + cnt])
            assert e < 2e-2, (part, e)
            off += cnt
        assert off == len(g) and sizes
    finally:
        net.close()


@pytest.mark.gpu
def test_training_reduces_the_loss(ctx):
    """A smooth target that depends on the input through a local stencil: Adam on the device
    lowers the masked MSE by at least 10x within 300 steps."""
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
        first = net.train(10, batch=4, lr=1e-3)
        last = net.train(300, batch=4, lr=1e-3)
        print("loss first %.3e last %.3e" % (first, last), flush=True)
        assert np.isfinite(last) and last < first / 10
    finally:
        net.close()
