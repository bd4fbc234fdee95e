"""CPU oracle of the learned-guess network (test infrastructure only: tests and smoke checks use
it to check the device kernels; no product path imports it).

The network has no counterpart in the reference.  This module restates, in float64 NumPy, the
forward pass, the masked-MSE loss and the backward pass of the device implementation, with the
same BF16 rounding points:
  * convolution weights are used as BF16 copies of the FP32 master weights;
  * every stored activation (convolution outputs after ReLU, pooled and upsampled tensors) is
    rounded to BF16;
  * the gradient that enters a convolution's backward (after the ReLU mask) is rounded to BF16;
  * everything else (accumulations, loss, head, pooling/upsampling gradients) is exact here
    (FP32 on the device).
With `bf16=False` no rounding is applied: that mode is checked against central finite
differences (tests/test_cnn_oracle.py), which pins the backward formulas of this file.

Architecture (channels, NHWC, 3x3 convolutions with zero padding 1, ReLU after every conv):
  x0 (4 input channels, padded to 32) -> c1a 32 -> c1b 32 = e1
  avgpool2 -> c2a 64 -> c2b 64 = e2
  avgpool2 -> c3a 128 -> c3b 128 = e3
  up2(e3) ++ e2 (192) -> c4a 64 -> c4b 64 = d2
  up2(d2) ++ e1 (96)  -> c5a 32 -> c5b 32 = d1
  y = x0[..., 3] + d1 . w_head + b_head
Parameter order (matches the device): for each conv W[9][N][K] then b[N]; then w_head[32],
b_head[1].  Tap t = 3*(dy+1) + (dx+1) reads input pixel (y+dy, x+dx).
"""
import numpy as np

C1, C2, C3 = 32, 64, 128
LAYERS = [("c1a", 32, C1), ("c1b", C1, C1), ("c2a", C1, C2), ("c2b", C2, C2), ("c3a", C2, C3),
          ("c3b", C3, C3), ("c4a", C3 + C2, C2), ("c4b", C2, C2), ("c5a", C2 + C1, C1), ("c5b", C1, C1)]


def bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(finite, r, u & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def num_params():
    return sum(9 * n * k + n for _, k, n in LAYERS) + C1 + 1


def unpack(flat):
    """Flat parameter vector -> dict of arrays (W as [9][N][K])."""
    p, off = {}, 0
    for name, k, n in LAYERS:
        p[name + ".W"] = flat[off:off + 9 * n * k].reshape(9, n, k).astype(np.float64)
        off += 9 * n * k
        p[name + ".b"] = flat[off:off + n].astype(np.float64)
        off += n
    p["head.w"] = flat[off:off + C1].astype(np.float64)
    off += C1
    p["head.b"] = flat[off:off + 1].astype(np.float64)
    off += 1
    assert off == len(flat)
    return p


def pack(g):
    out = []
    for name, _, _ in LAYERS:
        out += [g[name + ".W"].ravel(), g[name + ".b"].ravel()]
    out += [g["head.w"].ravel(), g["head.b"].ravel()]
    return np.concatenate(out)


def _shift(a, dy, dx):
    """s[y, x] = a[y + dy, x + dx] (zero outside)."""
    H, W = a.shape[:2]
    s = np.zeros_like(a)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    if ys < ye and xs < xe:
        s[ys:ye, xs:xe] = a[ys + dy:ye + dy, xs + dx:xe + dx]
    return s


TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def conv_fwd(x, Wt, b):
    """out[p][n] = sum_t sum_k x[p + d_t][k] W[t][n][k] + b[n] (no activation)."""
    out = np.zeros(x.shape[:2] + (Wt.shape[1],))
    for t, (dy, dx) in enumerate(TAPS):
        out += _shift(x, dy, dx) @ Wt[t].T
    return out + b


def conv_bwd(x, Wt, g):
    """Gradients of conv_fwd: (dW, db, dx) for upstream gradient g[p][n]."""
    dW = np.zeros_like(Wt)
    dx = np.zeros(x.shape)
    gf = g.reshape(-1, g.shape[-1])
    for t, (dy, ddx) in enumerate(TAPS):
        dW[t] = gf.T @ _shift(x, dy, ddx).reshape(-1, x.shape[-1])
        dx += _shift(g, -dy, -ddx) @ Wt[t]
    return dW, g.sum((0, 1)), dx


def pool(a):
    return 0.25 * (a[0::2, 0::2] + a[0::2, 1::2] + a[1::2, 0::2] + a[1::2, 1::2])


def pool_bwd(g):
    return 0.25 * np.repeat(np.repeat(g, 2, 0), 2, 1)


def up(a):
    return np.repeat(np.repeat(a, 2, 0), 2, 1)


def up_bwd(g):
    return g[0::2, 0::2] + g[0::2, 1::2] + g[1::2, 0::2] + g[1::2, 1::2]


def forward(flat, x4, use_bf16=True):
    """Forward pass.  Returns (y, acts) with acts[name] = (input, output) of every convolution
    (outputs after ReLU and BF16 rounding) plus acts["d1"] = the head's input."""
    R = bf16 if use_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    p = unpack(np.asarray(flat))
    Wq = {name: R(p[name + ".W"]) for name, _, _ in LAYERS}
    H, W = x4.shape[:2]
    x0 = np.zeros((H, W, 32))
    x0[..., :4] = R(x4)
    acts = {}

    def conv(name, x):
        a = R(np.maximum(conv_fwd(x, Wq[name], p[name + ".b"]), 0.0))
        acts[name] = (x, a)
        return a

    e1 = conv("c1b", conv("c1a", x0))
    e2 = conv("c2b", conv("c2a", R(pool(e1))))
    e3 = conv("c3b", conv("c3a", R(pool(e2))))
    d2 = conv("c4b", conv("c4a", np.concatenate([up(e3), e2], -1)))
    d1 = conv("c5b", conv("c5a", np.concatenate([up(d2), e1], -1)))
    y = x0[..., 3] + d1 @ p["head.w"] + p["head.b"][0]
    return y, acts


def backward(flat, acts, dy, use_bf16=True):
    """Parameter gradients (packed) for upstream dL/dy, given the forward activations `acts`
    (as returned by forward or acts_from_device)."""
    R = bf16 if use_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    p = unpack(np.asarray(flat))
    Wq = {name: R(p[name + ".W"]) for name, _, _ in LAYERS}
    d1 = acts["c5b"][1]
    g = {"head.w": (dy[..., None] * d1).sum((0, 1)), "head.b": np.array([dy.sum()])}

    def back(name, dout):
        x, a = acts[name]
        gz = R(np.where(a > 0, dout, 0.0))
        dW, db, dx = conv_bwd(x, Wq[name], gz)
        g[name + ".W"], g[name + ".b"] = dW, db
        return dx

    g_d1 = dy[..., None] * p["head.w"]
    g_u1 = back("c5a", back("c5b", g_d1))
    g_e1 = g_u1[..., C2:].copy()
    g_d2 = up_bwd(g_u1[..., :C2])
    g_u2 = back("c4a", back("c4b", g_d2))
    g_e2 = g_u2[..., C3:].copy()
    g_e3 = up_bwd(g_u2[..., :C3])
    g_p2 = back("c3a", back("c3b", g_e3))
    g_e2 = g_e2 + pool_bwd(g_p2)
    g_p1 = back("c2a", back("c2b", g_e2))
    g_e1 = g_e1 + pool_bwd(g_p1)
    back("c1a", back("c1b", g_e1))
    return pack(g)


def loss_grad_y(y, target, mask):
    """Masked MSE over all H*W pixels and its gradient dL/dy."""
    n = y.size
    e = np.where(mask != 0, y - target, 0.0)
    return float((e * e).sum() / n), 2.0 * e / n


def loss_and_grads(flat, x4, target, mask, use_bf16=True, acts_out=None):
    """Loss = sum over masked pixels of (y - t)^2 / (H W); returns (loss, y, flat gradient).

    x4 [H][W][4] input channels, target [H][W], mask [H][W] (nonzero = counted).  If acts_out
    is a dict it receives the forward activations (see forward)."""
    y, acts = forward(flat, x4, use_bf16)
    if acts_out is not None:
        acts_out.update(acts)
    loss, dy = loss_grad_y(y, target, mask)
    return loss, y, backward(flat, acts, dy, use_bf16)


# device tensors (sf_cnn_activation order) -> acts of forward()
DEVICE_TENSORS = ["x0", "e1a", "u1", "p1", "e2a", "u2", "p2", "e3a", "e3", "d2a", "d2", "d1a", "d1", "y"]


def acts_from_device(t, H, W):
    """t: dict name -> flat FP32 array read back from the device (names of DEVICE_TENSORS).
    Returns acts[name] = (input, output) of every convolution, as forward() does."""
    sh = {"x0": (H, W, 32), "e1a": (H, W, C1), "u1": (H, W, C2 + C1), "p1": (H // 2, W // 2, C1),
          "e2a": (H // 2, W // 2, C2), "u2": (H // 2, W // 2, C3 + C2), "p2": (H // 4, W // 4, C2),
          "e3a": (H // 4, W // 4, C3), "e3": (H // 4, W // 4, C3), "d2a": (H // 2, W // 2, C2),
          "d2": (H // 2, W // 2, C2), "d1a": (H, W, C1), "d1": (H, W, C1)}
    a = {k: np.asarray(t[k], np.float64).reshape(v) for k, v in sh.items()}
    e1, e2 = a["u1"][..., C2:], a["u2"][..., C3:]
    return {"c1a": (a["x0"], a["e1a"]), "c1b": (a["e1a"], e1), "c2a": (a["p1"], a["e2a"]), "c2b": (a["e2a"], e2),
            "c3a": (a["p2"], a["e3a"]), "c3b": (a["e3a"], a["e3"]), "c4a": (a["u2"], a["d2a"]),
            "c4b": (a["d2a"], a["d2"]), "c5a": (a["u1"], a["d1a"]), "c5b": (a["d1a"], a["d1"])}


def conv_layer(flat, name, x, use_bf16=True):
    """One convolution + ReLU (+ BF16 rounding) of layer `name` on input x."""
    R = bf16 if use_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    p = unpack(np.asarray(flat))
    return R(np.maximum(conv_fwd(x, R(p[name + ".W"]), p[name + ".b"]), 0.0))


def init_params(seed=0):
    """He-initialised parameters over the real fan-in (4 input channels for c1a)."""
    rng = np.random.default_rng(seed)
    g = {}
    for name, k, n in LAYERS:
        kr = 4 if name == "c1a" else k
        W = np.zeros((9, n, k))
        W[:, :, :kr] = rng.normal(0, np.sqrt(2.0 / (9 * kr)), (9, n, kr))
        g[name + ".W"] = W
        g[name + ".b"] = rng.normal(0, 0.05, n)
    g["head.w"] = rng.normal(0, np.sqrt(1.0 / C1), C1)
    g["head.b"] = np.array([0.01])
    return pack(g)
