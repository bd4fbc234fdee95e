"""Learned initial guess for the potential solve (host side of sf_cnn_* in the C-ABI).

The network maps the solver input of a context (rho, Dirichlet mask, boundary phi and the
Boltzmann inversion phi0 + kTe ln(rho / (e n0))) to phi / 100 V.  Training data are snapshots
of solved contexts; training runs on the device.  `predict(sim)` writes the guess into the
context's potential, which `sf_poisson_solve` then takes as its starting point; the solve
itself is unchanged and still runs to its residual tolerance.
"""
import ctypes as C

import numpy as np

from . import _lib


class LearnedGuess:
    def __init__(self, ctx, ni=0, nj=0, max_batch=4, seed=1):
        """ctx: a plugin.Context on a device; ni, nj: node grid (0: the context's mesh)."""
        self.ctx = ctx
        self.lib = _lib.load()
        h = C.c_void_p()
        _lib.check(ctx.h, self.lib.sf_cnn_create(ctx.h, ni, nj, max_batch, seed, C.byref(h)))
        self.ptr = h
        H, W, n = C.c_int(), C.c_int(), C.c_int64()
        self.lib.sf_cnn_dims(self.ptr, C.byref(H), C.byref(W), C.byref(n))
        self.H, self.W, self.nparams = H.value, W.value, n.value

    def close(self):
        if self.ptr:
            self.lib.sf_cnn_destroy(self.ptr)
            self.ptr = None

    def _check(self, rc):
        return _lib.check(self.ctx.h, rc)

    # -- data -------------------------------------------------------------------------------
    def snapshot(self, sim):
        """Adds the current rho / boundary / phi of a solved Simulation to the training set."""
        self._check(self.lib.sf_cnn_snapshot(self.ptr, sim.ctx.h))

    def add_sample(self, x4, target, mask):
        x4 = np.ascontiguousarray(x4, dtype=np.float32).reshape(self.H, self.W, 4)
        t = np.ascontiguousarray(target, dtype=np.float32).reshape(self.H, self.W)
        m = np.ascontiguousarray(mask, dtype=np.uint8).reshape(self.H, self.W)
        self._check(self.lib.sf_cnn_add_sample(self.ptr, _fp(x4), _fp(t), m.ctypes.data_as(C.POINTER(C.c_uint8))))

    def sample(self, idx):
        x4 = np.zeros((self.H, self.W, 4), np.float32)
        t = np.zeros((self.H, self.W), np.float32)
        m = np.zeros((self.H, self.W), np.uint8)
        self._check(self.lib.sf_cnn_get_sample(self.ptr, idx, _fp(x4), _fp(t), m.ctypes.data_as(C.POINTER(C.c_uint8))))
        return x4, t, m

    @property
    def num_samples(self):
        return self.lib.sf_cnn_num_samples(self.ptr)

    # -- parameters ---------------------------------------------------------------------------
    def params(self):
        buf = np.zeros(self.nparams, np.float32)
        self._check(self.lib.sf_cnn_params(self.ptr, 1, _fp(buf), None))
        return buf

    def set_params(self, flat):
        buf = np.ascontiguousarray(flat, dtype=np.float32)
        if buf.size != self.nparams:
            raise ValueError("expected %d parameters, got %d" % (self.nparams, buf.size))
        self._check(self.lib.sf_cnn_params(self.ptr, 0, _fp(buf), None))

    def save(self, path):
        np.save(path, self.params())

    def load(self, path):
        self.set_params(np.load(path, allow_pickle=False))

    # -- training / inference -----------------------------------------------------------------
    def train(self, steps, batch=4, lr=1e-3):
        loss = C.c_double()
        self._check(self.lib.sf_cnn_train(self.ptr, steps, batch, lr, C.byref(loss)))
        return loss.value

    def loss_and_grads(self, idx):
        loss = C.c_double()
        y = np.zeros((self.H, self.W), np.float32)
        g = np.zeros(self.nparams, np.float32)
        self._check(self.lib.sf_cnn_loss_and_grads(self.ptr, idx, C.byref(loss), _fp(y), _fp(g), None))
        return loss.value, y, g

    def activation(self, which):
        """Activation `which` (0..13, see sf_cnn_activation) of batch slot 0 after the last forward."""
        n = C.c_int64()
        self._check(self.lib.sf_cnn_activation(self.ptr, which, None, C.byref(n)))
        out = np.zeros(n.value, np.float32)
        self._check(self.lib.sf_cnn_activation(self.ptr, which, _fp(out), None))
        return out

    def predict(self, sim):
        """Writes the network's potential into sim's phi (the next solve's starting point)."""
        self._check(self.lib.sf_cnn_predict(self.ptr, sim.ctx.h))


def mfma_probe(ctx, a, b):
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(16, 32)
    b = np.ascontiguousarray(b, dtype=np.float32).reshape(32, 16)
    d = np.zeros((16, 16), np.float32)
    _lib.check(ctx.h, _lib.load().sf_cnn_mfma_probe(ctx.h, _fp(a), _fp(b), _fp(d)))
    return d


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def solve_from(sim, start, net=None):
    """Solves sim's potential for its current rho from a chosen start and restores the previous
    potential afterwards.  start: "zero" (phi = 0), "current" (the potential sim holds) or
    "learned" (net's guess).  Returns (SolveInfo, status)."""
    lib = _lib.load()
    prev = sim.phi()
    if start == "zero":
        zero = np.zeros_like(prev)
        sim.ctx.call("sf_phi_set", zero.ctypes.data_as(C.POINTER(C.c_double)))
    info = _lib.SolveInfo()
    sim.ctx.call("sf_rho_update")
    if start == "learned":
        used = C.c_int()
        rc = _lib.check(sim.ctx.h, lib.sf_poisson_solve_learned(sim.ctx.h, net.ptr, 0, C.byref(info), C.byref(used)))
    else:
        rc = sim.ctx.call("sf_poisson_solve", None, C.byref(info))
    sim.ctx.call("sf_phi_set", prev.ctypes.data_as(C.POINTER(C.c_double)))
    return info, rc


def generate_c2_snapshots(net, n_configs, steps, every, particles, ni, nj, seed=0, skip=0, log=None,
                          initial=False):
    """Runs C2-style simulations (O+ beam past a cylinder of radius U(0.02, 0.1) m biased at
    U(-200, 0) V) and adds a snapshot of (rho, boundary, phi) every `every` steps after `skip`
    steps, and with `initial` also the solved state of the preloaded plasma before the first step
    (the cold solve's case).  Returns the list of (radius, bias) used."""
    from . import deck as D
    from .plugin import Simulation
    rng = np.random.default_rng(seed)
    used = []
    for k in range(n_configs):
        radius, bias = float(rng.uniform(0.02, 0.1)), float(rng.uniform(-200.0, 0.0))
        deck = D.config_c2(particles=particles, radius=radius, bias=bias, ni=ni, nj=nj)
        sim = Simulation(deck, capacity=int(particles * 1.6) + 100_000)
        try:
            n, drift, temp = deck.preload
            sim.materials[0].preload(n, drift, temp, seed=int(rng.integers(1 << 30)))
            sim.initial_solve()
            if initial:
                net.snapshot(sim)
            for s in range(steps):
                sim.step()
                if s >= skip and (s - skip) % every == 0:
                    net.snapshot(sim)
            if log:
                log("config %d radius %.3f bias %.1f: %d samples" % (k, radius, bias, net.num_samples))
        finally:
            sim.close()
        used.append((radius, bias))
    return used
