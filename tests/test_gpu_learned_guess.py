"""What the learned initial guess gains, on the GPU, with the library's own training kernels.

The network is trained on snapshots of C2-style runs at a reduced 256x128-node resolution of the
same domain (cylinder radius U(0.02, 0.1) m, bias U(-200, 0) V).  On held-out decks (radius and
bias not in the training set) every potential solve is run to the same 1e-8 relative residual
from three starts: phi = 0 (plain), the previous step's phi (warm) and the network's guess
(learned, sf_poisson_solve_learned policy 0).  Asserted:

* every solve converges to the tolerance, whatever the start (plain PCG stays the judge);
* the learned start needs at least 3x fewer PCG iterations than phi = 0 on the held-out states;
* the learned solve ends on the same potential as the plain one: both stop at a 1e-8 relative
  nonlinear residual, which leaves up to ~2e-5 of the potential's scale between two solutions
  (1e-4 is asserted);
* a network with garbage weights still converges to the same potential, and policy 1 then keeps
  the previous potential as the start (it has the smaller residual).

The warm-start comparison is printed, not asserted: in a time loop the previous step's potential
is already a very good start."""
import ctypes as C

import numpy as np
import pytest

NI, NJ = 256, 128
PARTICLES = 400_000


def _held_out_state(radius, bias, seed, steps=12):
    from starfish_amd import deck as D
    from starfish_amd.plugin import Simulation
    deck = D.config_c2(particles=PARTICLES, radius=radius, bias=bias, ni=NI, nj=NJ)
    sim = Simulation(deck, capacity=int(PARTICLES * 1.6) + 100_000)
    n, drift, temp = deck.preload
    sim.materials[0].preload(n, drift, temp, seed=seed)
    sim.initial_solve()
    for _ in range(steps):
        sim.step()
    return sim


def _solve(sim, start, net, policy=0):
    from starfish_amd import _lib
    from starfish_amd.learned import solve_from
    if start != "learned" or policy == 0:
        return solve_from(sim, start, net)
    prev = sim.phi()
    info = _lib.SolveInfo()
    used = C.c_int()
    sim.ctx.call("sf_rho_update")
    rc = _lib.check(sim.ctx.h, _lib.load().sf_poisson_solve_learned(sim.ctx.h, net.ptr, policy, C.byref(info),
                                                                    C.byref(used)))
    sim.ctx.call("sf_phi_set", prev.ctypes.data_as(C.POINTER(C.c_double)))
    return info, rc, used.value


def _solved_phi(sim, start, net):
    """phi after a solve from `start` (the simulation keeps its previous potential afterwards)."""
    from starfish_amd import _lib
    prev = sim.phi()
    info = _lib.SolveInfo()
    sim.ctx.call("sf_rho_update")
    if start == "zero":
        z = np.zeros_like(prev)
        sim.ctx.call("sf_phi_set", z.ctypes.data_as(C.POINTER(C.c_double)))
        sim.ctx.call("sf_poisson_solve", None, C.byref(info))
    else:
        used = C.c_int()
        _lib.check(sim.ctx.h, _lib.load().sf_poisson_solve_learned(sim.ctx.h, net.ptr, 0, C.byref(info),
                                                                   C.byref(used)))
    phi = sim.phi()
    sim.ctx.call("sf_phi_set", prev.ctypes.data_as(C.POINTER(C.c_double)))
    return phi


@pytest.mark.gpu
def test_learned_guess_gains_over_plain_pcg_on_held_out_decks():
    from starfish_amd.learned import LearnedGuess, generate_c2_snapshots
    from starfish_amd.plugin import Context
    ctx = Context(0)
    net = LearnedGuess(ctx, ni=NI, nj=NJ, max_batch=4, seed=3)
    garbage = LearnedGuess(ctx, ni=NI, nj=NJ, max_batch=4, seed=4)
    sims = []
    try:
        used = generate_c2_snapshots(net, 8, 16, 2, PARTICLES, NI, NJ, seed=2024, skip=2)
        print("training samples %d from %s" % (net.num_samples, ["%.3f/%.0f" % u for u in used]), flush=True)
        loss = 0.0
        for r in range(4):
            loss = net.train(150, batch=4, lr=1e-3 if r < 3 else 3e-4)
        print("final loss %.3e" % loss, flush=True)
        # garbage: random weights blown up far beyond a trained network's scale
        rng = np.random.default_rng(9)
        garbage.set_params(rng.normal(size=garbage.nparams).astype(np.float32) * 0.5)
        ratios = []
        for radius, bias, seed in ((0.035, -150.0, 77), (0.085, -40.0, 78)):
            sim = _held_out_state(radius, bias, seed)
            sims.append(sim)
            info_p, rc_p = _solve(sim, "zero", net)
            info_w, rc_w = _solve(sim, "current", net)
            info_l, rc_l = _solve(sim, "learned", net)
            print("held-out r=%.3f bias=%.0f: PCG iterations plain %d, warm %d, learned %d; Newton %d/%d/%d; "
                  "rel res %.1e/%.1e/%.1e" % (radius, bias, info_p.pcg_iterations, info_w.pcg_iterations,
                                              info_l.pcg_iterations, info_p.newton_iterations,
                                              info_w.newton_iterations, info_l.newton_iterations,
                                              info_p.rel_residual, info_w.rel_residual, info_l.rel_residual),
                  flush=True)
            for info, rc in ((info_p, rc_p), (info_w, rc_w), (info_l, rc_l)):
                assert rc == 0 and info.rel_residual <= 1e-8
            ratios.append(info_p.pcg_iterations / max(1, info_l.pcg_iterations))
            phi_plain = _solved_phi(sim, "zero", net)
            phi_learn = _solved_phi(sim, "learned", net)
            scale = np.abs(phi_plain).max()
            assert np.abs(phi_learn - phi_plain).max() <= 1e-4 * scale
            # garbage weights: still converges to the same potential
            info_g, rc_g = _solve(sim, "learned", garbage)
            print("  garbage network: PCG %d, Newton %d, rel res %.1e" % (info_g.pcg_iterations,
                                                                        info_g.newton_iterations,
                                                                        info_g.rel_residual), flush=True)
            assert rc_g == 0 and info_g.rel_residual <= 1e-8
            phi_g = _solved_phi(sim, "learned", garbage)
            assert np.abs(phi_g - phi_plain).max() <= 1e-4 * scale
            # policy 1 compares residuals and keeps the (warm) previous potential over garbage
            info_1, rc_1, used_guess = _solve(sim, "learned", garbage, policy=1)
            assert rc_1 == 0 and info_1.rel_residual <= 1e-8 and used_guess == 0
        print("plain / learned PCG iteration ratios: %s" % ["%.1f" % r for r in ratios], flush=True)
        assert min(ratios) >= 3.0
    finally:
        for s in sims:
            s.close()
        garbage.close()
        net.close()
        ctx.close()
