#!/usr/bin/env python
"""Benchmark of the Starfish ES-PIC hot path on one MI355X.

Workload (default): O+ flow past a -100 V cylinder on a 1024x512-node XY mesh (dx = 1e-3 m),
2e7 macroparticles preloaded as a drifting Maxwellian (7000 m/s, 1000 K) plus left-edge
injection at 1e20 m^-2 s^-1, Boltzmann electrons (n0 = 1e10 m^-3, kTe = 2 eV), dt = 1e-7 s.
A step is one Starfish main-loop iteration: source injection, push + deposit + counting sort,
rho, Newton-PCG potential solve to a 1e-8 relative residual, E = -grad(phi).

Output: one JSON line.  `value` is particles pushed and deposited per second of device time of
the particle phase (push kernels, cell scan, scatter into cell order, normalisation), summed
over the timed steps with HIP events on the library's stream.  The full-step rate (all phases,
host clock), the potential-solve cost, the CPU baselines of the reference algorithm and the
errors against it on a sample of the same workload are reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
BYTES_PUSH = 48.0          # read + write of the 24 B particle state per particle pushed
BYTES_SCATTER = 48.0       # sort step: read the 24 B pushed state (its cell is the key), write it sorted
BYTES_STATE = 48.0         # read + write of the 24 B particle state (minimal push + deposit)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pmc_traffic_push_tile_fast_xy_c2.txt")


def measured_bytes_per_particle(kernel):
    """HBM bytes per particle of one push kernel from the committed FETCH_SIZE + WRITE_SIZE summary
    (its lines: kernel, particles per dispatch, bytes per dispatch, already scaled by the
    copy-kernel calibration); None when the summary or the kernel is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 3 and parts[0] == kernel:
                    return float(parts[2]) / float(parts[1])
    except (OSError, ValueError):
        pass
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--particles", type=float, default=2e7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads of the CPU baseline (0: all host cores)")
    ap.add_argument("--sort-interval", type=int, default=4)
    ap.add_argument("--learned-configs", type=int, default=8,
                    help="C2-style runs (random cylinder radius / bias) that provide training snapshots")
    ap.add_argument("--learned-train-steps", type=int, default=600)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--workload", choices=["c2", "c3", "mcc", "dsmc"], default="c2",
                    help="c2: the benchmark (default); c3: the RZ push at 2048x1024 with 1e8 Xe+; mcc: c3 with "
                         "charge exchange every step; dsmc: 5e7 N2 with NTC collisions every step (these three print "
                         "their own JSON line and use --steps / --warmup)")
    ap.add_argument("--cnn-only", action="store_true",
                    help="time only the learned guess's network (training steps at 1024x512, batch 4, on "
                         "synthetic samples) and print that JSON line")
    ap.add_argument("--host-sources", action="store_true",
                    help="sample the boundary injection on the host with the reference's random stream "
                         "(default: sampled on the device)")
    return ap.parse_args()


# (input channels, output channels, grid divisor) of the network's 3x3 convolutions, input channels
# padded to multiples of 32 as the kernels run them (sf_cnn.hip: conv_init)
CNN_LAYERS = [(32, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2), (64, 128, 4), (128, 128, 4),
              (192, 64, 2), (64, 64, 2), (96, 32, 1), (32, 32, 1)]


def cnn_flops(H, W, batch):
    """Matrix-core flops of one forward pass and of one training step (forward, backward-data for
    every layer but the first, backward-weights) at grid H x W and the given batch."""
    fwd = sum(2.0 * 9 * k * n * (H // d) * (W // d) * batch for k, n, d in CNN_LAYERS)
    first = 2.0 * 9 * 32 * 32 * H * W * batch
    return fwd, 3.0 * fwd - first


def cnn_throughput(ctx, ni, nj, batch=4, steps=20):
    """Training steps of the learned guess's network on synthetic samples at the benchmark's grid:
    device time per step (HIP events on the library's stream) and matrix-core TFLOP/s."""
    import ctypes as C
    import numpy as np
    from starfish_amd.learned import LearnedGuess
    net = LearnedGuess(ctx, ni=ni, nj=nj, max_batch=batch, seed=7)
    try:
        rng = np.random.default_rng(5)
        for _ in range(batch):
            net.add_sample(rng.normal(size=(net.H, net.W, 4)).astype(np.float32),
                           rng.normal(size=(net.H, net.W)).astype(np.float32), np.ones((net.H, net.W), np.uint8))
        net.train(2, batch=batch, lr=1e-4)
        ctx.call("sf_ctx_synchronize")
        t0 = time.perf_counter()
        net.train(steps, batch=batch, lr=1e-4)
        ctx.call("sf_ctx_synchronize")
        ms = (time.perf_counter() - t0) / steps * 1e3
        fwd, train = cnn_flops(net.H, net.W, batch)
        return {"grid": [net.H, net.W], "batch": batch, "steps": steps, "train_step_ms": ms,
                "train_step_tflops": train / (ms * 1e-3) / 1e12, "forward_gflop": fwd / 1e9,
                "train_step_gflop": train / 1e9, "peak_bf16_tflops": 2500.0}
    finally:
        net.close()


def other_workload(args):
    """The other configurations' hot kernels on one GPU: c3 (RZ 2048x1024 nodes, 1e8 Xe+ preloaded,
    push + deposit + sort every 4 steps + Boltzmann-electron solve), mcc (c3 plus charge exchange on
    a 1e19 m^-3 Xe background every step) and dsmc (512x256 cells, 5e7 N2, free flight + sort +
    NTC collisions every step).  Device time of the particle phase from HIP events, the MCC / DSMC
    pass from the host clock around a synchronised call."""
    import ctypes as C
    import numpy as np
    from starfish_amd import deck as D
    from starfish_amd.plugin import Simulation
    out = {"workload": args.workload, "steps": args.steps, "warmup": args.warmup, "n_gpus": 1}
    if args.workload in ("c3", "mcc"):
        deck = D.config_c3(particles=100_000_000)
        sim = Simulation(deck, capacity=110_000_000, track_ids=args.workload == "mcc")
    else:
        mesh = D.MeshSpec("XY", 513, 257, (0.0, 0.0), (1e-4, 1e-4), {})
        n = 50_000_000
        spwt = 1e22 * 512 * 256 * 1e-8 / n
        deck = D.Deck(mesh, materials=[D.MaterialSpec("N2", "kinetic_hip", 28.0, 0.0, spwt, 0.0)], solver=None,
                      num_it=1, dt=1e-8, seed=0)
        sim = Simulation(deck, capacity=n + 1_000_000, track_ids=True)
    try:
        mat = sim.materials[0]
        coll = None
        if args.workload in ("c3", "mcc"):
            npart, drift, temp = deck.preload
            mat.preload(npart, drift, temp, seed=3)
            sim.initial_solve()
            if args.workload == "mcc":
                from starfish_amd.interactions import MccHip
                coll = MccHip(sim, mat, "CEX", 5e-19, 131.0, seed=17)
                coll.set_target(np.full(sim.nn, 1e19), temp=np.full(sim.nn, 300.0))
        else:
            from starfish_amd.interactions import DsmcHip
            mat.preload(50_000_000, (500.0, 0.0, 0.0), 300.0, seed=5)
            coll = DsmcHip(sim, mat, mat, sig_cr_max=1e-16, seed=3, sigma="bird463", diam=(4.17e-10, 4.17e-10),
                           ref_temp=(273.0, 273.0), visc_temp_index=(0.74, 0.74))
        sim.ctx.call("sf_set_timing", 1)
        pushed, kern, coll_ms, ncoll, solve_ms = 0, np.zeros(4), 0.0, 0, []
        for step in range(args.warmup + args.steps):
            timed = step >= args.warmup
            np_step = mat.getNp()
            mat.update()
            tm = np.zeros(4)
            sim.ctx.call("sf_last_timings", tm.ctypes.data_as(C.POINTER(C.c_double)))
            if coll is not None:
                sim.ctx.call("sf_ctx_synchronize")
                t0 = time.perf_counter()
                nc = coll.perform(sim.it + 1, 0.0) if args.workload == "mcc" else coll.perform(sim.it + 1)
                sim.ctx.call("sf_ctx_synchronize")
                if timed:
                    coll_ms += (time.perf_counter() - t0) * 1e3
                    ncoll += nc
            if deck.solver is not None:
                sim.update_fields()
                if timed:
                    solve_ms.append(sim.solve_infos[-1].ms)
            sim.it += 1
            if timed:
                pushed += np_step
                kern += tm
        phase_s = kern.sum() * 1e-3
        out.update({"particles_per_step": pushed / args.steps, "particles_per_s": pushed / phase_s,
                    "particle_phase_ms": {"push": kern[0] / args.steps, "scan": kern[1] / args.steps,
                                          "scatter": kern[2] / args.steps, "normalize": kern[3] / args.steps},
                    "push_GBs": 48.0 * pushed / (kern[0] * 1e-3) / 1e9})
        if solve_ms:
            out["solve_ms"] = float(np.mean(solve_ms))
        if coll is not None:
            out.update({"collision_pass_ms": coll_ms / args.steps, "collisions_per_step": ncoll / args.steps,
                        "particles_per_s_collision_pass": pushed / (coll_ms * 1e-3)})
        return out
    finally:
        if coll is not None:
            coll.close()
        sim.close()


def learned_guess(sim, deck, args, rho0):
    """Trains the learned initial guess on snapshots of C2-style runs (cylinder radius
    U(0.02, 0.1) m, bias U(-200, 0) V; none is the benchmark's cylinder).  Reported: (1) the cold
    initial solve of the benchmark (preloaded plasma, phi = 0) against the same solve started from
    the network's guess, and (2) on the benchmark's own time-loop states, the solve from phi = 0
    (plain), from the previous step's phi (warm) and from the guess (learned).  Every solve runs to
    the same residual tolerance."""
    import ctypes as C
    import numpy as np
    from starfish_amd import _lib
    from starfish_amd.learned import LearnedGuess, generate_c2_snapshots, solve_from
    t0 = time.perf_counter()
    m = deck.mesh
    net = LearnedGuess(sim.ctx, max_batch=4, seed=1)
    try:
        generate_c2_snapshots(net, args.learned_configs, 24, 3, int(args.particles), m.ni, m.nj, seed=12345,
                              skip=3, initial=True)
        t_gen = time.perf_counter() - t0
        t1 = time.perf_counter()
        loss = 0.0
        for r in range(4):
            loss = net.train(args.learned_train_steps // 4, batch=4, lr=1e-3 if r < 3 else 3e-4)
        sim.ctx.call("sf_ctx_synchronize")
        t_train = time.perf_counter() - t1
        # (1) the cold initial solve, from phi = 0 and from the guess, on the initial rho
        prev_phi, prev_rho = sim.phi(), sim.rho()
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        cold = {}
        for start in ("zero", "learned"):
            sim.ctx.call("sf_rho_set", dp(rho0))
            sim.ctx.call("sf_phi_set", dp(np.zeros(sim.nn)))
            info = _lib.SolveInfo()
            if start == "zero":
                rc = sim.ctx.call("sf_poisson_solve", None, C.byref(info))
            else:
                used = C.c_int()
                rc = _lib.check(sim.ctx.h, _lib.load().sf_poisson_solve_learned(sim.ctx.h, net.ptr, 0, C.byref(info),
                                                                               C.byref(used)))
            cold[start] = {"pcg_iterations": info.pcg_iterations, "newton": info.newton_iterations, "ms": info.ms,
                           "rel_residual": info.rel_residual, "converged": rc == 0}
        sim.ctx.call("sf_rho_set", dp(prev_rho))
        sim.ctx.call("sf_phi_set", dp(prev_phi))
        cold["iteration_ratio"] = cold["zero"]["pcg_iterations"] / max(1, cold["learned"]["pcg_iterations"])
        cold["time_ratio"] = cold["zero"]["ms"] / max(1e-9, cold["learned"]["ms"])
        # (2) the time loop's states
        rows = []
        mat = sim.materials[0]
        for _ in range(4):
            sim.sample_sources()
            mat.update()
            row = {}
            for start in ("zero", "current", "learned"):
                info, rc = solve_from(sim, start, net)
                row[start] = (info.pcg_iterations, info.newton_iterations, info.ms, info.rel_residual, rc)
            rows.append(row)
            sim.update_fields()
            sim.it += 1

        def agg(key, k):
            return float(np.mean([r[key][k] for r in rows]))
        out = {"samples": net.num_samples, "configs": args.learned_configs, "train_steps": args.learned_train_steps,
               "final_loss": loss, "generation_s": round(t_gen, 2), "training_s": round(t_train, 2),
               "cold_initial_solve": cold, "states": len(rows),
               "plain": {"pcg_iterations": agg("zero", 0), "newton": agg("zero", 1), "ms": agg("zero", 2)},
               "warm": {"pcg_iterations": agg("current", 0), "newton": agg("current", 1), "ms": agg("current", 2)},
               "learned": {"pcg_iterations": agg("learned", 0), "newton": agg("learned", 1), "ms": agg("learned", 2)},
               "worst_rel_residual": max(max(r[k][3] for k in r) for r in rows),
               "all_converged": all(r[k][4] == 0 for r in rows for k in r) and cold["zero"]["converged"]
               and cold["learned"]["converged"]}
        out["iteration_ratio_plain_over_learned"] = out["plain"]["pcg_iterations"] / max(1.0, out["learned"]["pcg_iterations"])
        out["network_throughput"] = cnn_throughput(sim.ctx, m.ni, m.nj)
        return out, net
    except Exception:
        net.close()
        raise


def _rel_errs(g, r):
    """max and RMS error of g against r, relative to the largest / RMS magnitude of r"""
    import numpy as np
    d = np.asarray(g, np.float64) - np.asarray(r, np.float64)
    return {"max": float(np.abs(d).max() / max(np.abs(r).max(), 1e-300)),
            "rms": float(np.sqrt(np.mean(d * d)) / max(np.sqrt(np.mean(np.asarray(r, np.float64) ** 2)), 1e-300))}


def cpu_baseline(sim, deck, args):
    """The oracle (the C++ FP64 restatement of the reference CPU path: threaded mover, serial
    deposit and moment sampling, Newton with the reference's Gauss-Seidel) timed on this host, and
    used as the checker of the device on a sample of the same workload:

    * push+deposit particles/s on a 2e6-particle sample of the benchmark's population, mesh and E
      field, with every host core, and with one thread on a 2e5-particle sample;
    * one warm Newton+GS potential solve of the benchmark mesh from the device's last rho and phi
      (the reference's own tolerances), as the time loop would call it;
    * C1 (81x41, the reference's CPU deck) steps 101-200 with every host core, by phase;
    * errors of the device against the oracle at the benchmark's size: one push+deposit step of a
      2e5-particle sample from identical states (n, u, v, w, rho = q n, and T after computeFields),
      and phi from one FP64 Newton step with a direct solve of the reference's assembled system
      from the device's solution (Newton's quadratic convergence makes that step the error)."""
    import numpy as np
    import scipy.sparse as sps
    import scipy.sparse.linalg as spla
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import geometry, oracle_for
    from starfish_amd import deck as D
    from starfish_amd.plugin import Simulation
    nproc = args.cpu_threads or os.cpu_count() or 1
    geo = sim.geo
    ei, ej = sim.efield()
    mat = sim.materials[0]
    p = mat.particles()
    n_all = len(p["x"])
    out = {}

    def push_rate(threads, n, target_s, max_steps):
        o = oracle_for(deck, geo, None, nthreads=threads)
        o.set_field("efi", ei)
        o.set_field("efj", ej)
        sel = np.linspace(0, n_all - 1, n).astype(np.int64)
        pos = np.stack([p["x"][sel], p["y"][sel], np.zeros(n)], 1)
        vel = np.stack([p["u"][sel], p["v"][sel], p["w"][sel]], 1)
        o.load_particles(0, pos, vel)
        pushed, t_total, steps = 0, 0.0, 0
        while t_total < target_s and steps < max_steps:
            npre = o.num_particles(0)
            t0 = time.perf_counter()
            o.move()
            o.update_species_fields()
            t_total += time.perf_counter() - t0
            pushed += npre
            steps += 1
        o.close()
        return pushed / t_total, steps, t_total

    n_big = min(n_all, 2_000_000)
    rate_all, st_all, t_all = push_rate(nproc, n_big, 10.0, 40)
    n_one = min(n_all, 200_000)
    rate_one, st_one, t_one = push_rate(1, n_one, 4.0, 20)
    out["push_deposit"] = {
        "all_cores": {"particles_per_s": rate_all, "threads": nproc, "sample": n_big, "steps": st_all,
                      "seconds": round(t_all, 2)},
        "one_thread": {"particles_per_s": rate_one, "threads": 1, "sample": n_one, "steps": st_one,
                       "seconds": round(t_one, 2)}}

    # one warm Newton+GS solve of the benchmark mesh (Solver.solveNonLinearNR + LinearSolverGS)
    o = oracle_for(deck, geo, None, nthreads=nproc)
    o.set_field("rho", sim.rho())
    o.set_field("phi", sim.phi())
    t0 = time.perf_counter()
    o.solve_potential()
    t_solve = time.perf_counter() - t0
    o.close()
    out["newton_gs_warm_solve_ms"] = t_solve * 1e3

    # C1: the reference's CPU deck, steps 101-200, by phase (sources, move, deposit+samples, solve)
    c1 = D.config_c1()
    geo1, st1 = geometry(c1)
    o = oracle_for(c1, geo1, st1, nthreads=nproc)
    o.add_source(0, c1.boundary("inlet").points, False, c1.sources[0].mdot, c1.sources[0].v_drift,
                 c1.sources[0].temperature)
    o.initial_solve()
    o.step(100)
    t_a = o.timers()
    w0 = time.perf_counter()
    o.step(100)
    wall = time.perf_counter() - w0
    t_b = o.timers()
    np_c1 = o.num_particles(0)
    o.close()
    ph = (t_b - t_a) / 100 * 1e3
    out["c1_steps_101_200"] = {"threads": nproc, "ms_per_step": wall / 100 * 1e3, "particles_at_200": int(np_c1),
                               "ms_sources": ph[0], "ms_move": ph[1], "ms_deposit_samples": ph[2],
                               "ms_solve": ph[3]}

    # errors at the benchmark's size: one step of a 2e5-particle sample from identical states
    ns = min(n_all, 200_000)
    rng = np.random.default_rng(7)
    sel = np.sort(rng.choice(n_all, ns, replace=False))
    ids = np.arange(ns, dtype=np.int32)
    geo2, st2 = geometry(deck)
    o = oracle_for(deck, geo2, st2, nthreads=nproc)
    o.set_field("efi", ei)
    o.set_field("efj", ej)
    o.load_particles(0, np.stack([p["x"][sel], p["y"][sel], np.zeros(ns)], 1),
                     np.stack([p["u"][sel], p["v"][sel], p["w"][sel]], 1), ids=ids)
    small = Simulation(deck, capacity=2 * ns, track_ids=True)
    try:
        small.set_efield(ei, ej)
        sm = small.materials[0]
        sm.add_particles(p["x"][sel], p["y"][sel], p["u"][sel], p["v"][sel], p["w"][sel], ids=ids, rewind=False)
        # the first step pushes the appended sample and sorts it; the second runs the tiled kernels
        # from states that agree to FP32 (the oracle rounds its state as the device stores it)
        o.move()
        o.update_species_fields()
        o.round_state_fp32(0)
        sm.update()
        o.move()
        o.update_species_fields()
        sm.update()
        err = {"particles": {"device": sm.getNp(), "oracle": int(o.num_particles(0))}}
        for name in ("nd", "u", "v", "w"):
            err[name] = _rel_errs(sm.field(name), o.field(name, 0))
        err["rho"] = err["nd"]             # one species: rho = q n on both sides
        err["t"] = _rel_errs(sm.field("t"), o.field("t", 0))
    finally:
        small.close()
    # phi: one FP64 Newton step with a direct solve of the reference's assembled system
    A = o.operator(0).tocsr()
    fixed = o.fixed_nodes().astype(bool)
    o.close()
    phi, rho = sim.phi(), sim.rho()
    m = deck.mesh
    ni, nj = m.ni, m.nj
    nn = ni * nj
    I, J = np.arange(nn) % ni, np.arange(nn) // ni
    ft = geo.face_type
    right, top, left, bottom = ft[:nj], ft[nj:nj + ni], ft[nj + ni:2 * nj + ni], ft[2 * nj + ni:]
    face = ((I == 0) & (left[J] != 4) & (left[J] != 0)) | ((I == ni - 1) & (right[J] != 4) & (right[J] != 0)) | \
           ((J == 0) & (bottom[I] != 4) & (bottom[I] != 0)) | ((J == nj - 1) & (top[I] != 4) & (top[I] != 0))
    s = deck.solver
    eps = 8.85418782e-12 * s.eps_r
    b = -rho / eps
    b[fixed] = geo.node_bc[fixed]
    bx = np.where(fixed, 0, 1.60217646e-19 / eps * s.n0 * np.exp((phi - s.phi0) / s.Te0))
    P = bx / s.Te0
    b = b + bx
    b[face & ~fixed] = 0
    P[face & ~fixed] = 0
    F = A @ phi - b
    y = spla.spsolve((A - sps.diags(P)).tocsc(), F)
    err["phi"] = {"max": float(np.abs(y).max() / np.abs(phi).max()),
                  "rms": float(np.sqrt(np.mean(y * y)) / np.sqrt(np.mean(phi * phi))),
                  "max_volts": float(np.abs(y).max())}
    out["errors_vs_oracle"] = err
    base = {"value": rate_all, "unit": "particles/s", "cores": nproc, "kind": "port",
            "sample": "%d particles of the same population on the same 1024x512 mesh and E field, %d steps of "
                      "KineticMaterial.moveParticles (%d mover threads) + serial updateFields/updateSamples, %.1f s"
                      % (n_big, st_all, nproc, t_all)}
    base.update(out)
    return base


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the hot path on a single MI355X "
              "(--gpus %d is not supported; multi-GPU sharding is described in DESIGN.md)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import ctypes as C
    import numpy as np
    from starfish_amd import deck as D
    from starfish_amd.plugin import Simulation

    if args.workload != "c2":
        print(json.dumps(other_workload(args)), flush=True)
        return
    if args.cnn_only:
        from starfish_amd.plugin import Context
        ctx = Context(0)
        try:
            print(json.dumps({"cnn": cnn_throughput(ctx, 1024, 512)}), flush=True)
        finally:
            ctx.close()
        return
    npart = int(args.particles)
    deck = D.config_c2(steps=args.steps, particles=npart)
    sim = Simulation(deck, capacity=int(npart * 1.6) + 1_000_000, device_sources=not args.host_sources)
    mat = sim.materials[0]
    nparts, drift, temp = deck.preload
    mat.preload(nparts, drift, temp, seed=1)
    t0 = time.perf_counter()
    rc = sim.initial_solve()
    rho0 = sim.rho()
    cold = sim.solve_infos[-1]
    cold_info = {"ms": cold.ms, "newton": cold.newton_iterations, "pcg_iterations": cold.pcg_iterations,
                 "rel_residual": cold.rel_residual, "converged": rc == 0, "wall_s": time.perf_counter() - t0}
    sim.ctx.call("sf_set_timing", 1)
    sim.ctx.call("sf_species_set_sort_interval", mat.index, args.sort_interval)
    for _ in range(args.warmup):
        sim.step()
    sim.ctx.call("sf_ctx_synchronize")
    kern = np.zeros(4)
    pushed = 0
    # per-kernel totals: (ms, particles, launches): the tiled push (the particles it pushed itself),
    # the slow lists it left to the full reference loop (push_tile_slow), and the sort-step scatter
    k_push, k_queue, k_scatter = [0.0, 0, 0], [0.0, 0, 0], [0.0, 0, 0]
    solve_ms, pcg_its, newton_its = [], [], []
    worst_res = 0.0
    tail_ms = 0.0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        npre = mat.getNp()
        inj = C.c_int64(sim.sample_sources())
        mat.update()
        tm = np.zeros(4)
        sim.ctx.call("sf_last_timings", tm.ctypes.data_as(C.POINTER(C.c_double)))
        split = np.zeros(3)
        sim.ctx.call("sf_last_push_timings", split.ctypes.data_as(C.POINTER(C.c_double)))
        qd = np.zeros(2)
        sim.ctx.call("sf_last_push_queue", qd.ctypes.data_as(C.POINTER(C.c_double)))
        kern += tm
        npush = npre + inj.value
        nsorted = npush - int(split[2])     # particles in the cell-sorted part of the array
        nqueued = int(qd[1])
        pushed += npush
        tail_ms += split[1]
        k_push[0] += split[0]
        k_push[1] += nsorted - nqueued
        k_push[2] += 1
        k_queue[0] += qd[0]
        k_queue[1] += nqueued
        k_queue[2] += 1
        if tm[2] > 0:
            k_scatter[0] += tm[2]
            k_scatter[1] += npush
            k_scatter[2] += 1
        sim.update_fields()
        info = sim.solve_infos[-1]
        solve_ms.append(info.ms)
        pcg_its.append(info.pcg_iterations)
        newton_its.append(info.newton_iterations)
        worst_res = max(worst_res, info.rel_residual)
        sim.it += 1
    sim.ctx.call("sf_ctx_synchronize")
    wall = time.perf_counter() - t_start
    particle_s = kern.sum() * 1e-3
    value = pushed / particle_s

    def rate(k, nbytes):
        return nbytes * k[1] / (k[0] * 1e-3) / 1e9 if k[0] > 0 else 0.0

    per_launch = k_push[1] / max(1, k_push[2])
    bpp = measured_bytes_per_particle("push_tile_fast_xy")
    achieved = rate(k_push, BYTES_PUSH)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": round(bpp * per_launch) if bpp else None,
                "traffic_bytes_per_particle": round(bpp, 2) if bpp else None,
                "kernel": "push_tile_fast_xy", "bytes_per_particle": BYTES_PUSH,
                "particles_per_launch": round(per_launch),
                "algorithmic_bytes_per_launch": round(BYTES_PUSH * per_launch),
                "ms_per_launch": round(k_push[0] / max(1, k_push[2]), 4), "launches": k_push[2],
                "per_kernel_GBs": {"push_tile_fast_xy": round(achieved, 1),
                                   "push_tile_slow": round(rate(k_queue, BYTES_PUSH), 1),
                                   "copy_scatter": round(rate(k_scatter, BYTES_SCATTER), 1)},
                "slow_particles_per_step": round(k_queue[1] / max(1, k_queue[2])),
                "slow_ms_per_step": round(k_queue[0] / max(1, k_queue[2]), 4),
                "particle_phase_state_GBs": round(BYTES_STATE * pushed / particle_s / 1e9, 1)}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sim, deck, args)
    learned = None
    if not args.no_learned:
        learned, net = learned_guess(sim, deck, args, rho0)
        net.close()
    out = {
        "metric": "particle push+deposit/s (2D ES-PIC, 1024x512 XY, 2e7 particles)",
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+i64",
        "data": "synthetic (preloaded drifting Maxwellian, seed 1, plus boundary injection)",
        "config": {"workload": "ES-PIC O+ past a -100 V cylinder, XY 1024x512 nodes dx=1e-3 m, "
                               "%d particles preloaded, Boltzmann electrons n0=1e10 kTe=2 eV, dt=1e-7 s" % npart,
                   "particles_final": mat.getNp(), "state": "int32 cell + f32 offsets + f32 velocity (24 B)",
                   "precision": "push in FP32 relative to the cell away from segments and faces, FP64 reference "
                                "loop elsewhere; deposit rounded once in FP64 into int64 fixed point; PCG FP32 "
                                "with FP64 residual refinement and FP64 Newton",
                   "injection": ("host, java.util.Random replay" if args.host_sources else
                                 "sampled on the device (Philox4x32-10)"),
                   "deposit": "int64 fixed point, 7 moment channels", "solver": "Newton + FP32 PCG (block red-black SGS preconditioner) "
                   "with FP64 residual refinement, rel. residual 1e-8 every step"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "full_step_particles_per_s": pushed / wall,
        "particle_phase_ms": {"push_tiles": k_push[0] / args.steps, "push_slow": k_queue[0] / args.steps,
                              "push_appended": tail_ms / args.steps, "scan": kern[1] / args.steps,
                              "scatter": kern[2] / args.steps, "normalize": kern[3] / args.steps},
        "pcg": {"ms_per_solve": float(np.mean(solve_ms)), "pcg_iterations_per_solve": float(np.mean(pcg_its)),
                "newton_per_solve": float(np.mean(newton_its)), "worst_rel_residual": worst_res,
                "rel_tol": deck.solver.rel_tol, "cold_start": cold_info, "learned_guess": learned},
    }
    sim.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
