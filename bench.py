#!/usr/bin/env python
"""Benchmark of the Starfish ES-PIC hot path on one MI355X.

Workload (default): O+ flow past a -100 V cylinder on a 1024x512-node XY mesh (dx = 1e-3 m),
2e7 macroparticles preloaded as a drifting Maxwellian (7000 m/s, 1000 K) plus left-edge
injection at 1e20 m^-2 s^-1, Boltzmann electrons (n0 = 1e10 m^-3, kTe = 2 eV), dt = 1e-7 s.
A step is one Starfish main-loop iteration: source injection, push + deposit + counting sort,
rho, Newton-PCG potential solve to a 1e-8 relative residual, E = -grad(phi).

Output: one JSON line.  `value` is particles pushed and deposited per second of device time of
the particle phase (push/deposit pass, cell scan, push/scatter pass, normalisation), summed over
the timed steps with HIP events on the library's stream.  The full-step rate (all phases, host
clock) and the potential-solve cost are reported beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
BYTES_PUSH_INPLACE = 48.0   # every step: read + write the 24 B particle state in its slot
BYTES_SCATTER = 48.0        # sort step: read the 24 B pushed state (its cell is the key), write it sorted
BYTES_STATE = 48.0          # read + write of the 24 B particle state (minimal push + deposit)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pmc_traffic_c2_bench.txt")
TRAFFIC_PROFILE_PARTICLES = 19.8e6   # average particles per push dispatch in the profiled runs


def measured_bytes_per_particle(kernel):
    """HBM bytes per particle of one push kernel, from the committed FETCH_SIZE + WRITE_SIZE summary
    (already scaled by the copy-kernel calibration); None when the summary or the kernel is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                parts = line.split()
                if parts and parts[0] == kernel:
                    return float(parts[-1]) / TRAFFIC_PROFILE_PARTICLES
    except OSError:
        pass
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--particles", type=float, default=2e7)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--sort-interval", type=int, default=4)
    ap.add_argument("--learned-configs", type=int, default=8,
                    help="C2-style runs (random cylinder radius / bias) that provide training snapshots")
    ap.add_argument("--learned-train-steps", type=int, default=600)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--host-sources", action="store_true",
                    help="sample the boundary injection on the host with the reference's random stream "
                         "(default: sampled on the device)")
    return ap.parse_args()


def learned_guess(sim, deck, args):
    """Trains the learned initial guess on snapshots of C2-style runs (cylinder radius
    U(0.02, 0.1) m, bias U(-200, 0) V; none is the benchmark's cylinder) and compares, on the
    benchmark's own states, the solve from phi = 0 (plain), from the previous step's phi (warm)
    and from the network's guess (learned).  Every solve runs to the same residual tolerance."""
    import ctypes as C
    import numpy as np
    from starfish_amd.learned import LearnedGuess, generate_c2_snapshots, solve_from
    t0 = time.perf_counter()
    m = deck.mesh
    net = LearnedGuess(sim.ctx, max_batch=4, seed=1)
    try:
        generate_c2_snapshots(net, args.learned_configs, 24, 3, int(args.particles), m.ni, m.nj, seed=12345,
                              skip=3)
        t_gen = time.perf_counter() - t0
        t1 = time.perf_counter()
        loss = 0.0
        for r in range(4):
            loss = net.train(args.learned_train_steps // 4, batch=4, lr=1e-3 if r < 3 else 3e-4)
        sim.ctx.call("sf_ctx_synchronize")
        t_train = time.perf_counter() - t1
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
               "states": len(rows),
               "plain": {"pcg_iterations": agg("zero", 0), "newton": agg("zero", 1), "ms": agg("zero", 2)},
               "warm": {"pcg_iterations": agg("current", 0), "newton": agg("current", 1), "ms": agg("current", 2)},
               "learned": {"pcg_iterations": agg("learned", 0), "newton": agg("learned", 1), "ms": agg("learned", 2)},
               "worst_rel_residual": max(max(r[k][3] for k in r) for r in rows),
               "all_converged": all(r[k][4] == 0 for r in rows for k in r)}
        out["iteration_ratio_plain_over_learned"] = out["plain"]["pcg_iterations"] / max(1.0, out["learned"]["pcg_iterations"])
        return out, net
    except Exception:
        net.close()
        raise


def cpu_baseline(sim, deck, threads, target_s=15.0):
    """Times the oracle's restatement of the reference CPU path (threaded mover, serial deposit and
    moment sampling) on a sample of the same particle population, mesh and E field."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_for
    geo = sim.geo
    o = oracle_for(deck, geo, None, nthreads=threads)
    ei, ej = sim.efield()
    o.set_field("efi", ei)
    o.set_field("efj", ej)
    mat = sim.materials[0]
    p = mat.particles()
    n_all = len(p["x"])
    n = min(n_all, 2_000_000)
    sel = np.linspace(0, n_all - 1, n).astype(np.int64)
    pos = np.stack([p["x"][sel], p["y"][sel], np.zeros(n)], 1)
    vel = np.stack([p["u"][sel], p["v"][sel], p["w"][sel]], 1)
    o.load_particles(0, pos, vel)
    pushed, t_total, steps = 0, 0.0, 0
    while t_total < target_s and steps < 50:
        npre = o.num_particles(0)
        t0 = time.perf_counter()
        o.move()
        o.update_species_fields()
        t_total += time.perf_counter() - t0
        pushed += npre
        steps += 1
    o.close()
    return {"value": pushed / t_total, "unit": "particles/s", "cores": threads, "kind": "port",
            "sample": "%d particles of the same population on the same 1024x512 mesh and E field, %d steps of "
                      "KineticMaterial.moveParticles (%d mover threads) + serial updateFields/updateSamples, "
                      "%.1f s" % (n, steps, threads, t_total)}


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

    npart = int(args.particles)
    deck = D.config_c2(steps=args.steps, particles=npart)
    sim = Simulation(deck, capacity=int(npart * 1.6) + 1_000_000, device_sources=not args.host_sources)
    mat = sim.materials[0]
    nparts, drift, temp = deck.preload
    mat.preload(nparts, drift, temp, seed=1)
    t0 = time.perf_counter()
    rc = sim.initial_solve()
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
    # per-kernel totals: (ms, particles, launches) for the tiled push and the sort-step scatter
    k_push, k_scatter = [0.0, 0, 0], [0.0, 0, 0]
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
        kern += tm
        npush = npre + inj.value
        nsorted = npush - int(split[2])     # particles the tiled kernel pushed
        pushed += npush
        # the tiled kernels are timed alone, over the cell-sorted particles they push; the kernel
        # for particles appended since the last sort is in the particle phase but not here
        tail_ms += split[1]
        for k, ms, n in (((k_push, split[0], nsorted), (k_scatter, tm[2], npush)) if tm[2] > 0
                         else ((k_push, split[0], nsorted),)):
            k[0] += ms
            k[1] += n
            k[2] += 1
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

    kernels = [("push_deposit_tile_inplace", k_push, BYTES_PUSH_INPLACE),
               ("copy_scatter", k_scatter, BYTES_SCATTER)]
    dname, dk, dbytes = max(kernels, key=lambda e: e[1][0])
    achieved = rate(dk, dbytes)
    per_launch = dk[1] / max(1, dk[2])
    bpp = measured_bytes_per_particle(dname)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": round(bpp * per_launch) if bpp else None,
                "traffic_bytes_per_particle": round(bpp, 2) if bpp else None,
                "algorithmic_bytes_per_launch": round(dbytes * per_launch),
                "particles_per_launch": round(per_launch),
                "kernel": dname, "bytes_per_particle": dbytes,
                "ms_per_launch": round(dk[0] / max(1, dk[2]), 4), "launches": dk[2],
                "per_kernel_GBs": {name: round(rate(k, b), 1) for name, k, b in kernels},
                "particle_phase_state_GBs": round(BYTES_STATE * pushed / particle_s / 1e9, 1)}
    cpu = None
    if not args.no_cpu_baseline:
        thr = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(sim, deck, thr)
    learned = None
    if not args.no_learned:
        learned, net = learned_guess(sim, deck, args)
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
        "dtype": "f64",
        "data": "synthetic (preloaded drifting Maxwellian, seed 1, plus boundary injection)",
        "config": {"workload": "ES-PIC O+ past a -100 V cylinder, XY 1024x512 nodes dx=1e-3 m, "
                               "%d particles preloaded, Boltzmann electrons n0=1e10 kTe=2 eV, dt=1e-7 s" % npart,
                   "particles_final": mat.getNp(), "state": "int32 cell + f32 offsets + f32 velocity (24 B)",
                   "injection": ("host, java.util.Random replay" if args.host_sources else
                                 "sampled on the device (Philox4x32-10)"),
                   "deposit": "int64 fixed point, 7 moment channels", "solver": "Newton + FP32 PCG (block red-black SGS preconditioner) "
                   "with FP64 residual refinement, rel. residual 1e-8 every step"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "full_step_particles_per_s": pushed / wall,
        "particle_phase_ms": {"push_deposit": kern[0] / args.steps,
                              "push_deposit_appended": tail_ms / args.steps, "scan": kern[1] / args.steps,
                              "scatter": kern[2] / args.steps, "normalize": kern[3] / args.steps},
        "pcg": {"ms_per_solve": float(np.mean(solve_ms)), "pcg_iterations_per_solve": float(np.mean(pcg_its)),
                "newton_per_solve": float(np.mean(newton_its)), "worst_rel_residual": worst_res,
                "rel_tol": deck.solver.rel_tol, "cold_start": cold_info, "learned_guess": learned},
    }
    sim.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
