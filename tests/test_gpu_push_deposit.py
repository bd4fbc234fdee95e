"""One-step push + deposit parity on the GPU against the FP64 oracle from identical particle
sets (KineticMaterial.updateFields, core/materials/KineticMaterial.java:117-197, 1553-1595).

Tolerance: 1e-5 relative (positions relative to the domain extent, velocities to the largest
speed, node fields to their maximum).  The first step runs the unsorted-tail kernel, the second
the LDS-tiled kernel on the cell-sorted array."""
import numpy as np
import pytest

from helpers import geometry, open_cell_particles, oracle_for, rel_err
from starfish_amd import deck as D

TOL = 1e-5


def _compare_particles(deck, gp, op):
    m = deck.mesh
    gi = np.argsort(gp["id"])
    oi = np.argsort(op["id"])
    gid, oid = gp["id"][gi], op["id"][oi]
    common = np.intersect1d(gid, oid)
    mism = len(gid) + len(oid) - 2 * len(common)
    gsel = np.isin(gid, common)
    osel = np.isin(oid, common)
    ext = max((m.ni - 1) * m.dh[0], (m.nj - 1) * m.dh[1])
    ex = rel_err(gp["x"][gi][gsel], op["pos"][oi][osel][:, 0], ext)
    ey = rel_err(gp["y"][gi][gsel], op["pos"][oi][osel][:, 1], ext)
    vmax = np.abs(op["vel"]).max()
    eu = rel_err(gp["u"][gi][gsel], op["vel"][oi][osel][:, 0], vmax)
    ev = rel_err(gp["v"][gi][gsel], op["vel"][oi][osel][:, 1], vmax)
    ew = rel_err(gp["w"][gi][gsel], op["vel"][oi][osel][:, 2], vmax)
    return mism, max(ex, ey), max(eu, ev, ew)


@pytest.mark.gpu
@pytest.mark.parametrize("interval", [1, 4])
def test_push_deposit_one_step_matches_oracle(interval):
    from starfish_amd.plugin import Simulation
    n = 20000
    deck = D.config_c1()
    geo, st = geometry(deck)
    o = oracle_for(deck, geo, st)
    o.initial_solve()                        # E of the charged cylinder with no ions
    efi, efj = o.field("efi"), o.field("efj")
    x, y, v, _, _ = open_cell_particles(deck, geo, n, seed=3)
    ids = np.arange(n, dtype=np.int32)
    pos3 = np.stack([x, y, np.zeros(n)], 1)
    o.load_particles(0, pos3, v, ids=ids, rewind=False)

    sim = Simulation(deck, capacity=4 * n, track_ids=True)
    try:
        sim.set_efield(efi, efj)
        mat = sim.materials[0]
        sim.ctx.call("sf_species_set_sort_interval", mat.index, interval)
        mat.add_particles(x, y, v[:, 0], v[:, 1], v[:, 2], ids=ids, rewind=False)
        for step in range(3):
            o.move()
            o.update_species_fields()
            mat.update()
            gp = mat.particles(ids=True)
            op = o.particles(0)
            print("step", step, "gpu np", len(gp["id"]), "oracle np", len(op["id"]), flush=True)
            mism, epos, evel = _compare_particles(deck, gp, op)
            assert mism <= 2, "particle fates differ for %d particles" % mism
            assert epos < TOL, "position error %g" % epos
            assert evel < TOL, "velocity error %g" % evel
            for name in ("nd", "u", "v", "w", "count-sum", "u-sum", "uu-sum", "ww-sum"):
                g = mat.field(name)
                r = o.field(name, 0)
                e = rel_err(g, r)
                assert e < TOL, "%s error %g" % (name, e)
            # mover sums (mass, momentum, |v|) as in KineticMaterial.moveParticles
            gs, os_ = mat.stats(), o.stats(0)
            assert abs(gs[0] - os_[0]) <= 2e-5 * abs(os_[0])
            assert np.abs(gs[1:] - os_[1:]).max() <= 1e-4 * np.abs(os_[1:]).max()
            # cell-sorted order after a sort step (counting sort)
            if interval == 1:
                key = np.floor(gp["lc1"]).clip(0, deck.mesh.nj - 2) * (deck.mesh.ni - 1) + \
                    np.floor(gp["lc0"]).clip(0, deck.mesh.ni - 2)
                assert np.all(np.diff(key) >= 0), "particles are not sorted by cell"
    finally:
        sim.close()


@pytest.mark.gpu
def test_fields_do_not_depend_on_the_sort_interval():
    """Pushes are independent per particle and deposits are integer sums: sorting every step,
    every 3rd step or every 5th step gives bit-identical fields and particle sets."""
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    geo, st = geometry(deck)
    o = oracle_for(deck, geo, st)
    o.initial_solve()
    n = 60000
    x, y, v, _, _ = open_cell_particles(deck, geo, n, seed=5)
    ids = np.arange(n, dtype=np.int32)
    res = []
    for interval in (1, 3, 5):
        sim = Simulation(deck, capacity=2 * n, track_ids=True)
        try:
            sim.set_efield(o.field("efi"), o.field("efj"))
            mat = sim.materials[0]
            sim.ctx.call("sf_species_set_sort_interval", mat.index, interval)
            mat.add_particles(x, y, v[:, 0], v[:, 1], v[:, 2], ids=ids, rewind=False)
            for _ in range(7):
                mat.update()
            gp = mat.particles(ids=True)
            order = np.argsort(gp["id"])
            res.append([mat.field(k) for k in ("nd", "u", "count-sum", "uu-sum", "mpc-sum")] +
                       [gp[k][order] for k in ("id", "x", "y", "u", "v", "w")])
        finally:
            sim.close()
    for other in res[1:]:
        for a, b in zip(res[0], other):
            assert np.array_equal(a, b)


@pytest.mark.gpu
def test_deposit_is_bit_identical_run_to_run():
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    geo, st = geometry(deck)
    n = 50000
    x, y, v, _, _ = open_cell_particles(deck, geo, n, seed=11)
    outs = []
    for rep in range(2):
        sim = Simulation(deck, capacity=2 * n)
        try:
            mat = sim.materials[0]
            mat.add_particles(x, y, v[:, 0], v[:, 1], v[:, 2], rewind=False)
            mat.update()
            mat.update()
            outs.append([mat.field(k) for k in ("nd", "u", "v", "w", "count-sum", "uu-sum")])
        finally:
            sim.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("interval", [1, 4])
def test_boris_push_matches_oracle(interval):
    """A magnetic field of up to 0.3 T that is zero on the left third of the domain: the
    addParticle rewind and three pushes use Boris where the gathered B is nonzero and the
    electric kick elsewhere (UpdateVelocityBoris, KineticMaterial.java:336-353, 779-794, 847-893).
    The GPU keeps B in FP32, so the oracle gets the same FP32-rounded values; tolerance 1e-5."""
    from starfish_amd.plugin import Simulation
    n = 20000
    deck = D.config_c1()
    geo, st = geometry(deck)
    m = deck.mesh
    o = oracle_for(deck, geo, st)
    o.initial_solve()
    efi, efj = o.field("efi"), o.field("efj")
    I, J = np.meshgrid(np.arange(m.ni), np.arange(m.nj))
    on = (I > m.ni // 3).ravel()
    bfi = (0.3 * np.cos(np.pi * I / (m.ni - 1))).ravel() * on
    bfj = (0.2 * np.sin(np.pi * J / (m.nj - 1))).ravel() * on
    bfi, bfj = bfi.astype(np.float32).astype(np.float64), bfj.astype(np.float32).astype(np.float64)
    o.set_field("bfi", bfi)
    o.set_field("bfj", bfj)
    x, y, v, _, _ = open_cell_particles(deck, geo, n, seed=21)
    ids = np.arange(n, dtype=np.int32)
    o.load_particles(0, np.stack([x, y, np.zeros(n)], 1), v, ids=ids, rewind=True)
    sim = Simulation(deck, capacity=4 * n, track_ids=True)
    try:
        sim.set_efield(efi, efj)
        sim.set_bfield(bfi, bfj)
        mat = sim.materials[0]
        sim.ctx.call("sf_species_set_sort_interval", mat.index, interval)
        mat.add_particles(x, y, v[:, 0], v[:, 1], v[:, 2], ids=ids, rewind=True)
        mism, epos, evel = _compare_particles(deck, mat.particles(ids=True), o.particles(0))
        print("after the rewind: velocity error %.2e" % evel, flush=True)
        assert mism == 0 and evel < TOL
        for step in range(3):
            o.move()
            o.update_species_fields()
            mat.update()
            gp, op = mat.particles(ids=True), o.particles(0)
            mism, epos, evel = _compare_particles(deck, gp, op)
            print("step %d: np %d / %d, position error %.2e, velocity error %.2e" %
                  (step, len(gp["id"]), len(op["id"]), epos, evel), flush=True)
            assert mism <= 2 and epos < TOL and evel < TOL
            for name in ("nd", "u", "v", "w", "ww-sum"):
                e = rel_err(mat.field(name), o.field(name, 0))
                assert e < TOL, "%s error %g" % (name, e)
        # Boris turned w for the particles in the field (the electric kick never changes w in XY)
        w0 = v[op["id"], 2]
        assert np.mean(op["vel"][:, 2] != w0) > 0.3
        sim.set_bfield(None, None)
        mat.update()          # the same species continues on the electric-only path
    finally:
        sim.close()
