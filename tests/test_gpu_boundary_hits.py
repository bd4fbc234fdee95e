"""Per-boundary impact tallies on the GPU against the oracle (Material.addSurfaceMassDeposit
from KineticMaterial.ProcessBoundary, core/materials/KineticMaterial.java:596-600), and the
stats CSV flux columns built on them (Material.getFluxInst, Material.java:806-813;
core/diagnostics/StatsModule.java:141-147): the same particle sets are pushed by the device
and the oracle, and the counts per boundary agree."""
import csv
import ctypes as C
import math

import numpy as np
import pytest

from helpers import geometry, open_cell_particles, oracle_for
from starfish_amd import deck as D


def _hits(sim, mat, clear=True):
    nb = len(sim.deck.boundaries)
    h = np.zeros(nb, dtype=np.int64)
    sim.ctx.call("sf_species_boundary_hits", mat.index, nb, h.ctypes.data_as(C.POINTER(C.c_int64)), 1 if clear else 0)
    return h


@pytest.mark.gpu
def test_aimed_ring_is_tallied_and_written_as_flux(tmp_path):
    from starfish_amd.output import StatsWriter, _boundary_area
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    names = [b.name for b in deck.boundaries]
    sim = Simulation(deck, capacity=10_000)
    try:
        sim.set_efield(np.zeros(sim.nn), np.zeros(sim.nn))
        mat = sim.materials[0]
        rng = np.random.default_rng(3)
        n = 300
        ang = rng.uniform(math.radians(10), math.radians(170), n)
        r, speed = 0.08, 0.04 / deck.dt
        mat.add_particles(r * np.cos(ang), r * np.sin(ang), -speed * np.cos(ang), -speed * np.sin(ang), np.zeros(n),
                          rewind=False)
        stats = StatsWriter(sim, str(tmp_path / "stats.csv"))
        stats.start()
        assert mat.update() == 0
        assert mat.getNp() == 0
        stats.save_stats()          # reads and clears the tallies
        stats.finish()
        rows = list(csv.reader(open(tmp_path / "stats.csv")))
        col = rows[0].index("flux.cylinder.%s (#/m^2/s)" % mat.name)
        area = _boundary_area(deck.boundary("cylinder"), deck.mesh.domain)
        expect = n * deck.materials[0].spwt / (area * deck.dt)
        assert abs(float(rows[1][col]) / expect - 1) < 1e-3          # %.4g
        assert float(rows[1][rows[0].index("flux.inlet.%s (#/m^2/s)" % mat.name)]) == 0.0
        assert _hits(sim, mat)[names.index("cylinder")] == 0        # cleared by the stats line
    finally:
        sim.close()


@pytest.mark.gpu
def test_boundary_hits_match_oracle_in_the_c1_field():
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    geo, st = geometry(deck)
    o = oracle_for(deck, geo, st)
    o.initial_solve()
    efi, efj = o.field("efi"), o.field("efj")
    n = 100_000
    x, y, v, _, _ = open_cell_particles(deck, geo, n, seed=21, drift=(0.0, 0.0, 0.0), temp=1e6)
    ids = np.arange(n, dtype=np.int32)
    o.load_particles(0, np.stack([x, y, np.zeros(n)], 1), v, ids=ids, rewind=False)
    sim = Simulation(deck, capacity=2 * n, track_ids=True)
    try:
        sim.set_efield(efi, efj)
        mat = sim.materials[0]
        mat.add_particles(x, y, v[:, 0], v[:, 1], v[:, 2], ids=ids, rewind=False)
        for step in range(3):
            o.move()
            o.update_species_fields()
            o.round_state_fp32(0)
            assert mat.update() == 0
        hd, ho = _hits(sim, mat), o.boundary_hits(0)
        print("absorbed per boundary: device %s, oracle %s" % (hd.tolist(), ho.tolist()), flush=True)
        assert hd[deck.boundaries.index(deck.boundary("cylinder"))] > 100
        # fates agree particle for particle up to FP32-storage ties at a surface
        assert np.abs(hd - ho).sum() <= 2
        assert abs(mat.getNp() - o.num_particles(0)) <= 2
    finally:
        sim.close()
        o.close()
