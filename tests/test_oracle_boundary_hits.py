"""Pins of the oracle's per-boundary impact tally (Material.addSurfaceMassDeposit from
KineticMaterial.ProcessBoundary, core/materials/KineticMaterial.java:596-600; read back through
Material.getFluxInst for StatsModule's flux columns, Material.java:806-813) by trajectories
whose fate is known: particles aimed radially at the cylinder of the tutorial deck are all
absorbed on it; particles moving away from it, or leaving through the left mesh face (no
segment), are not counted; reading with clear restarts the count."""
import math

import numpy as np

from helpers import geometry, oracle_for
from starfish_amd import deck as D


def _ring(n, r, speed, inward, rng):
    ang = rng.uniform(math.radians(10), math.radians(170), n)
    pos = np.stack([r * np.cos(ang), r * np.sin(ang), np.zeros(n)], 1)
    d = -1.0 if inward else 1.0
    vel = np.stack([d * speed * np.cos(ang), d * speed * np.sin(ang), np.zeros(n)], 1)
    return pos, vel


def test_aimed_particles_are_tallied_on_the_cylinder_only():
    deck = D.config_c1()
    geo, st = geometry(deck)
    names = [b.name for b in deck.boundaries]
    sb = list(geo.seg_bnd[:geo.nseg])
    assert geo.nbnd == len(names) and sb.count(names.index("cylinder")) == 20 and sb.count(names.index("inlet")) == 1
    o = oracle_for(deck, geo, st)
    nn = deck.mesh.ni * deck.mesh.nj
    o.set_field("efi", np.zeros(nn))
    o.set_field("efj", np.zeros(nn))
    rng = np.random.default_rng(3)
    # 0.03 m outside r = 0.05 m, moving 0.04 m per step toward the axis: every one crosses the surface
    pin, vin = _ring(300, 0.08, 0.04 / deck.dt, True, rng)
    pout, vout = _ring(200, 0.08, 0.02 / deck.dt, False, rng)
    # near the left face, moving out of the mesh (a mesh face, not a segment)
    pl = np.stack([np.full(100, -0.145), rng.uniform(0.1, 0.3, 100), np.zeros(100)], 1)
    vl = np.stack([np.full(100, -0.02 / deck.dt), np.zeros(100), np.zeros(100)], 1)
    o.load_particles(0, np.concatenate([pin, pout, pl]), np.concatenate([vin, vout, vl]), rewind=False)
    o.move()
    hits = o.boundary_hits(0)
    assert hits[names.index("cylinder")] == 300
    assert hits[names.index("inlet")] == 0
    assert o.num_particles(0) == 200            # the outward ring survives, the rest is gone
    # the count accumulates until a read clears it
    o.move()
    assert o.boundary_hits(0, clear=True)[names.index("cylinder")] == 300
    assert o.boundary_hits(0)[names.index("cylinder")] == 0
