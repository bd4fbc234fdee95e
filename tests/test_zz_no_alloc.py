"""sf_particles_add_noalloc (staging in the library's fixed device array) against
sf_particles_add (per-call staging allocations) through the C-ABI: the same accepted particles,
slots, rewound velocities and ids, bit for bit, on two calls in a row, the second one smaller."""
import ctypes as C

import numpy as np
import pytest

from helpers import geometry, open_cell_particles
from starfish_amd import deck as D

FIELDS = ("x", "y", "lc0", "lc1", "u", "v", "w")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def _batch(deck, geo, n, seed):
    """n particles in open cells plus rejected ones: non-finite velocities and positions outside
    the mesh (KineticMaterial.java:1356-1361, DomainModule.java:106-117)."""
    x, y, v, _, _ = open_cell_particles(deck, geo, n, seed=seed)
    x, y = x.copy(), y.copy()
    v = np.ascontiguousarray(v)
    v[3, 0] = np.nan
    v[n // 2, 2] = np.inf
    x[7] = deck.mesh.x0[0] - 0.5
    y[n - 2] = deck.mesh.x0[1] + 10.0 * deck.mesh.nj * deck.mesh.dh[1]
    return x, y, v


def _add(sim, entry, x, y, v, ids):
    a = [np.ascontiguousarray(q, dtype=np.float64) for q in (x, y, v[:, 0], v[:, 1], v[:, 2])]
    ida = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
    return sim.ctx.call(entry, sim.materials[0].index, len(x), *[_dp(q) for q in a], _ip(ida), 1)


@pytest.mark.gpu
def test_noalloc_add_matches_add_on_two_calls():
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    geo, _ = geometry(deck)
    rng = np.random.default_rng(5)
    nn = deck.mesh.ni * deck.mesh.nj
    ei, ej = rng.normal(size=nn) * 300.0, rng.normal(size=nn) * 300.0
    sims = {e: Simulation(deck, capacity=20000, track_ids=True) for e in ("sf_particles_add", "sf_particles_add_noalloc")}
    try:
        for s in sims.values():
            s.set_efield(ei, ej)
        # first call: 9000 particles with ids; second: 1200 particles, ids from the species counter
        for n, seed, with_ids in ((9000, 11, True), (1200, 12, False)):
            x, y, v = _batch(deck, geo, n, seed)
            ids = np.arange(100000, 100000 + n, dtype=np.int32) if with_ids else None
            for e, s in sims.items():
                assert _add(s, e, x, y, v, ids) == 0
            a, b = (sims[e].materials[0] for e in ("sf_particles_add", "sf_particles_add_noalloc"))
            assert a.getNp() == b.getNp()
            pa, pb = a.particles(ids=True), b.particles(ids=True)
            for k in FIELDS + ("id",):
                assert np.array_equal(pa[k], pb[k]), k
        assert sims["sf_particles_add"].materials[0].getNp() == 9000 - 4 + 1200 - 4
        # both continue identically: one push + deposit from the added state
        for s in sims.values():
            assert s.materials[0].update() == 0
        a, b = (sims[e].materials[0] for e in ("sf_particles_add", "sf_particles_add_noalloc"))
        assert np.array_equal(a.field("nd"), b.field("nd"))
    finally:
        for s in sims.values():
            s.close()


@pytest.mark.gpu
def test_noalloc_add_refuses_more_than_its_staging_array():
    from starfish_amd import _lib
    from starfish_amd._lib import StarfishError
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    sim = Simulation(deck, capacity=2 * _lib.SF_NOALLOC_MAX_ADD + 10)
    try:
        n = _lib.SF_NOALLOC_MAX_ADD + 1
        z = np.zeros(n)
        with pytest.raises(StarfishError) as ex:
            sim.ctx.call("sf_particles_add_noalloc", sim.materials[0].index, n, _dp(z), _dp(z), _dp(z), _dp(z),
                         _dp(z), None, 1)
        assert ex.value.code == -1 and "staging" in str(ex.value)
        assert sim.materials[0].getNp() == 0
    finally:
        sim.close()
