"""Failure paths of the C-ABI on the GPU, forced from the host so that every kernel still exits
normally (the reference's counterparts: non-finite particles dropped at insertion,
KineticMaterial.java:1356-1361; solver non-convergence warnings, Solver.java:384-385; Log.error
for bad input, LoggerModule.java:170-176; a restart file of another material count,
RestartModule.java:153-169)."""
import ctypes as C

import numpy as np
import pytest

from helpers import geometry, open_cell_particles
from starfish_amd import deck as D


@pytest.mark.gpu
def test_non_finite_particles_are_dropped_at_insertion():
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    geo, _ = geometry(deck)
    sim = Simulation(deck, capacity=1000, track_ids=True)
    try:
        mat = sim.materials[0]
        x, y, v, _, _ = open_cell_particles(deck, geo, 10, seed=1)
        v[2, 0] = np.nan
        v[5, 1] = np.inf
        v[7, 2] = -np.inf
        mat.add_particles(x, y, v[:, 0], v[:, 1], v[:, 2], ids=np.arange(10, dtype=np.int32), rewind=True)
        assert mat.getNp() == 7
        p = mat.particles(ids=True)
        assert sorted(p["id"].tolist()) == [0, 1, 3, 4, 6, 8, 9]
        assert np.all(np.isfinite(p["u"])) and np.all(np.isfinite(p["v"])) and np.all(np.isfinite(p["w"]))
        sim.set_efield(np.zeros(sim.nn), np.zeros(sim.nn))
        assert mat.update() == 0
        assert np.all(np.isfinite(mat.field("nd")))
    finally:
        sim.close()


@pytest.mark.gpu
def test_non_finite_field_is_refused():
    from starfish_amd._lib import StarfishError
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    sim = Simulation(deck, capacity=100)
    try:
        e = np.zeros(sim.nn)
        e[17] = np.nan
        with pytest.raises(StarfishError) as ex:
            sim.set_efield(e, np.zeros(sim.nn))
        assert ex.value.code == -1 and "non-finite" in str(ex.value)
    finally:
        sim.close()


@pytest.mark.gpu
def test_iteration_cap_returns_not_converged_warning():
    from starfish_amd import _lib
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    deck.solver.max_it = 2
    deck.solver.nl_max_it = 1
    sim = Simulation(deck, capacity=100)
    try:
        info = _lib.SolveInfo()
        rc = sim.ctx.call("sf_poisson_solve", None, C.byref(info))
        assert rc == _lib.SF_WARN_NOT_CONVERGED
        assert info.rel_residual > deck.solver.rel_tol
        assert info.newton_iterations <= 1 and info.pcg_iterations <= 2
        # the potential is still finite and the next, uncapped solve converges from it
        assert np.all(np.isfinite(sim.phi()))
    finally:
        sim.close()
    deck = D.config_c1()
    sim = Simulation(deck, capacity=100)
    try:
        info = _lib.SolveInfo()
        assert sim.ctx.call("sf_poisson_solve", None, C.byref(info)) == 0
        assert info.rel_residual <= deck.solver.rel_tol
    finally:
        sim.close()


@pytest.mark.gpu
def test_capacity_overflow_is_an_error_and_leaves_the_species_intact():
    from starfish_amd._lib import StarfishError
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    geo, _ = geometry(deck)
    sim = Simulation(deck, capacity=100, track_ids=True)
    try:
        mat = sim.materials[0]
        x, y, v, _, _ = open_cell_particles(deck, geo, 150, seed=2)
        with pytest.raises(StarfishError) as ex:
            mat.add_particles(x, y, v[:, 0], v[:, 1], v[:, 2], rewind=False)
        assert ex.value.code == -4          # SF_ERR_OOM
        assert mat.getNp() == 0
        mat.add_particles(x[:60], y[:60], v[:60, 0], v[:60, 1], v[:60, 2], rewind=False)
        assert mat.getNp() == 60
        # device injection past the capacity: the same error, nothing added
        cnt = C.c_int64()
        with pytest.raises(StarfishError) as ex:
            for _ in range(200):
                sim.ctx.call("sf_sources_sample_device", C.c_uint64(1), C.byref(cnt))
        assert ex.value.code == -4
        n_before = mat.getNp()
        assert n_before <= 100
        sim.set_efield(np.zeros(sim.nn), np.zeros(sim.nn))
        assert mat.update() == 0
        assert mat.getNp() <= n_before
    finally:
        sim.close()


@pytest.mark.gpu
def test_restart_with_another_material_count_is_refused(tmp_path):
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    sim = Simulation(deck, capacity=1000)
    try:
        path = str(tmp_path / "restart.bin")
        sim.save_restart(path)
    finally:
        sim.close()
    deck2 = D.config_c1()
    deck2.materials.append(D.MaterialSpec("O", "solid", 16.0, 0.0, 1.0, 0.0))
    sim2 = Simulation(deck2, capacity=1000)
    try:
        with pytest.raises(ValueError):
            sim2.load_restart(path)
    finally:
        sim2.close()


@pytest.mark.gpu
def test_duplicate_collision_streams_are_refused():
    from starfish_amd._lib import StarfishError
    from starfish_amd.interactions import MccHip
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    sim = Simulation(deck, capacity=100, track_ids=True)
    try:
        mat = sim.materials[0]
        a = MccHip(sim, mat, "CEX", 1e-19, 16.0, seed=5, tag=3)
        with pytest.raises(StarfishError):
            MccHip(sim, mat, "MEX", 1e-19, 16.0, seed=5, tag=3)
        b = MccHip(sim, mat, "MEX", 1e-19, 16.0, seed=5)      # automatic tag: a free one
        c = MccHip(sim, mat, "MEX", 1e-19, 16.0, seed=6, tag=3)   # another seed: allowed
        for o in (a, b, c):
            o.close()
        d = MccHip(sim, mat, "CEX", 1e-19, 16.0, seed=5, tag=3)   # released on close
        d.close()
    finally:
        sim.close()
