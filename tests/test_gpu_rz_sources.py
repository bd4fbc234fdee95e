"""RZ push + deposit parity (rotation of KineticMaterial.rotateToRZ, Ruyten weights of
Field2D.scatter) and parity of the host Maxwellian source + GPU addParticle rewind."""
import numpy as np
import pytest

from helpers import geometry, oracle_for, rel_err
from starfish_amd import deck as D

TOL = 1e-5


def rz_deck(domain="RZ"):
    """RZ: r along i (axis on the left face); ZR: the same geometry with the axes exchanged."""
    pts = [(0.03, 0.02), (0.04, 0.02), (0.04, 0.025), (0.03, 0.025), (0.03, 0.02)]
    if domain == "RZ":
        mesh = D.MeshSpec("RZ", 129, 97, (0.0, 0.0), (5e-4, 5e-4),
                          {"LEFT": ("SYMMETRY", 0.0), "BOTTOM": ("DIRICHLET", 0.0), "TOP": ("DIRICHLET", -40.0)})
    else:
        mesh = D.MeshSpec("ZR", 97, 129, (0.0, 0.0), (5e-4, 5e-4),
                          {"BOTTOM": ("SYMMETRY", 0.0), "LEFT": ("DIRICHLET", 0.0), "RIGHT": ("DIRICHLET", -40.0)})
        pts = [(y, x) for x, y in pts][::-1]
    return D.Deck(mesh,
                  boundaries=[D.BoundarySpec("electrode", "DIRICHLET", -20.0, pts, True, "SS")],
                  materials=[D.MaterialSpec("Xe+", "kinetic_hip", 131.0, 1.0, 1e5, 0.0)],
                  solver=D.SolverSpec("poisson_hip", "GS", False, 1e12, 2.0, 0.0), num_it=1, dt=2e-8, seed=1)


@pytest.mark.gpu
@pytest.mark.parametrize("domain", ["RZ", "ZR"])
def test_rz_push_deposit_matches_oracle(domain):
    from starfish_amd.plugin import Simulation
    deck = rz_deck(domain)
    geo, st = geometry(deck)
    o = oracle_for(deck, geo, st)
    o.initial_solve()
    m = deck.mesh
    rng = np.random.default_rng(21)
    n = 20000
    nt = geo.node_type.reshape(m.nj, m.ni)
    openc = (nt[:-1, :-1] == -1) & (nt[1:, :-1] == -1) & (nt[:-1, 1:] == -1) & (nt[1:, 1:] == -1)
    cj, ci = np.nonzero(openc)
    k = rng.integers(0, len(ci), n)
    lc0 = ci[k] + rng.random(n).astype(np.float32)
    lc1 = cj[k] + rng.random(n).astype(np.float32)
    r, z = lc0 * m.dh[0], lc1 * m.dh[1]
    v = (rng.normal(size=(n, 3)) * 3000.0).astype(np.float32).astype(np.float64)
    ids = np.arange(n, dtype=np.int32)
    o.load_particles(0, np.stack([r, z, np.zeros(n)], 1), v, ids=ids)
    sim = Simulation(deck, capacity=2 * n, track_ids=True)
    try:
        sim.set_efield(o.field("efi"), o.field("efj"))
        mat = sim.materials[0]
        mat.add_particles(r, z, v[:, 0], v[:, 1], v[:, 2], ids=ids, rewind=False)
        for step in range(2):
            o.move()
            o.update_species_fields()
            mat.update()
            gp = mat.particles(ids=True)
            op = o.particles(0)
            gi, oi = np.argsort(gp["id"]), np.argsort(op["id"])
            print("rz step", step, len(gi), len(oi), flush=True)
            assert np.array_equal(gp["id"][gi], op["id"][oi])
            ext = (m.ni - 1) * m.dh[0]
            assert rel_err(gp["x"][gi], op["pos"][oi][:, 0], ext) < TOL
            assert rel_err(gp["y"][gi], op["pos"][oi][:, 1], ext) < TOL
            vmax = np.abs(op["vel"]).max()
            for c, name in enumerate("uvw"):
                assert rel_err(gp[name][gi], op["vel"][oi][:, c], vmax) < TOL
            for name in ("nd", "u", "v", "w", "count-sum", "ww-sum"):
                assert rel_err(mat.field(name), o.field(name, 0)) < TOL, name
    finally:
        sim.close()


@pytest.mark.gpu
def test_maxwellian_injection_matches_oracle():
    """Same shared generator state, same draws: the host source of the library and the oracle's
    MaxwellianSource produce the same particles; the GPU applies addParticle's half-step rewind."""
    from starfish_amd.plugin import Simulation
    deck = D.config_c1()
    geo, st = geometry(deck)
    o = oracle_for(deck, geo, st)
    s = deck.sources[0]
    o.add_source(0, deck.boundary("inlet").points, False, s.mdot, s.v_drift, s.temperature)
    o.initial_solve()
    sim = Simulation(deck, capacity=10000, track_ids=True)
    try:
        sim.initial_solve()
        # identical E for the rewind: take the oracle's (the GPU solve agrees to ~1e-8 anyway)
        sim.set_efield(o.field("efi"), o.field("efj"))
        import ctypes as C
        import oracle as O
        # both generators continue from the state the geometry's Monte-Carlo volumes left
        sim.ctx.call("sf_ctx_set_rng", st, 1)
        O.set_rng_state(o, st)
        inj = C.c_int64()
        sim.ctx.call("sf_sources_sample", C.byref(inj))
        # the oracle samples inside its step; compare after one push on both sides
        mat = sim.materials[0]
        mat.update()
        o.step(1)
        gp = mat.particles(ids=True)
        op = o.particles(0)
        print("injected", inj.value, "gpu", len(gp["id"]), "oracle", len(op["id"]), flush=True)
        assert inj.value > 400
        assert len(gp["id"]) == len(op["id"])
        gi, oi = np.lexsort((gp["y"], gp["x"])), np.lexsort((op["pos"][:, 1], op["pos"][:, 0]))
        assert rel_err(gp["x"][gi], op["pos"][oi][:, 0], 0.8) < TOL
        assert rel_err(gp["y"][gi], op["pos"][oi][:, 1], 0.8) < TOL
        assert rel_err(gp["u"][gi], op["vel"][oi][:, 0]) < TOL
        assert rel_err(gp["v"][gi], op["vel"][oi][:, 1], np.abs(op["vel"]).max()) < TOL
    finally:
        sim.close()
