"""Pins of the oracle's MCC restatement (collisions/MCC.java:167-297): the counter-based
generator against the published Philox4x32-10 known-answer vectors, and the collision pass
against closed forms (collision probability 1 - exp(-sigma g dt n), the target's Maxwellian
moments with the 1.2 factor of Utils.SampleMaxw3D, the energy source, elastic kinematics)."""
import math

import numpy as np
import pytest

from helpers import geometry, oracle_for
from starfish_amd import deck as D

QE, AMU, K_B = 1.60217646e-19, 1.66053886e-27, 1.3806503e-23


def test_philox_known_answer_vectors():
    import oracle as O
    # Random123 kat_vectors, philox4x32 with 10 rounds: counter, key -> output
    kat = [([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
           ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
           ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
            [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1])]
    for ctr, key, out in kat:
        assert O.philox4x32_10(ctr, key) == out


def _box(n, v0, seed=0):
    mesh = D.MeshSpec("XY", 41, 21, (0.0, 0.0), (0.01, 0.01), {})
    deck = D.Deck(mesh, materials=[D.MaterialSpec("Xe+", "kinetic_hip", 131.0, 1.0, 1e6, 0.0)], solver=None,
                  num_it=1, dt=1e-6, seed=0)
    geo, _ = geometry(deck)
    o = oracle_for(deck, geo)
    rng = np.random.default_rng(seed)
    lc = np.stack([rng.uniform(1, 39, n), rng.uniform(1, 19, n)], 1)
    pos = np.concatenate([lc * 0.01, np.zeros((n, 1))], 1)
    vel = np.tile(np.asarray(v0, float), (n, 1))
    o.load_particles(0, pos, vel, ids=np.arange(n, dtype=np.int32))
    return deck, o


def _target(nn, den, T=0.0, u=0.0):
    return {"den": np.full(nn, den), "u": np.full(nn, u), "v": np.zeros(nn), "w": np.zeros(nn),
            "T": np.full(nn, T)}


def _acc(nn):
    return {"count_sum": np.zeros(nn), "real_sum": np.zeros(nn), "S": np.zeros(nn)}


def test_cex_probability_energy_source_and_counts():
    """Cold target at rest: every particle has g = v0, so the collided fraction is binomial
    around P = 1 - exp(-sigma v0 dt n); CEX leaves collided particles at rest; the energy source
    sums to ncol mpw (0.5 m v0) / (dt dx dy) (E = 0.5 m |v| as the reference writes it)."""
    n, v0, sig, den = 20000, 2000.0, 5e-19, 1e20
    deck, o = _box(n, [v0, 0, 0])
    nn = 41 * 21
    acc = _acc(nn)
    ncol = o.mcc_perform(0, _target(nn, den), 131 * AMU, sig, 0, -1.0, 12345, 0, 1, 1, acc)
    P = 1 - math.exp(-sig * v0 * deck.dt * den)
    assert abs(ncol / n - P) < 5 * math.sqrt(P * (1 - P) / n)
    vel = o.particles(0)["vel"]
    hit = np.all(vel == 0, axis=1)
    assert hit.sum() == ncol and np.all(vel[~hit] == [v0, 0, 0])
    assert acc["count_sum"].sum() == ncol and np.isclose(acc["real_sum"].sum(), ncol * 1e6)
    m = 131 * AMU
    assert np.isclose(acc["S"].sum(), ncol * 1e6 * 0.5 * m * v0 / (deck.dt * 0.01 * 0.01), rtol=1e-12)


def test_cex_target_maxwellian_moments():
    """Warm target streaming at u: collided velocities have mean (u, 0, 0) and per-component
    variance (1.2 vth)^2 / 2 with vth = sqrt(2 k T / m) (Irwin-Hall sum of 12 uniforms, variance 1)."""
    n, T, u = 40000, 300.0, 150.0
    deck, o = _box(n, [3000.0, 0, 0], seed=1)
    nn = 41 * 21
    m = 131 * AMU
    ncol = o.mcc_perform(0, _target(nn, 1e22, T=T, u=u), m, 5e-19, 0, -1.0, 7, 0, 3, 1, _acc(nn))
    vel = o.particles(0)["vel"]
    hit = vel[:, 0] != 3000.0
    assert hit.sum() == ncol and ncol > 10000
    vth = math.sqrt(2 * K_B * T / m)
    var = (1.2 * vth) ** 2 * 0.5
    for c, mean in ((0, u), (1, 0.0), (2, 0.0)):
        x = vel[hit, c]
        assert abs(x.mean() - mean) < 5 * math.sqrt(var / ncol)
        assert abs(x.var() / var - 1) < 0.05


def test_mex_elastic_kinematics():
    """Elastic (MEX) collision with a cold target at rest: the source's post-collision velocity
    lies on the sphere of radius rm2 g around the centre-of-mass velocity rm1 v0."""
    n, v0 = 5000, 2500.0
    deck, o = _box(n, [v0, 0, 0], seed=2)
    nn = 41 * 21
    ms, mt = 131 * AMU, 40 * AMU
    ncol = o.mcc_perform(0, _target(nn, 1e22), mt, 5e-19, 1, -1.0, 99, 1, 1, 1, _acc(nn))
    vel = o.particles(0)["vel"]
    hit = np.any(vel != [v0, 0, 0], axis=1)
    assert hit.sum() == ncol and ncol > 1000
    rm1, rm2 = ms / (ms + mt), mt / (ms + mt)
    r = np.linalg.norm(vel[hit] - [rm1 * v0, 0, 0], axis=1)
    assert np.allclose(r, rm2 * v0, rtol=1e-12)
    # isotropic in the centre-of-mass frame: the cosine along x averages to zero
    assert abs(((vel[hit, 0] - rm1 * v0) / (rm2 * v0)).mean()) < 5 / math.sqrt(3 * ncol)


def test_no_collisions_where_the_target_density_is_zero():
    n = 2000
    deck, o = _box(n, [1000.0, 0, 0], seed=3)
    nn = 41 * 21
    tgt = _target(nn, 1e25)
    tgt["den"][:] = 0.0
    assert o.mcc_perform(0, tgt, 131 * AMU, 1e-18, 0, -1.0, 5, 0, 1, 1, _acc(nn)) == 0
