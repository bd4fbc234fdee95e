"""Pins of the oracle's DSMC restatement (collisions/DSMC.java:167-430, SigmaPlus.java:38-78,
Utils.gamma): Bird's Gamma approximation against math.gamma, the VHS scaling of the
cross-section, exact conservation of momentum and energy by elastic collisions, the number of
selected pairs and the remainder of the NTC formula, and the sig_cr_max update."""
import math

import numpy as np
import pytest

from helpers import geometry, oracle_for
from starfish_amd import deck as D

AMU, K_B = 1.66053886e-27, 1.3806503e-23
N2 = {"sigma_type": 1, "diam": (4.17e-10, 4.17e-10), "ref_temp": (273.0, 273.0), "visc_temp_index": (0.74, 0.74),
      "vss_alpha": (1.0, 1.0)}


def _box(n_per_cell, vel_fn, spwt=1e12, ni=9, nj=5, dx=1e-4, seed=0):
    mesh = D.MeshSpec("XY", ni, nj, (0.0, 0.0), (dx, dx), {})
    deck = D.Deck(mesh, materials=[D.MaterialSpec("N2", "kinetic_hip", 28.0, 0.0, spwt, 0.0)], solver=None,
                  num_it=1, dt=1e-8, seed=0)
    geo, _ = geometry(deck)
    o = oracle_for(deck, geo)
    rng = np.random.default_rng(seed)
    pos, vel = [], []
    for j in range(nj - 1):
        for i in range(ni - 1):
            for _ in range(n_per_cell):
                pos.append([(i + rng.uniform(0.01, 0.99)) * dx, (j + rng.uniform(0.01, 0.99)) * dx, 0.0])
                vel.append(vel_fn(rng))
    pos, vel = np.array(pos), np.array(vel)
    o.load_particles(0, pos, vel, ids=np.arange(len(pos), dtype=np.int32))
    ncell = (ni - 1) * (nj - 1)
    return deck, o, ncell


def test_bird_gamma_matches_the_gamma_function():
    """Bird's polynomial Gamma: exact at 1 and 2, within 2e-4 of math.gamma elsewhere (a wrong
    coefficient or recursion step is off by far more)."""
    from oracle import load
    lib = load()
    assert lib.orc_bird_gamma(1.0) == 1.0 and lib.orc_bird_gamma(2.0) == 1.0
    for x in (0.7, 1.25, 1.5, 1.76, 2.5, 3.2, 4.9):
        assert abs(lib.orc_bird_gamma(x) / math.gamma(x) - 1) < 2e-4, x


def _pair_once(v1, v2, pair, sig_cr_max):
    """Two particles in one cell: the only pair, so nsel and acceptance follow the formula."""
    deck, o, ncell = _box(0, None, spwt=4e14)
    dx = 1e-4
    o.load_particles(0, np.array([[0.3 * dx, 0.4 * dx, 0], [0.6 * dx, 0.2 * dx, 0]]), np.array([v1, v2]),
                     ids=np.array([0, 1], dtype=np.int32))
    st = {"sig_cr_max": np.full(ncell, sig_cr_max), "rem": np.zeros(ncell)}
    acc = {"count_sum": np.zeros(9 * 5), "real_sum": np.zeros(9 * 5)}
    n = o.dsmc_perform(0, 0, pair, 11, 0, 1, 1, st, acc)
    return deck, o, st, acc, n


def test_ntc_selection_remainder_and_sig_cr_max_update():
    """nsel_f = (np spwt / V) np dt sig_cr_max / 2 for one species: with 2 particles, spwt 4e14,
    V = dx^2, dt = 1e-8 and sig_cr_max = 3.7e-15 it is 2.96, so 2 pairs are tested and 0.96 stays
    in rem.  sig_cr_max then becomes sigma(cr) cr of the only pair, with Bird's eq. 4.63
    sigma = pi d^2 (2 k T_ref / (m_r cr^2))^(w - 1/2) / Gamma(5/2 - w) (d = the two diam summed
    over 2), checked with math.gamma at 1e-4 and for its cr^(1 - 2w) scaling between two speeds."""
    mr = 14 * AMU
    sigma0 = math.pi * (2 * 4.17e-10) ** 2 / 4
    res = []
    for v1, v2 in (([600.0, -100.0, 50.0], [-300.0, 200.0, 0.0]), ([60.0, -10.0, 5.0], [-30.0, 20.0, 0.0])):
        deck, o, st, acc, n = _pair_once(v1, v2, N2, 3.7e-15)
        V = 1e-4 * 1e-4
        nsel_f = (2 * 4e14 / V) * 2 * 1e-8 * 3.7e-15 / 2
        assert int(nsel_f) == 2 and abs(st["rem"][0] - (nsel_f - 2)) < 1e-12
        cr = np.linalg.norm(np.subtract(v1, v2))
        expect = cr * sigma0 * (2 * K_B * 273.0 / (mr * cr * cr)) ** (0.74 - 0.5) / math.gamma(2.5 - 0.74)
        assert abs(st["sig_cr_max"][0] / expect - 1) < 1e-4
        assert st["sig_cr_max"][1:].max() == 3.7e-15            # cells without pairs keep their value
        res.append((cr, st["sig_cr_max"][0] / cr))
    (c1, s1), (c2, s2) = res
    assert abs(s2 / s1 / (c2 / c1) ** (1 - 2 * 0.74) - 1) < 1e-12


def test_elastic_collisions_conserve_momentum_and_energy():
    """Many collisions among equal-weight particles: total momentum and kinetic energy are
    unchanged to rounding (a sign or index error in ModelElastic breaks both)."""
    deck, o, ncell = _box(40, lambda r: r.normal(size=3) * 400.0 + [300, 0, 0], spwt=5e14, seed=1)
    v0 = o.particles(0)["vel"].copy()
    st = {"sig_cr_max": np.full(ncell, 1e-15), "rem": np.zeros(ncell)}
    acc = {"count_sum": np.zeros(45), "real_sum": np.zeros(45)}
    n = 0
    for it in range(1, 6):
        n += o.dsmc_perform(0, 0, N2, 5, 0, it, 1, st, acc)
    v1 = o.particles(0)["vel"]
    assert n > 200 and np.mean(np.any(v1 != v0, axis=1)) > 0.3
    assert np.allclose(v1.sum(0), v0.sum(0), rtol=0, atol=1e-9 * np.abs(v0).sum())
    assert abs((v1 ** 2).sum() / (v0 ** 2).sum() - 1) < 1e-12
    assert acc["count_sum"].sum() == n and np.isclose(acc["real_sum"].sum(), n * 5e14)


def test_vss_collisions_conserve_energy_and_differ_from_vhs():
    pair = dict(N2, vss_alpha=(1.4, 1.4))
    deck, o, ncell = _box(30, lambda r: r.normal(size=3) * 500.0, spwt=5e14, seed=2)
    v0 = o.particles(0)["vel"].copy()
    st = {"sig_cr_max": np.full(ncell, 1e-15), "rem": np.zeros(ncell)}
    acc = {"count_sum": np.zeros(45), "real_sum": np.zeros(45)}
    n = sum(o.dsmc_perform(0, 0, pair, 9, 0, it, 1, st, acc) for it in range(1, 4))
    v1 = o.particles(0)["vel"]
    assert n > 50
    assert np.allclose(v1.sum(0), v0.sum(0), rtol=0, atol=1e-9 * np.abs(v0).sum())
    assert abs((v1 ** 2).sum() / (v0 ** 2).sum() - 1) < 1e-12
