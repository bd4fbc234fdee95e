"""Restart files (core/common/RestartModule.java:120-165, KineticMaterial.saveRestartData /
loadRestartData, core/materials/KineticMaterial.java:904-1000): the bytes are parsed here with an
independent big-endian reader and compared with the device state, and a fresh simulation loaded
from the file holds the same particles and sample count."""
import struct

import numpy as np
import pytest

from starfish_amd import deck as D


def _read_block(data, off, ni, nj):
    num_samples, n = struct.unpack_from(">iq", data, off)
    off += 12
    rec = np.frombuffer(data, dtype=np.dtype(">f8, >f8, >f8, >f8, >f8, >f8, >f8, >f8, >f8, >f8, >f8, >i4, >i4"),
                        count=n, offset=off)
    off += n * (11 * 8 + 8)
    fields = np.frombuffer(data, dtype=">f8", count=9 * ni * nj, offset=off).reshape(9, ni, nj)
    off += 9 * ni * nj * 8
    return num_samples, rec, fields, off


@pytest.mark.gpu
def test_restart_round_trip(tmp_path):
    from starfish_amd.plugin import Simulation
    deck = D.config_c1(steps=20)
    m = deck.mesh
    path = str(tmp_path / "restart.bin")
    sim = Simulation(deck, capacity=100000, track_ids=True)
    try:
        sim.initial_solve()
        sim.run(15)
        mat = sim.materials[0]
        sim.save_restart(path)
        p = mat.particles(ids=True)
        ref_fields = [mat.field(k) for k in ("nd", "nd-ave", "t", "u", "v", "w", "u-ave", "v-ave", "w-ave")]
        data = open(path, "rb").read()
        it, nmat = struct.unpack_from(">ii", data, 0)
        assert it == sim.it and nmat == len(deck.materials)
        ns, rec, fields, off = _read_block(data, 8, m.ni, m.nj)
        assert off == len(data)
        assert len(rec) == len(p["x"]) > 1000
        assert ns == 15
        cols = [rec["f%d" % k].astype(np.float64) for k in range(11)]
        for k, name in ((0, "x"), (1, "u"), (2, "y"), (3, "v"), (5, "w"), (6, "lc0"), (7, "lc1")):
            assert np.array_equal(cols[k], p[name]), name
        assert not cols[4].any() and not cols[8].any()
        assert np.all(cols[9] == deck.materials[0].spwt)
        assert np.allclose(cols[10], 16.0 * D.AMU, rtol=1e-15)
        assert np.array_equal(rec["f12"].astype(np.int32), p["id"])
        for q in range(9):
            assert np.array_equal(fields[q], ref_fields[q].reshape(m.nj, m.ni).T), q
    finally:
        sim.close()
    sim2 = Simulation(deck, capacity=100000, track_ids=True)
    try:
        it = sim2.load_restart(path)
        assert it == 16 and sim2.it == 16
        q = sim2.materials[0].particles(ids=True)
        order_a = np.lexsort((p["u"], p["y"], p["x"]))
        order_b = np.lexsort((q["u"], q["y"], q["x"]))
        assert len(q["x"]) == len(p["x"])
        for name in ("x", "y", "u", "v", "w"):
            assert np.allclose(q[name][order_b], p[name][order_a], rtol=1e-12, atol=1e-12), name
        # addParticle gives loaded particles new ids (KineticMaterial.java:797)
        assert np.array_equal(np.sort(q["id"]), np.arange(len(q["id"])))
    finally:
        sim2.close()
