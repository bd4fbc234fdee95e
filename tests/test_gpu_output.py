"""VTK particle output and the stats CSV written from device state (VTKWriter.writeParticles,
StatsModule.saveStats): the VTK arrays decode to exactly the downloaded particles, the CSV columns
are the reference's and carry the mover sums, particle counts and the source's added mass."""
import base64
import csv
import ctypes as C
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from starfish_amd import deck as D


def _decode(arr):
    raw = base64.b64decode(arr.text.strip())
    nbytes = struct.unpack("<i", raw[:4])[0]
    assert len(raw) == 4 + nbytes + 4 and raw[-4:] == b"\0" * 4
    return np.frombuffer(raw[4:4 + nbytes], dtype="<f8" if arr.get("type") == "Float64" else "<i4")


@pytest.mark.gpu
def test_vtk_particles_and_stats_csv(tmp_path):
    from starfish_amd.output import StatsWriter, java_format_g, write_vtk_particles
    from starfish_amd.plugin import Simulation
    deck = D.config_c1(steps=4)
    sim = Simulation(deck, capacity=200_000)
    try:
        mat = sim.materials[0]
        sim.initial_solve()
        stats = StatsWriter(sim, str(tmp_path / "stats.csv"))
        stats.start()
        injected = 0
        for _ in range(4):
            injected += sim.sample_sources()
            mat.update()
            sim.update_fields()
            sim.it += 1
            stats.print_stats()
        stats.finish()
        added = C.c_int64()
        sim.ctx.call("sf_source_added", 0, C.byref(added))
        print("injected", injected, "added", added.value, "np", mat.getNp(), flush=True)
        assert added.value == injected > 0

        rows = list(csv.reader(open(tmp_path / "stats.csv")))
        s, b, m = deck.sources[0].name, [x.name for x in deck.boundaries], mat.name
        assert rows[0][:8] == ["it", "time", "mass." + m, "mom.%s.x" % m, "mom.%s.y" % m, "mom.%s.z" % m,
                               "energy." + m, "mp." + m]
        assert rows[0][8:10] == ["source.%s (kg/s)" % s, "source.%s (A/m^2/s)" % s]
        assert rows[0][10:] == ["flux.%s.%s (#/m^2/s)" % (x, m) for x in b]
        assert len(rows) == 5 and all(len(r) == len(rows[0]) for r in rows)
        last = rows[-1]
        assert last[0] == str(sim.it) and last[1] == java_format_g(sim.it * deck.dt)
        st = mat.stats()
        assert last[2:7] == [java_format_g(v, 4) for v in st]
        assert int(last[7]) == mat.getNp()
        total_mass = sum(float(r[8]) for r in rows[1:])
        expect = injected * deck.materials[0].spwt * mat.mass
        assert abs(total_mass - expect) <= 1e-3 * expect   # %.4g per line

        path = write_vtk_particles(sim, mat, str(tmp_path / "parts.vtp"), animation=True)
        assert path.endswith("parts_%06d.vtp" % sim.it)
        root = ET.parse(path).getroot()
        assert root.get("type") == "PolyData" and root.get("byte_order") == "LittleEndian"
        piece = root.find("PolyData/Piece")
        arrays = {a.get("Name"): _decode(a) for a in root.iter("DataArray")}
        p = mat.particles()
        n = len(p["x"])
        assert int(piece.get("NumberOfPoints")) == n > 1000
        pos = arrays["pos"].reshape(n, 3)
        vel = arrays["velocity"].reshape(n, 3)
        assert np.array_equal(pos[:, 0], p["x"]) and np.array_equal(pos[:, 1], p["y"]) and not pos[:, 2].any()
        assert np.array_equal(vel, np.stack([p["u"], p["v"], p["w"]], 1))
        assert not arrays["radius"].any() and len(arrays["radius"]) == n
        sub = write_vtk_particles(sim, mat, str(tmp_path / "sub.vtp"), particle_count=n // 10)
        ns = int(ET.parse(sub).getroot().find("PolyData/Piece").get("NumberOfPoints"))
        assert abs(ns - n / 10) < 5 * np.sqrt(n / 10)
    finally:
        sim.close()
