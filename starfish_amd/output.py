"""Output writers over the device-resident state: the VTK particle file and the stats CSV.

Host-side restatements of the reference's writers, fed through the C-ABI downloads:
* `write_vtk_particles` follows VTKWriter.writeParticles (core/io/VTKWriter.java:1010-1099) with
  its data-array encoding (outputDataArray, VTKWriter.java:1218-1291) and file naming
  (Writer.splitFileName, core/io/Writer.java:422-448);
* `StatsWriter` follows StatsModule (core/diagnostics/StatsModule.java:36-154).
Numbers are printed the way java.util.Formatter prints them (`java_format_g`).
"""
import base64
import math
import struct
from decimal import Context, Decimal, ROUND_HALF_UP

import ctypes as C
import numpy as np


def java_format_g(x, precision=6):
    """java.util.Formatter's %g (general scientific) conversion of a double.

    The value's shortest round-trip decimal digits are rounded half-up to `precision` significant
    digits (0 counts as 1).  With E the decimal exponent of the rounded value, -4 <= E < precision
    prints in plain notation with precision-1-E digits after the point, otherwise in scientific
    notation d.ddd with precision-1 digits and an exponent of at least two digits (e+05).  Trailing
    zeros are kept, unlike C's %g.  Zero prints as 0.000 (precision-1 zeros), NaN as NaN and the
    infinities as Infinity / -Infinity."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Infinity" if x > 0 else "-Infinity"
    prec = 1 if precision == 0 else precision
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        return sign + ("0." + "0" * (prec - 1) if prec > 1 else "0")
    d = Context(prec=prec, rounding=ROUND_HALF_UP).plus(Decimal(repr(abs(x))))
    e = d.adjusted()
    if -4 <= e < prec:
        return sign + format(d, ".%df" % (prec - 1 - e))
    mant = format(d.scaleb(-e), ".%df" % (prec - 1))
    return sign + "%se%s%02d" % (mant, "-" if e < 0 else "+", abs(e))


def split_file_name(name):
    """Writer.splitFileName: [name up to the last '.', the rest, directory, base name].  Without a
    '.' (or with one only at position 0) the split falls before the last character, as in the
    reference."""
    p1 = len(name) - 1
    while p1 >= 0 and name[p1] not in "/\\":
        p1 -= 1
    p2 = len(name) - 1
    while p2 >= 0 and name[p2] != ".":
        p2 -= 1
    if p2 <= 0:
        p2 = len(name) - 1
    if p1 <= 0:
        p1 = 0
    return [name[:p2], name[p2:], name[:p1], name[(p1 + 1) if p1 > 0 else 0:p2]]


def _data_array(lines, name, data, ncomp, fmt, little_endian=True):
    """VTKWriter.outputDataArray for Float64 (double[]) or Int32 (int[]) data."""
    data = np.asarray(data)
    is_int = data.dtype.kind in "iu"
    vtype = "Int32" if is_int else "Float64"
    nc = 'NumberOfComponents="%d"' % ncomp
    if fmt == "binary":
        e = "<" if little_endian else ">"
        vals = data.astype(e + ("i4" if is_int else "f8"))
        nbytes = vals.nbytes
        # ByteBuffer.allocate(num_bytes + Long.BYTES): the 4-byte length, the values, and 4 bytes
        # that stay zero
        raw = struct.pack(e + "i", nbytes) + vals.tobytes() + b"\0" * 4
        lines.append('<DataArray Name="%s" type="%s" %s format="binary">' % (name, vtype, nc))
        lines.append(base64.b64encode(raw).decode("ascii"))
        lines.append("</DataArray>")
    elif fmt == "ascii":
        lines.append('<DataArray Name="%s" type="%s" %s format="ascii">' % (name, vtype, nc))
        if is_int:
            body = "".join("%d " % v for v in data.ravel())
        else:
            body = "".join(java_format_g(v) + " " for v in data.ravel())
        lines.append(body + "\n</DataArray>")
    else:
        raise ValueError("output format %r: binary and ascii are supported" % fmt)


def write_vtk_particles(sim, material, file_name, particle_count=None, animation=False, output_format="binary",
                        seed=0):
    """VTKWriter.writeParticles for a kinetic material on the device: a PolyData file with the
    points at (x, y, 0), and point data 'velocity' (3 components) and 'radius' (diam/2 of the
    material, 0 unless the material has a diameter).

    With particle_count below the particle count, each particle is kept with probability
    particle_count / np, as in the reference.  The reference draws from Starfish.rnd() in its
    particle order; here the draws come from a generator seeded with `seed`, over the cell-sorted
    device order, so the subset is a different sample of the same size distribution.  The
    reference's `rotate` option needs the azimuthal angle, which this state does not keep, and is
    not offered.  Returns the path written."""
    mat = sim.material(material) if isinstance(material, str) else material
    p = mat.particles()
    n = len(p["x"])
    keep = np.ones(n, bool)
    if particle_count is not None and n > 0:
        prob = float(particle_count) / n
        if prob < 1.0:
            keep = np.random.default_rng(seed).random(n) < prob
    x, y = p["x"][keep], p["y"][keep]
    vel = np.stack([p["u"][keep], p["v"][keep], p["w"][keep]], 1).ravel()
    pos = np.stack([x, y, np.zeros_like(x)], 1).ravel()
    radius = np.full(len(x), 0.5 * float(getattr(mat.spec, "diam", 0.0) or 0.0))
    parts = split_file_name(file_name)
    name = parts[0] + ("_%06d" % sim.it if animation else "") + parts[1]
    endianess = ' byte_order="LittleEndian"' if output_format == "binary" else ""
    lines = ['<?xml version="1.0"?>', '<VTKFile type="PolyData"%s>' % endianess, "<PolyData>",
             '<Piece NumberOfPoints="%d" NumberOfVerts="0" NumberOfLines="0" NumberOfStrips="0" '
             'NumberOfPolys="0">' % len(x), "<Points>"]
    _data_array(lines, "pos", pos, 3, output_format)
    lines += ["</Points>", "<PointData>"]
    _data_array(lines, "velocity", vel, 3, output_format)
    _data_array(lines, "radius", radius, 1, output_format)
    lines += ["</PointData>", "</Piece>", "</PolyData>", "</VTKFile>"]
    with open(name, "w", newline="\n") as f:
        f.write("\n".join(lines) + "\n")
    return name


class StatsWriter:
    """StatsModule: one CSV line per `skip` steps with, per material, the mover sums (mass,
    momentum, energy) and the macroparticle count; per source, the mass it added since the last
    line and the matching current density; per boundary and material, the flux of particles the
    boundary absorbed since the last line over its area and the interval (Material.getFluxInst:
    the device tallies absorbed particles per boundary, sf_species_boundary_hits, and the count
    times spwt is the reference's deposit sum).  Mass, momentum and energy are those of the last
    push (sf_species_stats), the reference's getMassSum / getMomentumSum / getEnergySum."""

    def __init__(self, sim, file_name="starfish_stats.csv", skip=1):
        self.sim = sim
        self.file_name = file_name
        self.skip = skip
        self.f = None
        self.added = [0] * len(sim.deck.sources)

    def start(self):
        """StatsModule.start: opens the file and writes the header."""
        sim = self.sim
        self.f = open(self.file_name, "w", newline="\n")
        h = "it,time"
        for m in sim.materials:
            h += ",mass.%s,mom.%s.x,mom.%s.y,mom.%s.z,energy.%s" % ((m.name,) * 5)
            if _kinetic(m):
                h += ",mp.%s" % m.name
        for s in sim.deck.sources:
            h += ",source.%s (kg/s)" % s.name
            h += ",source.%s (A/m^2/s)" % s.name
        for b in sim.deck.boundaries:
            for m in sim.materials:
                h += ",flux.%s.%s (#/m^2/s)" % (b.name, m.name)
        self.f.write(h + "\n")
        self.added = [self._source_added(k) for k in range(len(sim.deck.sources))]

    def _source_added(self, k):
        c = C.c_int64()
        self.sim.ctx.call("sf_source_added", k, C.byref(c))
        return c.value

    def print_stats(self):
        """StatsModule.printStats: a line every `skip` iterations."""
        if self.skip > 0 and self.sim.it % self.skip != 0:
            return
        self.save_stats()

    def save_stats(self):
        """StatsModule.saveStats."""
        sim = self.sim
        dt = sim.deck.dt
        line = "%d,%s" % (sim.it, java_format_g(sim.it * dt))
        for m in sim.materials:
            st = m.stats() if _kinetic(m) else np.zeros(5)
            line += "".join("," + java_format_g(v, 4) for v in st[:5])
            if _kinetic(m):
                line += ",%d" % m.getNp()
        for k, s in enumerate(sim.deck.sources):
            mat = sim.material(s.material)
            now = self._source_added(k)
            mass_gen = (now - self.added[k]) * mat.spec.spwt * mat.mass
            self.added[k] = now
            area = _boundary_area(sim.deck.boundary(s.boundary), sim.deck.mesh.domain)
            cur = mass_gen * mat.charge / (mat.mass * area * self.skip * dt) if area > 0 else float("nan")
            line += "," + java_format_g(mass_gen, 4)
            line += "," + java_format_g(cur, 4)
        nb = len(sim.deck.boundaries)
        hits = {}
        for m in sim.materials:
            if _kinetic(m):
                h = np.zeros(max(1, nb), dtype=np.int64)
                sim.ctx.call("sf_species_boundary_hits", m.index, nb, h.ctypes.data_as(C.POINTER(C.c_int64)), 1)
                hits[m.name] = h
        t = self.skip * dt
        for b_i, b in enumerate(sim.deck.boundaries):
            area = _boundary_area(b, sim.deck.mesh.domain)
            for m in sim.materials:
                dep = float(hits[m.name][b_i]) * m.spec.spwt if m.name in hits else 0.0
                flux = dep / (area * t) if area > 0 else (float("nan") if dep == 0 else float("inf"))
                line += "," + java_format_g(flux, 4)
        self.f.write(line + "\n")
        self.f.flush()

    def finish(self):
        if self.f is not None:
            self.f.close()
            self.f = None


def _kinetic(m):
    from .plugin import KineticHipMaterial
    return isinstance(m, KineticHipMaterial)


def _boundary_area(b, domain):
    """Spline.area (core/boundaries/Spline.java:576-579) over linear segments
    (LinearSegment.area, LinearSegment.java:54-79): the length in XY (unit depth), the lateral
    area pi (r1 + r2) |ds| of each conical frustum in RZ (r = x) and ZR (r = y)."""
    pts = np.asarray(b.points, dtype=float)
    if len(pts) < 2:
        return 0.0
    p1, p2 = pts[:-1], pts[1:]
    ds = np.hypot(p2[:, 0] - p1[:, 0], p2[:, 1] - p1[:, 1])
    if domain == "XY":
        return float(ds.sum())
    r = 0 if domain == "RZ" else 1
    return float(np.sum(np.pi * np.abs(p1[:, r] + p2[:, r]) * ds))
