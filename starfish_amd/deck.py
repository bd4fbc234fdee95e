"""Starfish input decks for the hot path: XML parsing (the subset the path reads) and the
benchmark configurations, as plain data.

The XML surface follows the reference: a key is read from an attribute or from a child
element of the same name (core/io/InputParser.java:80-139), `<load>` splices another file
(InputParser.java:45-78), and the blocks are <domain>/<mesh>/<mesh-bc>
(core/domain/DomainModule.java:141-240), <boundaries>/<boundary>/<path>
(core/boundaries/BoundaryModule.java, core/boundaries/Spline.java:126-285), <materials>
(core/materials/MaterialsModule.java:77-95, Material.java:62-80), <sources>/<boundary_source>
(core/source/SourceModule.java), <solver> (core/solver/SolverModule.java:71-86,
pic/PoissonSolver.java:49-101) and <time> (core/common/TimeModule.java).
"""
import math
import os
import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

QE = 1.60217646e-19
AMU = 1.66053886e-27
K_B = 1.3806503e-23

FACES = ("RIGHT", "TOP", "LEFT", "BOTTOM")     # Mesh.Face order


@dataclass
class MeshSpec:
    domain: str = "XY"
    ni: int = 2
    nj: int = 2
    x0: Tuple[float, float] = (0.0, 0.0)
    dh: Tuple[float, float] = (1.0, 1.0)
    mesh_bc: dict = field(default_factory=dict)   # face -> (type name, value)

    def face_arrays(self):
        """face types / values in the C-ABI layout RIGHT[nj], TOP[ni], LEFT[nj], BOTTOM[ni]"""
        from ._lib import FACE
        types, values = [], []
        for f in FACES:
            n = self.nj if f in ("RIGHT", "LEFT") else self.ni
            t, v = self.mesh_bc.get(f, ("OPEN", 0.0))
            types += [FACE[t.upper()]] * n
            values += [float(v)] * n
        return types, values


@dataclass
class BoundarySpec:
    name: str
    type: str = "DIRICHLET"          # Boundary.BoundaryType (SOLID maps to DIRICHLET)
    value: float = 0.0
    points: List[Tuple[float, float]] = field(default_factory=list)
    reverse: bool = False
    material: Optional[str] = None


@dataclass
class MaterialSpec:
    name: str
    kind: str = "kinetic"            # kinetic, kinetic_hip, solid, boltzmann
    molwt: float = 1.0
    charge: float = 0.0
    spwt: float = 1.0
    nd_back: float = 0.0
    init_T: float = 0.0              # init T=... (Material init_vals.T): computeFields' empty-node value


@dataclass
class SourceSpec:
    name: str
    kind: str = "maxwellian"
    material: str = ""
    boundary: str = ""
    mdot: float = 0.0
    v_drift: float = 0.0
    temperature: float = 0.0
    v_az: float = 0.0                # MaxwellianSource v_az (MaxwellianSource.java:62, 117)
    start_it: int = 0                # Source start_it / stop_it (end_it) (Source.java:68-71)
    stop_it: int = -1


@dataclass
class SolverSpec:
    type: str = "poisson"
    method: str = "GS"
    linear: bool = False
    n0: float = 1e15
    Te0: float = 1.0
    phi0: float = 0.0
    eps_r: float = 1.0
    max_it: int = 5000
    tol: float = 1e-6
    nl_max_it: int = 50
    nl_tol: float = 1e-4
    rel_tol: float = 1e-8
    learned_guess: bool = False
    nl_tol_set: bool = False         # nl_tol given in the deck (the GPU solve then also stops on it)
    skip: int = 1                    # PoissonSolver skip: solve when (it - 1) % skip == 0
    initial_only: bool = False       # SolverModule initial_only: solve only on the first call


@dataclass
class Deck:
    mesh: MeshSpec
    boundaries: List[BoundarySpec] = field(default_factory=list)
    materials: List[MaterialSpec] = field(default_factory=list)
    sources: List[SourceSpec] = field(default_factory=list)
    solver: Optional[SolverSpec] = None
    num_it: int = 0
    dt: float = 1e-7
    seed: int = 0

    def boundary(self, name):
        for b in self.boundaries:
            if b.name.lower() == name.lower():
                return b
        raise KeyError("boundary %s not found" % name)

    def material(self, name):
        for m in self.materials:
            if m.name.lower() == name.lower():
                return m
        raise KeyError("material %s not found" % name)


# ------------------------------------------------------------------------------------------
# XML
# ------------------------------------------------------------------------------------------
def _get(el, key, default=None):
    """InputParser lookup: attribute first, then a child element (case-insensitive)."""
    for k, v in el.attrib.items():
        if k.lower() == key.lower():
            return v.strip()
    for ch in el:
        if ch.tag.lower() == key.lower():
            return (ch.text or "").strip()
    if default is None:
        raise KeyError("missing <%s> in <%s>" % (key, el.tag))
    return default


def _getf(el, key, default=None):
    v = _get(el, key, None if default is None else str(default))
    return float(v)


def _splice(el, base_dir):
    """expands <load>file</load> children in place (InputParser.java:45-78)"""
    out = []
    for ch in list(el):
        if ch.tag.lower() == "load":
            sub = ET.parse(os.path.join(base_dir, ch.text.strip())).getroot()
            out.append(_splice(sub, base_dir))
        else:
            out.append(_splice(ch, base_dir))
    el[:] = out
    return el


def parse_path(path):
    """Linear SVG-like spline path "M x,y L x,y x,y ..." (Spline.setPath); returns points."""
    toks = [t for t in re.split(r"[,\s]+", path.strip()) if t]
    pts, mode, k = [], None, 0
    while k < len(toks):
        t = toks[k]
        if t in ("M", "m", "L", "l"):
            mode = t.upper()
            k += 1
            continue
        if t in ("Z", "z"):
            k += 1
            continue
        if t.upper() in ("C", "S"):
            raise ValueError("cubic spline segments are not supported by this path")
        x, y = float(toks[k]), float(toks[k + 1])
        k += 2
        pts.append((x, y))
    if len(pts) < 2:
        raise ValueError("path needs at least two points: %r" % path)
    return pts


def parse_deck(path):
    """Parses a Starfish simulation file (with its <load>ed files) into a Deck."""
    base = os.path.dirname(os.path.abspath(path))
    root = _splice(ET.parse(path).getroot(), base)
    mesh = None
    bnds, mats, srcs = [], [], []
    solver = None
    num_it, dt, seed = 0, 1e-7, 0
    parent = {c: p for p in root.iter() for c in p}
    for el in root.iter():
        tag = el.tag.lower()
        if tag == "material" and parent.get(el) is not None and parent[el].tag.lower() != "materials":
            continue                 # <material> inside <boundary> names the boundary's material
        if tag == "boundary" and parent.get(el) is not None and parent[el].tag.lower() != "boundaries":
            continue                 # <boundary> inside a source names the source's boundary
        if tag == "domain":
            dom = el.attrib.get("type", "xy").upper()
            for m in el:
                if m.tag.lower() != "mesh":
                    continue
                if _get(m, "type").upper() != "UNIFORM":
                    raise ValueError("only uniform meshes are supported")
                nodes = [int(float(v)) for v in re.split(r"[,\s]+", _get(m, "nodes")) if v]
                origin = [float(v) for v in re.split(r"[,\s]+", _get(m, "origin")) if v]
                spacing = [float(v) for v in re.split(r"[,\s]+", _get(m, "spacing")) if v]
                bc = {}
                for b in m:
                    if b.tag.lower() == "mesh-bc":
                        bc[_get(b, "wall").upper()] = (_get(b, "type").upper(), float(_get(b, "value", "0")))
                mesh = MeshSpec(dom, nodes[0], nodes[1], tuple(origin), tuple(spacing), bc)
        elif tag == "boundary":
            tname = _get(el, "type", "solid").upper()
            btype = {"SOLID": "DIRICHLET", "SYMMETRY": "VIRTUAL"}.get(tname, tname)
            mat = _get(el, "material", "") or None
            rev = _get(el, "reverse", "false").lower() == "true"
            bnds.append(BoundarySpec(_get(el, "name"), btype, float(_get(el, "value", "0")),
                                     parse_path(_get(el, "path")), rev, mat))
        elif tag == "material":
            kind = _get(el, "type").lower()
            nd_back, init_T = 0.0, 0.0
            init = _get(el, "init", "")
            for item in init.split(","):              # Material.setInitValues (Material.java:580-608)
                if "=" in item:
                    k, v = item.split("=")
                    if k.strip().lower() == "nd_back":
                        nd_back = float(v)
                    elif k.strip().lower() == "t":
                        init_T = float(v)
            mats.append(MaterialSpec(_get(el, "name"), kind, float(_get(el, "molwt", "1")),
                                     float(_get(el, "charge", "0")), float(_get(el, "spwt", "1")), nd_back,
                                     init_T))
        elif tag == "boundary_source":
            if _get(el, "circuit_model", "false").lower() == "true":
                raise ValueError("source %s: <circuit_model> is not supported by this path" % _get(el, "name"))
            start_it = int(float(_get(el, "start_it", "0")))
            stop_it = int(float(_get(el, "stop_it", "-1")))
            stop_it = int(float(_get(el, "end_it", str(stop_it))))
            srcs.append(SourceSpec(_get(el, "name"), _get(el, "type").lower(), _get(el, "material"),
                                   _get(el, "boundary"), float(_get(el, "mdot", "0")),
                                   float(_get(el, "v_drift", "0")), float(_get(el, "temperature", "0")),
                                   float(_get(el, "v_az", "0")), start_it, stop_it))
            srcs[-1].current = float(_get(el, "current", "0"))
            srcs[-1].den = float(_get(el, "den", "0"))
        elif tag == "solver":
            s = SolverSpec(type=_get(el, "type").lower())
            s.method = _get(el, "method", "GS").upper()
            s.linear = _get(el, "linear", "false").lower() == "true"
            s.n0 = float(_get(el, "n0", "-1" if not s.linear else "1e15"))
            s.Te0 = float(_get(el, "Te0", "1"))
            s.phi0 = float(_get(el, "phi0", "0"))
            s.eps_r = float(_get(el, "eps_r", "1"))
            s.max_it = int(float(_get(el, "max_it", "5000")))
            s.tol = float(_get(el, "tol", "1e-6"))
            s.nl_max_it = int(float(_get(el, "nl_max_it", "50")))
            s.nl_tol_set = _get(el, "nl_tol", "") != ""
            s.nl_tol = float(_get(el, "nl_tol", "1e-4"))
            s.rel_tol = float(_get(el, "rel_tol", "1e-8"))
            s.learned_guess = _get(el, "learned_guess", "false").lower() == "true"
            s.skip = int(float(_get(el, "skip", "1")))                  # PoissonSolver.java:101
            s.initial_only = _get(el, "initial_only", "false").lower() == "true"   # SolverModule.java:72
            if s.skip < 1:
                raise ValueError("<skip> must be at least 1")
            # keys whose reference behaviour this path does not provide are refused, not dropped
            if _get(el, "qn_switch", "false").lower() == "true":
                raise ValueError("<qn_switch> (PoissonSolver.java:162-193) is not supported by this path")
            if _get(el, "n0_pos", "") != "":
                raise ValueError("<n0_pos> (sampled Boltzmann reference density) is not supported by this path")
            if int(float(_get(el, "rho_moat", "0"))) != 0:
                raise ValueError("<rho_moat> is not supported by this path")
            solver = s
        elif tag == "time":
            num_it = int(float(_get(el, "num_it", "0")))
            dt = float(_get(el, "dt", "1e-7"))
        elif tag == "starfish":
            if el.attrib.get("randomize", "true").lower() == "false":
                seed = 0
    if mesh is None:
        raise ValueError("deck has no <domain>/<mesh>")
    deck = Deck(mesh, bnds, mats, srcs, solver, num_it, dt, seed)
    for src in srcs:
        resolve_source_rate(deck, src, getattr(src, "current", 0.0), getattr(src, "den", 0.0))
    return deck


def boundary_area(deck, b):
    """Spline area of a linear boundary (Spline.computeGeometry): its length in XY (depth 1); the
    surface of revolution (conical frustra about the axis) in RZ and ZR."""
    area = 0.0
    for (x1, y1), (x2, y2) in zip(b.points[:-1], b.points[1:]):
        length = math.hypot(x2 - x1, y2 - y1)
        dom = deck.mesh.domain.upper()
        if dom == "XY":
            area += length
        else:
            r1, r2 = (x1, x2) if dom == "RZ" else (y1, y2)
            area += math.pi * (r1 + r2) * length
    return area


def resolve_source_rate(deck, src, current=0.0, den=0.0):
    """MaxwellianSource's rate keys (MaxwellianSource.java:46-75): exactly one of <mdot>, <current>
    and <den>; current gives mdot = |current / charge * mass|, den gives
    mdot = den A (v_drift + v_mean/4) mass with v_mean = sqrt(8 K T / (pi mass))."""
    n = sum(1 for v in (src.mdot, current, den) if v > 0)
    if n != 1:
        raise ValueError("source %s: only one of <mdot>, <current>, and <den> can be specified" % src.name)
    mat = deck.material(src.material)
    mass = mat.molwt * AMU
    if current != 0.0:
        if mat.charge == 0.0:
            raise ValueError("source %s: cannot use <current> with neutral materials" % src.name)
        src.mdot = abs(current / (mat.charge * QE) * mass)
    if den > 0:
        v_mean = math.sqrt(8 * K_B * src.temperature / (math.pi * mass))
        src.mdot = den * boundary_area(deck, deck.boundary(src.boundary)) * (src.v_drift + v_mean / 4) * mass


# ------------------------------------------------------------------------------------------
# benchmark configurations
# ------------------------------------------------------------------------------------------
def circle_path(r, n, cx=0.0, cy=0.0):
    """closed polyline from (r,0) clockwise, as the tutorial cylinder path
    (dat/examples/tutorial/step1/boundaries.xml: 21 points, 20 segments)"""
    return [(cx + r * math.cos(-2 * math.pi * k / n), cy + r * math.sin(-2 * math.pi * k / n)) for k in range(n + 1)]


def ion_source_mdot(flux, length, molwt):
    """mdot (kg/s) of a boundary source with number flux `flux` (m^-2 s^-1) over `length` (XY, depth 1)"""
    return flux * length * molwt * AMU


def config_c1(seed=1, steps=200):
    """Flow of O+ past a charged cylinder: 81x41 XY nodes, dx = dy = 0.01 m, origin (-0.15, 0);
    left wall Dirichlet 0 V, bottom symmetry; cylinder r = 0.05 m at -100 V; O+ drifting
    Maxwellian (7000 m/s, 1000 K) through the left edge at a number flux of 1e20 m^-2 s^-1;
    Boltzmann electrons n0 = 1e10 m^-3, kTe = 2 eV, phi0 = 0; dt = 1e-7 s."""
    mesh = MeshSpec("XY", 81, 41, (-0.15, 0.0), (0.01, 0.01),
                    {"LEFT": ("DIRICHLET", 0.0), "BOTTOM": ("SYMMETRY", 0.0)})
    L = 0.4
    mdot = ion_source_mdot(1e20, L, 16.0)
    spwt = 8e9
    deck = Deck(mesh,
                boundaries=[BoundarySpec("cylinder", "DIRICHLET", -100.0, circle_path(0.05, 20), False, "SS"),
                            BoundarySpec("inlet", "VIRTUAL", 0.0, [(-0.15, L), (-0.15, 0.0)], False, None)],
                materials=[MaterialSpec("O+", "kinetic_hip", 16.0, 1.0, spwt, 0.0),
                           MaterialSpec("SS", "solid", 52.3, 0.0, 1.0, 0.0)],
                sources=[SourceSpec("ions", "maxwellian", "O+", "inlet", mdot, 7000.0, 1000.0)],
                solver=SolverSpec("poisson_hip", "GS", False, 1e10, 2.0, 0.0),
                num_it=steps, dt=1e-7, seed=seed)
    return deck


def config_c2(steps=500, particles=20_000_000, radius=0.05, bias=-100.0, ni=1024, nj=512):
    """Same physics at 1024x512 nodes, dx = 1e-3 m: 2e7 O+ preloaded uniformly over the open cells
    with the drifting Maxwellian plus left-edge injection.  radius / bias vary the cylinder and
    ni, nj the resolution of the same 1.023 m x 0.511 m domain."""
    Lx, Ly = 1023e-3, 511e-3
    mesh = MeshSpec("XY", ni, nj, (-0.15, 0.0), (round(Lx / (ni - 1), 12), round(Ly / (nj - 1), 12)),
                    {"LEFT": ("DIRICHLET", 0.0), "BOTTOM": ("SYMMETRY", 0.0)})
    n_ion = 1e20 / 7000.0                          # density of the inflowing beam
    area = Lx * Ly - 0.5 * math.pi * radius ** 2
    spwt = n_ion * area / particles
    mdot = ion_source_mdot(1e20, Ly, 16.0)
    deck = Deck(mesh,
                boundaries=[BoundarySpec("cylinder", "DIRICHLET", bias, circle_path(radius, 128), False, "SS"),
                            BoundarySpec("inlet", "VIRTUAL", 0.0, [(-0.15, Ly), (-0.15, 0.0)], False, None)],
                materials=[MaterialSpec("O+", "kinetic_hip", 16.0, 1.0, spwt, 0.0),
                           MaterialSpec("SS", "solid", 52.3, 0.0, 1.0, 0.0)],
                sources=[SourceSpec("ions", "maxwellian", "O+", "inlet", mdot, 7000.0, 1000.0)],
                solver=SolverSpec("poisson_hip", "GS", False, 1e10, 2.0, 0.0),
                num_it=steps, dt=1e-7, seed=1)
    deck.preload = (particles, (7000.0, 0.0, 0.0), 1000.0)
    return deck


def config_c3(particles=100_000_000):
    """Axisymmetric RZ mesh 2048x1024 nodes at 1e-4 m (r along i from the axis): Xe+ (131 amu)
    between two ring electrodes held at 0 V and -400 V (about 20 V/cm over the 0.2 m gap);
    axis symmetry, outer radius open."""
    ni, nj = 2048, 1024
    dr = dz = 1e-4
    mesh = MeshSpec("RZ", ni, nj, (0.0, 0.0), (dr, dz),
                    {"LEFT": ("SYMMETRY", 0.0), "BOTTOM": ("DIRICHLET", 0.0), "TOP": ("DIRICHLET", -200.0)})
    R, Z = (ni - 1) * dr, (nj - 1) * dz
    spwt = 1e12 * math.pi * R * R * Z / particles
    deck = Deck(mesh,
                boundaries=[BoundarySpec("electrode", "DIRICHLET", -100.0,
                                         [(0.08, 0.04), (0.12, 0.04), (0.12, 0.06), (0.08, 0.06), (0.08, 0.04)],
                                         True, "SS")],
                materials=[MaterialSpec("Xe+", "kinetic_hip", 131.0, 1.0, spwt, 0.0)],
                sources=[],
                solver=SolverSpec("poisson_hip", "PCG", False, 1e12, 2.0, 0.0),
                num_it=10, dt=1e-8, seed=1)
    deck.preload = (particles, (0.0, 0.0, 0.0), 300.0)
    return deck
