"""Host-side mirror of the Starfish plugin surface for the GPU hot path.

Starfish selects implementations through registries filled by plugins:
`Plugin.register()` (core/common/Plugin.java:14-21) calls
`MaterialsModule.registerMaterialType(name, MaterialParser)` (MaterialsModule.java:91-95) and
`SolverModule.registerSolver(name, SolverFactory)` (SolverModule.java:109-112); the XML `type`
attribute then picks the implementation.  This module keeps that shape: `HipPlugin.register`
adds the "KINETIC_HIP" material and the "POISSON_HIP" solver to a `Registry`, and
`Simulation` runs Starfish's main loop (core/common/Starfish.java:67-121) over them, calling the
C-ABI of libstarfish_hip.so.  All arithmetic of the hot path runs in that library on the GPU.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import BND, DOMAIN, FIELD, check
from .deck import Deck


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None


class Context:
    """Owns an sf_ctx (one HIP stream on one device)."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.sf_ctx_create(device, C.byref(h))
        if rc != 0:
            raise _lib.StarfishError(rc, "sf_ctx_create failed (is a GPU visible?)")
        self.h = h

    def close(self):
        if self.h:
            self.lib.sf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args, allow_warning=True):
        return check(self.h, getattr(self.lib, name)(self.h, *args), allow_warning)


class Geometry:
    """Node types, node volumes and segment lists of a mesh (Mesh.setBoundaries + initNodes)."""

    def __init__(self, ctx, deck: Deck):
        m = deck.mesh
        self.mesh = m
        md = _lib.MeshDesc(DOMAIN[m.domain.upper()], m.ni, m.nj, (C.c_double * 2)(*m.x0),
                           (C.c_double * 2)(*m.dh))
        self.md = md
        ft, fv = m.face_arrays()
        self.face_type = np.array(ft, dtype=np.int32)
        self.face_value = np.array(fv, dtype=np.float64)
        bt, bv, bm, bn, bp, br = [], [], [], [], [], []
        for b in deck.boundaries:
            bt.append(BND[b.type.upper()])
            bv.append(b.value)
            bm.append(1 if b.material else 0)
            bn.append(len(b.points))
            for p in b.points:
                bp += [p[0], p[1]]
            br.append(1 if b.reverse else 0)
        arr = lambda x, t: np.array(x if x else [0], dtype=t)
        g = C.c_void_p()
        ctx.call("sf_geometry_build", C.byref(md), _ip(self.face_type), _dp(self.face_value), len(bt),
                 _ip(arr(bt, np.int32)), _dp(arr(bv, np.float64)), _ip(arr(bm, np.int32)), _ip(arr(bn, np.int32)),
                 _dp(arr(bp, np.float64)), _ip(arr(br, np.int32)), C.byref(g))
        nseg, nlist = C.c_int(), C.c_int()
        ctx.lib.sf_geometry_sizes(g, C.byref(nseg), C.byref(nlist))
        nn = m.ni * m.nj
        self.node_type = np.zeros(nn, np.int32)
        self.node_bc = np.zeros(nn)
        self.node_vol = np.zeros(nn)
        self.nseg = nseg.value
        self.seg_xy = np.zeros(max(1, 4 * self.nseg))
        self.seg_btype = np.zeros(max(1, self.nseg), np.int32)
        self.seg_absorb = np.zeros(max(1, self.nseg), np.int32)
        self.node_seg_ptr = np.zeros(nn + 1, np.int32)
        self.node_seg_idx = np.zeros(max(1, nlist.value), np.int32)
        ctx.lib.sf_geometry_export(g, _ip(self.node_type), _dp(self.node_bc), _dp(self.node_vol),
                                   _dp(self.seg_xy), _ip(self.seg_btype), _ip(self.seg_absorb),
                                   _ip(self.node_seg_ptr), _ip(self.node_seg_idx))
        self.seg_bnd = np.zeros(max(1, self.nseg), np.int32)     # boundary of each segment (deck order)
        ctx.lib.sf_geometry_segment_boundaries(g, _ip(self.seg_bnd))
        self.nbnd = len(deck.boundaries)
        ctx.lib.sf_geometry_free(g)


class KineticHipMaterial:
    """`<material type="kinetic_hip">`: a KineticMaterial whose particles live on the GPU
    (same keys as `kinetic`: molwt, charge, spwt, init nd_back; KineticMaterial.java:42-62)."""

    def __init__(self, sim, spec, capacity, track_ids=False, vmax=0.0):
        self.sim = sim
        self.name = spec.name
        self.spec = spec
        if not vmax:
            vmax = speed_hint(sim.deck, spec)
        d = _lib.SpeciesDesc(spec.charge, spec.molwt, spec.spwt, spec.nd_back, vmax, int(capacity),
                             1 if track_ids else 0, getattr(spec, "init_T", 0.0))
        idx = C.c_int()
        sim.ctx.call("sf_species_create", C.byref(d), C.byref(idx))
        self.index = idx.value
        self.mass = spec.molwt * _lib_consts()["AMU"]
        self.charge = spec.charge * _lib_consts()["QE"]

    # KineticMaterial.updateFields() (KineticMaterial.java:117-163) + Material.update nd_back;
    # returns SF_WARN_MOMENT_RANGE (2) when a particle overflowed the fixed-point moments
    def update(self):
        return self.sim.ctx.call("sf_push_deposit", self.index)

    def getNp(self):
        n = C.c_int64()
        self.sim.ctx.call("sf_particles_count", self.index, C.byref(n))
        return n.value

    def add_particles(self, x, y, u, v, w, ids=None, rewind=True):
        """KineticMaterial.addParticle (rewind=True) or a raw state load (rewind=False)"""
        n = len(x)
        a = [np.ascontiguousarray(q, dtype=np.float64) for q in (x, y, u, v, w)]
        ida = np.ascontiguousarray(ids, dtype=np.int32) if ids is not None else None
        self.sim.ctx.call("sf_particles_add", self.index, n, *[_dp(q) for q in a], _ip(ida), 1 if rewind else 0)

    def preload(self, n, drift, temp, seed):
        dr = np.array(drift, dtype=np.float64)
        self.sim.ctx.call("sf_particles_preload", self.index, int(n), _dp(dr), float(temp), C.c_uint64(seed))

    def particles(self, ids=False):
        n = self.getNp()
        out = {k: np.zeros(n) for k in ("x", "y", "lc0", "lc1", "u", "v", "w")}
        idarr = np.zeros(n, np.int32) if ids else None
        self.sim.ctx.call("sf_particles_download", self.index, *[_dp(out[k]) for k in
                                                                 ("x", "y", "lc0", "lc1", "u", "v", "w")],
                          _ip(idarr))
        if ids:
            out["id"] = idarr
        return out

    def field(self, name):
        out = np.zeros(self.sim.nn)
        self.sim.ctx.call("sf_species_field", self.index, FIELD[name], _dp(out))
        return out

    def stats(self):
        out = np.zeros(5)
        self.sim.ctx.call("sf_species_stats", self.index, _dp(out))
        return out

    def saveRestartData(self, out):
        """KineticMaterial.saveRestartData: appends this species' restart block to a binary file."""
        n = C.c_int64()
        self.sim.ctx.call("sf_species_restart_size", self.index, C.byref(n))
        buf = (C.c_uint8 * n.value)()
        self.sim.ctx.call("sf_species_restart_write", self.index, buf, n.value)
        out.write(bytes(buf))

    def loadRestartData(self, data, offset):
        """KineticMaterial.loadRestartData from data[offset:]; returns the offset after the block."""
        view = (C.c_uint8 * (len(data) - offset)).from_buffer_copy(data, offset)
        used = C.c_int64()
        self.sim.ctx.call("sf_species_restart_read", self.index, view, len(data) - offset, C.byref(used))
        return offset + used.value

    def clear_samples(self):
        self.sim.ctx.call("sf_species_clear_samples", self.index)


def _lib_consts():
    from . import deck
    return {"AMU": deck.AMU, "QE": deck.QE}


def speed_hint(deck, spec):
    """Speed hint of a species (sf_species_desc.vmax) from the deck: twice the drift plus three
    widened thermal speeds of its sources and preload (Utils.SampleMaxw3D widens by 1.2).  It sets
    the fixed-point scale of the mover sums (momentum and speed); the deposit channels take their
    scales from the particles at every push.  0 keeps the library's 1e5 m/s when the deck names
    no population."""
    import math
    from .deck import AMU, K_B
    mass = spec.molwt * AMU
    v = 0.0
    for s in deck.sources:
        if s.material.lower() == spec.name.lower():
            v = max(v, s.v_drift + 3 * 1.2 * math.sqrt(2 * K_B * s.temperature / mass))
    pre = getattr(deck, "preload", None)
    if pre is not None and deck.materials and deck.materials[0] is spec:
        _, drift, temp = pre
        v = max(v, math.sqrt(sum(d * d for d in drift)) + 3 * 1.2 * math.sqrt(2 * K_B * temp / mass))
    return 2 * v


class PoissonHipSolver:
    """`<solver type="poisson_hip">`: PoissonSolver keys (n0, Te0, phi0, eps_r, linear, max_it,
    nl_max_it, nl_tol, skip; PoissonSolver.java:49-113, SolverModule.java:71-86) plus rel_tol.
    The solve always runs to rel_tol; nl_tol adds the reference's Newton stop only when the deck
    sets it.  SolverModule.updateFields applies initial_only (Simulation.update_fields)."""

    def __init__(self, sim, spec):
        self.sim = sim
        self.spec = spec
        nl_tol = spec.nl_tol if getattr(spec, "nl_tol_set", False) else 0.0
        d = _lib.PoissonDesc(1 if spec.linear else 0, spec.n0, spec.Te0, spec.phi0, spec.eps_r, spec.max_it,
                             spec.rel_tol, spec.nl_max_it, nl_tol)
        sim.ctx.call("sf_poisson_setup", C.byref(d))
        self.skip = max(1, int(getattr(spec, "skip", 1)))
        self.first = True
        self.initial_only = bool(getattr(spec, "initial_only", False))
        self.last = _lib.SolveInfo()
        self.solved = False                         # whether the last update() ran a solve

    def update(self):                               # PoissonSolver.update
        self.sim.ctx.call("sf_rho_update")
        self.solved = False
        if (self.sim.it - 1) % self.skip != 0:     # (Starfish.getIt()-1) % skip (PoissonSolver.java:113)
            return 0
        self.last = _lib.SolveInfo()
        rc = self.sim.ctx.call("sf_poisson_solve", None, C.byref(self.last))
        self.solved = True
        return rc

    def updateGradientField(self):                  # PotentialSolver.updateGradientField
        self.sim.ctx.call("sf_efield_update")


class BoltzmannElectronsHip:
    """`<material type="boltzmann_electrons_hip">`: BoltzmannElectronsMaterial with the BOLTZMANN
    model (core/materials/BoltzmannElectronsMaterial.java:30-60, 179-215).  The density is
    n0 exp((phi - phi0)/kTe0), floored at 1 m^-3, from the device-resident potential; n0/kTe0 <= 0
    take n0, kTe0 and phi0 from the potential solver.  Like the reference it is not part of rho
    (SolverModule.java:157).  The reference recomputes it every step for output; here it is
    computed when read."""

    def __init__(self, sim, spec, capacity=0, track_ids=False, vmax=0.0, n0=-1.0, kTe0=-1.0, phi0=0.0):
        self.sim = sim
        self.name = spec.name
        self.spec = spec
        self.den0, self.kTe0, self.phi0 = n0, kTe0, phi0

    def update(self):
        pass

    def getNp(self):
        return 0

    def field(self, name):
        if name == "nd":
            out = np.zeros(self.sim.nn)
            self.sim.ctx.call("sf_boltzmann_density", self.den0, self.kTe0, self.phi0, _dp(out))
            return out
        if name == "t":
            from .deck import K_B, QE
            kte = self.kTe0 if (self.den0 > 0 and self.kTe0 > 0) else self.sim.deck.solver.Te0
            return np.full(self.sim.nn, kte * (QE / K_B))    # kTe0 * Constants.EVtoK
        raise KeyError(name)


class Registry:
    """Type-name registries of MaterialsModule and SolverModule."""

    def __init__(self):
        self.materials = {}
        self.solvers = {}

    def registerMaterialType(self, name, parser):
        self.materials[name.upper()] = parser

    def registerSolver(self, name, factory):
        self.solvers[name.upper()] = factory


class HipPlugin:
    """Plugin.register(): adds the GPU material and solver types."""

    def register(self, registry: Registry):
        registry.registerMaterialType("KINETIC_HIP", KineticHipMaterial)
        registry.registerMaterialType("BOLTZMANN_ELECTRONS_HIP", BoltzmannElectronsHip)
        registry.registerSolver("POISSON_HIP", PoissonHipSolver)


class Simulation:
    """Starfish.start + MainLoop over the registered GPU types, driven from a Deck."""

    def __init__(self, deck: Deck, device=0, capacity=None, track_ids=False, vmax=0.0, plugins=None,
                 device_sources=False):
        # device_sources: sample boundary injection on the device (sf_sources_sample_device) instead
        # of replaying the reference's java.util.Random stream on the host
        self.deck = deck
        self.device_sources = device_sources
        self.registry = Registry()
        for p in (plugins or [HipPlugin()]):
            p.register(self.registry)
        self.ctx = Context(device)
        m = deck.mesh
        self.nn = m.ni * m.nj
        self.ctx.call("sf_ctx_set_dt", deck.dt)
        self.ctx.call("sf_ctx_set_rng", deck.seed, 0)
        self.geo = Geometry(self.ctx, deck)
        g = self.geo
        self.ctx.call("sf_mesh_setup", C.byref(g.md), _ip(g.face_type), _dp(g.face_value), _ip(g.node_type),
                      _dp(g.node_bc), _dp(g.node_vol), g.nseg, _dp(g.seg_xy), _ip(g.seg_btype), _ip(g.seg_absorb),
                      _ip(g.node_seg_ptr), _ip(g.node_seg_idx))
        # per-boundary impact tallies of every species (StatsModule's flux columns)
        self.ctx.call("sf_mesh_segment_boundaries", g.nseg, _ip(g.seg_bnd), g.nbnd)
        self.materials = []
        for spec in deck.materials:
            kind = spec.kind.upper()
            if kind in self.registry.materials:
                cap = capacity if capacity is not None else 1_000_000
                self.materials.append(self.registry.materials[kind](self, spec, cap, track_ids, vmax))
        self.solver = None
        if deck.solver is not None and deck.solver.type.upper() in self.registry.solvers:
            self.solver = self.registry.solvers[deck.solver.type.upper()](self, deck.solver)
        for s in deck.sources:
            mat = self.material(s.material)
            b = deck.boundary(s.boundary)
            pts = np.array([c for p in b.points for c in p], dtype=np.float64)
            src = C.c_int()
            self.ctx.call("sf_source_maxwellian", mat.index, len(b.points), _dp(pts), 1 if b.reverse else 0,
                          s.mdot, s.v_drift, s.temperature, C.byref(src))
            v_az, start_it, stop_it = getattr(s, "v_az", 0.0), getattr(s, "start_it", 0), getattr(s, "stop_it", -1)
            if v_az != 0.0 or start_it != 0 or stop_it >= 0:
                self.ctx.call("sf_source_options", src.value, v_az, start_it, stop_it)
        self.it = 1
        self.solve_infos = []

    def material(self, name):
        for m in self.materials:
            if m.name.lower() == name.lower():
                return m
        raise KeyError(name)

    def rng_state(self):
        s = C.c_int64()
        self.ctx.call("sf_ctx_get_rng_state", C.byref(s))
        return s.value

    def update_fields(self):                        # SolverModule.updateFields (SolverModule.java:123-139)
        if self.solver is None:
            return
        if getattr(self.solver, "initial_only", False) and not self.solver.first:
            self.ctx.call("sf_rho_update")
            return 0
        rc = self.solver.update()
        if getattr(self.solver, "solved", True):
            self.solve_infos.append(self.solver.last)
        self.solver.updateGradientField()
        self.solver.first = False
        return rc

    def initial_solve(self):
        return self.update_fields()

    def sample_sources(self):                       # SourceModule.sampleSources
        self.ctx.call("sf_ctx_set_it", self.it)
        injected = C.c_int64()
        if self.device_sources:
            self.ctx.call("sf_sources_sample_device", C.c_uint64(self.deck.seed), C.byref(injected))
        else:
            self.ctx.call("sf_sources_sample", C.byref(injected))
        return injected.value

    def step(self):                                 # Starfish.MainLoop body
        self.sample_sources()                       # also sets the step counter of the library
        for m in self.materials:
            m.update()
        self.update_fields()
        self.it += 1

    def run(self, nsteps):
        for _ in range(nsteps):
            self.step()

    def phi(self):
        out = np.zeros(self.nn)
        self.ctx.call("sf_phi_get", _dp(out))
        return out

    def rho(self):
        out = np.zeros(self.nn)
        self.ctx.call("sf_rho_get", _dp(out))
        return out

    def efield(self):
        ei, ej = np.zeros(self.nn), np.zeros(self.nn)
        self.ctx.call("sf_efield_get", _dp(ei), _dp(ej))
        return ei, ej

    def set_efield(self, ei, ej):
        self.ctx.call("sf_efield_set", _dp(np.ascontiguousarray(ei, np.float64)),
                      _dp(np.ascontiguousarray(ej, np.float64)))

    def save_restart(self, path):
        """RestartModule.saveRestartData (core/common/RestartModule.java:120-140): the step, the
        material count, then each material's block (kinetic GPU species; other materials write
        none, as the reference's Material base class)."""
        import struct
        with open(path, "wb") as f:
            f.write(struct.pack(">ii", self.it, len(self.deck.materials)))
            for m in self.materials:
                if hasattr(m, "saveRestartData"):
                    m.saveRestartData(f)

    def load_restart(self, path):
        """RestartModule.loadRestartData (RestartModule.java:145-165): restores the step and the
        kinetic species; refuses a file written for another material count."""
        import struct
        data = open(path, "rb").read()
        it, nmat = struct.unpack(">ii", data[:8])
        if nmat != len(self.deck.materials):
            raise ValueError("incompatible restart file: %d materials, this deck has %d" % (nmat, len(self.deck.materials)))
        off = 8
        for m in self.materials:
            if hasattr(m, "loadRestartData"):
                off = m.loadRestartData(data, off)
        self.it = it
        return it

    def set_bfield(self, bi, bj):
        """Loads the domain's magnetic field (DomainModule "bfi"/"bfj"); None removes it."""
        if bi is None:
            self.ctx.call("sf_bfield_set", None, None)
            return
        self.ctx.call("sf_bfield_set", _dp(np.ascontiguousarray(bi, np.float64)),
                      _dp(np.ascontiguousarray(bj, np.float64)))

    def close(self):
        self.ctx.close()
