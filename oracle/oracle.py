"""ctypes wrapper of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg as the checker / CPU baseline.  The product package never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class MeshDesc(C.Structure):
    _fields_ = [("domain", C.c_int), ("ni", C.c_int), ("nj", C.c_int), ("x0", C.c_double * 2),
                ("dh", C.c_double * 2), ("face_type", IP), ("face_value", DP), ("node_type", IP),
                ("node_bc", DP), ("node_vol", DP), ("nseg", C.c_int), ("seg_xy", DP), ("seg_btype", IP),
                ("seg_absorb", IP), ("node_seg_ptr", IP), ("node_seg_idx", IP), ("seed", C.c_int64),
                ("nthreads", C.c_int), ("dt", C.c_double), ("seg_bnd", IP), ("nbnd", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    sig = {
        "orc_create": (P, [C.POINTER(MeshDesc)]),
        "orc_destroy": (None, [P]),
        "orc_add_species": (C.c_int, [P, C.c_double, C.c_double, C.c_double, C.c_double]),
        "orc_boundary_hits": (C.c_int, [P, C.c_int, C.POINTER(C.c_int64), C.c_int, C.c_int]),
        "orc_set_solver": (None, [P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.c_int, C.c_double]),
        "orc_add_source": (C.c_int, [P, C.c_int, C.c_int, DP, C.c_int, C.c_double, C.c_double, C.c_double]),
        "orc_load_particles": (None, [P, C.c_int, C.c_int64, DP, DP, DP, IP, C.c_int]),
        "orc_num_particles": (C.c_int64, [P, C.c_int]),
        "orc_get_particles": (None, [P, C.c_int, DP, DP, DP, DP, IP]),
        "orc_set_field": (C.c_int, [P, C.c_int, C.c_char_p, DP]),
        "orc_get_field": (C.c_int, [P, C.c_int, C.c_char_p, DP]),
        "orc_move": (None, [P]),
        "orc_update_species_fields": (None, [P]),
        "orc_compute_fields": (None, [P]),
        "orc_update_rho": (None, [P]),
        "orc_solve_potential": (C.c_int, [P]),
        "orc_update_efield": (None, [P]),
        "orc_initial_solve": (None, [P]),
        "orc_step": (None, [P, C.c_int]),
        "orc_source_options": (None, [P, C.c_int, C.c_double, C.c_int, C.c_int]),
        "orc_species_init_T": (None, [P, C.c_int, C.c_double]),
        "orc_set_it": (None, [P, C.c_int]),
        "orc_sample_sources": (None, [P]),
        "orc_get_it": (C.c_int, [P]),
        "orc_get_stats": (None, [P, C.c_int, DP]),
        "orc_get_timers": (None, [P, DP]),
        "orc_last_lin_iterations": (C.c_int, [P]),
        "orc_operator_nnz": (C.c_int64, [P, C.c_int]),
        "orc_operator_csr": (None, [P, C.c_int, C.POINTER(C.c_int64), IP, DP]),
        "orc_fixed_nodes": (C.c_int, [P, IP]),
        "orc_set_rng_state": (None, [P, C.c_int64]),
        "orc_round_state_fp32": (None, [P, C.c_int]),
        "orc_boltzmann_density": (None, [C.c_int64, DP, C.c_double, C.c_double, C.c_double, DP]),
        "orc_bird_gamma": (C.c_double, [C.c_double]),
        "orc_dsmc_perform": (C.c_int64, [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                         C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                         C.c_uint32, C.c_int, C.c_int, DP, DP, DP, DP]),
        "orc_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "orc_mcc_perform": (C.c_int64, [P, C.c_int, DP, DP, DP, DP, DP, C.c_double, C.c_double, C.c_int, C.c_double,
                                        C.c_uint64, C.c_uint32, C.c_int, C.c_int, DP, DP, DP]),
        "orc_javarandom_doubles": (None, [C.c_int64, C.c_int, DP]),
        "orc_javarandom_ints": (None, [C.c_int64, C.c_int, IP]),
        "orc_javarandom_ints_bounded": (None, [C.c_int64, C.c_int, C.c_int, IP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(DP) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(IP) if a is not None else None


class Oracle:
    """One oracle simulation over a mesh geometry (arrays as produced by Starfish's Mesh)."""

    def __init__(self, mesh, face_type, face_value, node_type, node_bc, node_vol, seg_xy, seg_btype, seg_absorb,
                 node_seg_ptr, node_seg_idx, nseg, seed=0, nthreads=1, dt=1e-7, seg_bnd=None, nbnd=0):
        self.lib = load()
        dom = {"XY": 0, "RZ": 1, "ZR": 2}[mesh.domain.upper()]
        self._keep = [np.ascontiguousarray(a) for a in (face_type, face_value, node_type, node_bc, node_vol, seg_xy,
                                                         seg_btype, seg_absorb, node_seg_ptr, node_seg_idx)]
        k = self._keep
        self.ni, self.nj = mesh.ni, mesh.nj
        self.nn = mesh.ni * mesh.nj
        d = MeshDesc(dom, mesh.ni, mesh.nj, (C.c_double * 2)(*mesh.x0), (C.c_double * 2)(*mesh.dh),
                     _ip(k[0].astype(np.int32)), _dp(k[1]), _ip(k[2]), _dp(k[3]), _dp(k[4]), nseg, _dp(k[5]),
                     _ip(k[6]), _ip(k[7]), _ip(k[8]), _ip(k[9]), seed, nthreads, dt)
        if seg_bnd is not None:
            self._seg_bnd = np.ascontiguousarray(seg_bnd, dtype=np.int32)
            d.seg_bnd = _ip(self._seg_bnd)
            d.nbnd = int(nbnd)
        self.nbnd = int(nbnd) if seg_bnd is not None else 0
        self._ft = k[0].astype(np.int32)
        d.face_type = _ip(self._ft)
        self.h = self.lib.orc_create(C.byref(d))

    @classmethod
    def from_geometry(cls, geo, seed=0, nthreads=1, dt=1e-7):
        return cls(geo.mesh, geo.face_type, geo.face_value, geo.node_type, geo.node_bc, geo.node_vol, geo.seg_xy,
                   geo.seg_btype, geo.seg_absorb, geo.node_seg_ptr, geo.node_seg_idx, geo.nseg, seed, nthreads, dt,
                   getattr(geo, "seg_bnd", None), getattr(geo, "nbnd", 0))

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def boundary_hits(self, sp, clear=False):
        """particles of species sp absorbed on each boundary since the last clear"""
        out = np.zeros(max(1, self.nbnd), dtype=np.int64)
        assert self.lib.orc_boundary_hits(self.h, sp, out.ctypes.data_as(C.POINTER(C.c_int64)), self.nbnd,
                                          1 if clear else 0) == 0
        return out[:self.nbnd]

    def add_species(self, charge, molwt, spwt, nd_back=0.0):
        return self.lib.orc_add_species(self.h, charge, molwt, spwt, nd_back)

    def set_solver(self, spec):
        method = 1 if spec.method.upper() == "PCG" else 0
        self.lib.orc_set_solver(self.h, method, 1 if spec.linear else 0, spec.n0, spec.Te0, spec.phi0, spec.eps_r,
                                spec.max_it, spec.tol, spec.nl_max_it, spec.nl_tol)

    def add_source(self, sp, points, reverse, mdot, v_drift, temp):
        pts = np.array([c for p in points for c in p], dtype=np.float64)
        return self.lib.orc_add_source(self.h, sp, len(points), _dp(pts), 1 if reverse else 0, mdot, v_drift, temp)

    def source_options(self, src, v_az=0.0, start_it=0, stop_it=-1):
        self.lib.orc_source_options(self.h, src, v_az, start_it, stop_it)

    def set_init_T(self, sp, T):
        self.lib.orc_species_init_T(self.h, sp, T)

    @property
    def it(self):
        return self.lib.orc_get_it(self.h)

    @it.setter
    def it(self, value):
        self.lib.orc_set_it(self.h, int(value))

    def load_particles(self, sp, pos3, vel3, ids=None, rewind=False, mpw=None):
        pos3 = np.ascontiguousarray(pos3, np.float64)
        vel3 = np.ascontiguousarray(vel3, np.float64)
        ida = np.ascontiguousarray(ids, np.int32) if ids is not None else None
        mp = np.ascontiguousarray(mpw, np.float64) if mpw is not None else None
        self.lib.orc_load_particles(self.h, sp, len(pos3), _dp(pos3), _dp(vel3), _dp(mp), _ip(ida),
                                    1 if rewind else 0)

    def num_particles(self, sp):
        return self.lib.orc_num_particles(self.h, sp)

    def particles(self, sp):
        n = self.num_particles(sp)
        pos, vel, lc = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 2))
        mpw, ids = np.zeros(n), np.zeros(n, np.int32)
        self.lib.orc_get_particles(self.h, sp, _dp(pos), _dp(vel), _dp(lc), _dp(mpw), _ip(ids))
        return {"pos": pos, "vel": vel, "lc": lc, "mpw": mpw, "id": ids}

    def set_field(self, name, data, sp=-1):
        a = np.ascontiguousarray(data, np.float64)
        assert self.lib.orc_set_field(self.h, sp, name.encode(), _dp(a)) == 0

    def field(self, name, sp=-1):
        out = np.zeros(self.nn)
        assert self.lib.orc_get_field(self.h, sp, name.encode(), _dp(out)) == 0
        return out

    def move(self):
        self.lib.orc_move(self.h)

    def dsmc_perform(self, spa, spb, pair, seed, tag, it, frequency, state, acc):
        """One DSMC pass (collisions/DSMC.java:167-430).  pair: dict sigma_type, c0, diam, ref_temp,
        visc_temp_index, vss_alpha (each of the last four a pair of values); state: dict of per-cell
        sig_cr_max and rem (in/out); acc: dict of count_sum and real_sum node fields (accumulated)."""
        d, t, w, a = pair["diam"], pair["ref_temp"], pair["visc_temp_index"], pair["vss_alpha"]
        return self.lib.orc_dsmc_perform(self.h, spa, spb, pair["sigma_type"], pair.get("c0", 0.0), d[0], d[1], t[0], t[1],
                                         w[0], w[1], a[0], a[1], seed, tag, it, frequency, _dp(state["sig_cr_max"]),
                                         _dp(state["rem"]), _dp(acc["count_sum"]), _dp(acc["real_sum"]))

    def mcc_perform(self, sp, target, target_mass, sigma, model, max_T, seed, tag, it, frequency, acc):
        """One MCC pass (collisions/MCC.java:167-221) over species sp.  target: dict of node fields
        den, u, v, w, T; acc: dict of count_sum, real_sum, S arrays, accumulated in place."""
        f = {k: np.ascontiguousarray(target[k], np.float64) for k in ("den", "u", "v", "w", "T")}
        return self.lib.orc_mcc_perform(self.h, sp, _dp(f["den"]), _dp(f["u"]), _dp(f["v"]), _dp(f["w"]),
                                        _dp(f["T"]), target_mass, sigma, model, max_T, seed, tag, it, frequency,
                                        _dp(acc["count_sum"]), _dp(acc["real_sum"]), _dp(acc["S"]))

    def update_species_fields(self):
        self.lib.orc_update_species_fields(self.h)

    def compute_fields(self):
        self.lib.orc_compute_fields(self.h)

    def update_rho(self):
        self.lib.orc_update_rho(self.h)

    def solve_potential(self):
        return self.lib.orc_solve_potential(self.h)

    def update_efield(self):
        self.lib.orc_update_efield(self.h)

    def initial_solve(self):
        self.lib.orc_initial_solve(self.h)

    def sample_sources(self):
        self.lib.orc_sample_sources(self.h)

    def step(self, n=1):
        self.lib.orc_step(self.h, n)

    def round_state_fp32(self, sp):
        """Rounds offsets and velocities to FP32, as the device stores the particle state."""
        self.lib.orc_round_state_fp32(self.h, sp)

    def stats(self, sp):
        out = np.zeros(5)
        self.lib.orc_get_stats(self.h, sp, _dp(out))
        return out

    def timers(self):
        out = np.zeros(4)
        self.lib.orc_get_timers(self.h, _dp(out))
        return out

    def operator(self, which=0):
        """assembled matrix (0 A, 1 Gi, 2 Gj) as a scipy CSR matrix"""
        import scipy.sparse as sp
        nnz = self.lib.orc_operator_nnz(self.h, which)
        rp = np.zeros(self.nn + 1, np.int64)
        col = np.zeros(nnz, np.int32)
        val = np.zeros(nnz)
        self.lib.orc_operator_csr(self.h, which, rp.ctypes.data_as(C.POINTER(C.c_int64)), _ip(col), _dp(val))
        return sp.csr_matrix((val, col, rp), shape=(self.nn, self.nn))

    def fixed_nodes(self):
        out = np.zeros(self.nn, np.int32)
        self.lib.orc_fixed_nodes(self.h, _ip(out))
        return out


def philox4x32_10(ctr, key):
    """Philox4x32-10 block (Salmon et al., SC'11) of a 4-word counter under a 2-word key."""
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    load().orc_philox4x32_10(c, k, o)
    return list(o)


def boltzmann_density(phi, den0, kTe0, phi0):
    """Boltzmann electron density of the reference's BOLTZMANN electron model on every node."""
    phi = np.ascontiguousarray(phi, np.float64)
    out = np.zeros_like(phi)
    load().orc_boltzmann_density(phi.size, _dp(phi), den0, kTe0, phi0, _dp(out))
    return out


def java_doubles(seed, n):
    out = np.zeros(n)
    load().orc_javarandom_doubles(seed, n, _dp(out))
    return out


def java_ints(seed, n, bound=None):
    out = np.zeros(n, np.int32)
    if bound is None:
        load().orc_javarandom_ints(seed, n, _ip(out))
    else:
        load().orc_javarandom_ints_bounded(seed, bound, n, _ip(out))
    return out


def set_rng_state(orc, state):
    load().orc_set_rng_state(orc.h, state)
