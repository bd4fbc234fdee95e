"""Volume interactions on the device: the `<mcc>` block of the reference's interactions module
(collisions/MCC.java:45-124, InteractionsModule.performInteractions, interactions/
InteractionsModule.java:115-130).  The collision pass runs in libstarfish_hip.so
(sf_mcc_perform); this module holds the parameters and the target material's node fields."""
import ctypes as C

import numpy as np

from . import _lib

TAG_AUTO = 0xFFFFFFFF   # SF_TAG_AUTO

MODELS = {"CEX": 0, "MEX": 1, "ELASTIC": 1}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class MccHip:
    """MCC of a kinetic GPU species with a background target given by node fields.

    sim: plugin.Simulation; source: the KineticHipMaterial (created with track_ids=True);
    model: "CEX" or "MEX"/"ELASTIC"; sigma: constant cross-section (m^2); target_molwt: amu;
    frequency, stop_it, max_target_temp: the reference's keys; seed and tag key the per-particle
    Philox stream (tag None: the library picks a tag no other live MCC of this seed uses)."""

    FIELDS = {"col-count-sum": 0, "col-real-sum": 1, "col-count": 2, "nu": 3, "S": 4}

    def __init__(self, sim, source, model, sigma, target_molwt, frequency=1, stop_it=-1, max_target_temp=-1.0,
                 seed=0, tag=None):
        self.sim = sim
        self.lib = _lib.load()
        tag = TAG_AUTO if tag is None else tag
        kind = model.upper()
        if kind not in MODELS:
            raise ValueError("unsupported MCC model %s (CEX and MEX/ELASTIC are provided)" % model)
        d = _lib.MccDesc(source.index, MODELS[kind], sigma, target_molwt, frequency, stop_it, max_target_temp, seed,
                         tag)
        h = C.c_void_p()
        sim.ctx.call("sf_mcc_create", C.byref(d), C.byref(h))
        self.h = h

    def _call(self, name, *args):
        return _lib.check(self.sim.ctx.h, getattr(self.lib, name)(self.h, *args))

    def set_target(self, den, u=None, v=None, w=None, temp=None):
        """Target material's density, stream velocity and temperature on the nodes."""
        arr = [None if a is None else np.ascontiguousarray(np.broadcast_to(a, (self.sim.nn,)), np.float64)
               for a in (den, u, v, w, temp)]
        self._call("sf_mcc_set_target", *[_dp(a) for a in arr])

    def perform(self, it, time):
        """MCC.perform at step it and time `time`; returns the number of collisions."""
        n = C.c_int64()
        self._call("sf_mcc_perform", it, time, C.byref(n))
        return n.value

    def field(self, name):
        out = np.zeros(self.sim.nn)
        self._call("sf_mcc_field", self.FIELDS[name], _dp(out))
        return out

    def clearSamples(self, clear_source=False):
        self._call("sf_mcc_clear_samples", 1 if clear_source else 0)

    def close(self):
        if self.h:
            self.lib.sf_mcc_destroy(self.h)
            self.h = None


class DsmcHip:
    """`<dsmc>` of the reference's collisions plugin (collisions/DSMC.java:45-124) between two
    GPU species (the same one for a single-species gas), both created with track_ids=True.

    sigma: "const" (sigma_c0 in m^2) or "bird463" (VHS from each material's diam, ref_temp and
    visc_temp_index, given as pairs); vss_alpha: pair of the materials' <vss_alpha>;
    sig_cr_max: initial value; frequency: the reference's key; seed and tag key the per-cell
    Philox streams (tag None: the library picks a free one)."""

    FIELDS = {"col-count-sum": 0, "col-real-sum": 1, "col-count": 2, "nu": 3}
    SIGMA = {"CONST": 0, "BIRD463": 1}

    def __init__(self, sim, mat_a, mat_b, sigma="bird463", sigma_c0=0.0, diam=(0.0, 0.0), ref_temp=(275.0, 275.0),
                 visc_temp_index=(0.85, 0.85), vss_alpha=(1.0, 1.0), sig_cr_max=1e-16, frequency=1, seed=0, tag=None):
        self.sim = sim
        self.lib = _lib.load()
        tag = TAG_AUTO if tag is None else tag
        kind = sigma.upper()
        if kind not in self.SIGMA:
            raise ValueError("unsupported DSMC cross-section %s (const and bird463 are provided)" % sigma)
        d = _lib.DsmcDesc(mat_a.index, mat_b.index, self.SIGMA[kind], sigma_c0, diam[0], diam[1], ref_temp[0],
                          ref_temp[1], visc_temp_index[0], visc_temp_index[1], vss_alpha[0], vss_alpha[1], sig_cr_max,
                          frequency, seed, tag)
        h = C.c_void_p()
        sim.ctx.call("sf_dsmc_create", C.byref(d), C.byref(h))
        self.h = h

    def _call(self, name, *args):
        return _lib.check(self.sim.ctx.h, getattr(self.lib, name)(self.h, *args))

    def perform(self, it):
        n = C.c_int64()
        self._call("sf_dsmc_perform", it, C.byref(n))
        return n.value

    def field(self, name):
        out = np.zeros(self.sim.nn)
        self._call("sf_dsmc_field", self.FIELDS[name], _dp(out))
        return out

    def sig_cr_max(self):
        m = self.sim.deck.mesh
        out = np.zeros((m.ni - 1) * (m.nj - 1))
        self._call("sf_dsmc_field", 4, _dp(out))
        return out

    def clearSamples(self):
        self._call("sf_dsmc_clear_samples")

    def close(self):
        if self.h:
            self.lib.sf_dsmc_destroy(self.h)
            self.h = None
