"""Device-side boundary injection (sf_sources_sample_device) against the host sampler that
replays the reference's java.util.Random stream (MaxwellianSource.sampleParticle,
sources/MaxwellianSource.java:93-125): identical particle counts per step, and the same
distributions of position along the inlet and of velocity (drift + Utils.SampleMaxw3D with the
1.2 factor), checked with two-sample Kolmogorov-Smirnov statistics and closed-form moments."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import stats

from starfish_amd import deck as D

K_B, AMU = 1.3806503e-23, 1.66053886e-27


def _inject(device, steps):
    from starfish_amd.plugin import Simulation
    deck = D.config_c1(steps=steps)
    sim = Simulation(deck, capacity=200000)
    counts = []
    try:
        mat = sim.materials[0]
        for _ in range(steps):
            n = C.c_int64()
            if device:
                sim.ctx.call("sf_sources_sample_device", C.c_uint64(12345), C.byref(n))
            else:
                sim.ctx.call("sf_sources_sample", C.byref(n))
            counts.append(n.value)
        return deck, counts, mat.particles()
    finally:
        sim.close()


@pytest.mark.gpu
def test_device_injection_matches_host_sampler():
    steps = 60
    deck, ch, ph = _inject(False, steps)
    _, cd, pd = _inject(True, steps)
    print("particles host %d device %d" % (sum(ch), sum(cd)), flush=True)
    assert ch == cd and sum(cd) > 20000
    assert len(pd["x"]) == sum(cd)
    src = deck.sources[0]
    vth = math.sqrt(2 * K_B * src.temperature / (16.0 * AMU))
    sig = 1.2 * vth * math.sqrt(0.5)            # Utils.SampleMaxw3D: (vth 1.2) sqrt(1/2) N(0,1)
    for name in ("y", "u", "v", "w"):
        d = stats.ks_2samp(ph[name], pd[name])
        print("%s: KS D %.4f p %.3g" % (name, d.statistic, d.pvalue), flush=True)
        assert d.pvalue > 1e-4, name
    n = len(pd["u"])
    assert np.all(pd["u"] > 0)                  # inward normal of the left inlet is +x
    assert abs(pd["u"].mean() - src.v_drift) < 5 * sig / math.sqrt(n)
    for name in ("v", "w"):
        assert abs(pd[name].mean()) < 5 * sig / math.sqrt(n)
        assert abs(pd[name].std() / sig - 1) < 0.03
    # positions: on the inlet line (x = -0.15 plus the 1e-6 dt nudge), uniform along it
    inlet = deck.boundary("inlet").points
    y0, y1 = min(p[1] for p in inlet), max(p[1] for p in inlet)
    assert np.all(np.abs(pd["x"] - inlet[0][0]) < 1e-6) and pd["y"].min() >= y0 and pd["y"].max() <= y1
    assert stats.kstest((pd["y"] - y0) / (y1 - y0), "uniform").pvalue > 1e-4
