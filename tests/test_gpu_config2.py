"""The viscoelastic shot at full size (4000 x 1200, DH = 5 m, FD8 Holberg,
DT = 0.4 ms, L = 2, CPML 20 + free surface, 800 receivers) through the
autotuned kernels, shapes and cache policy: after the first 120 of its 6000
time steps the wavefields, memory variables, CPML state and the recorded
samples equal the FP32 oracle bit for bit.  The full 6000 steps stay finite
and carry energy to the receivers."""
import os
import sys

import numpy as np
import pytest

import dnhip
import orc

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
import bench  # noqa: E402  (the benchmark's workload definition)

FIELDS = [(dnhip.F_VX, "vx"), (dnhip.F_VY, "vy"), (dnhip.F_SXX, "sxx"), (dnhip.F_SYY, "syy"),
          (dnhip.F_SXY, "sxy")]
NSTEPS = 120


def test_config2_full_size_prefix_exact_and_full_run_finite():
    p, m, src, sig, rec = bench.workload(2)
    g = dnhip.Context(p)
    g.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    vt, st = g.autotune(20)
    print(f"[config2] shapes {vt} {st}, cache policy {g.store_hint(0)} {g.store_hint(1)}", flush=True)
    g.forward(NSTEPS)
    orc.set_threads(min(16, os.cpu_count() or 1))
    o = orc.Oracle(p, m["vp"], m["vs"], m["rho"], m["taus"], m["taup"], prec="f32")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward(NSTEPS)
    for fid, name in FIELDS:
        assert np.array_equal(g.field(fid), o.field(fid)), name
    for l in (1, 2):
        for which in range(3):
            assert np.array_equal(g.memvar(which, l), o.memvar(which, l)), ("memvar", which, l)
    for pid in range(9):
        assert np.array_equal(g.psi(pid), o.psi(pid)), ("psi", pid)
    for comp in (0, 1):
        assert np.array_equal(g.seismogram(comp)[:, :NSTEPS], o.seismogram(comp)[:, :NSTEPS]), comp
    o.close()
    g.forward()
    seis = g.seismogram(1)
    assert np.isfinite(seis).all() and np.abs(seis).max() > 0
