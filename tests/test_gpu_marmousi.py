"""The reference's shipped Marmousi example through the drop-in executable:
its parameter file (tests/golden/marmousi/DENISE_marm_OBC.inp: 500 x 174,
DH 20 m, DT 2 ms, 6 s, FD8, CPML 10, free surface, the band-limited spike
source QUELLART 6 with its 15 Hz order-5 Butterworth low-pass, 400 OBC
receivers) and its receiver file, with three of the 100 shots of its
source file (1, 50, 100) and a synthetic model where MFILE points (the
example's model is not shipped: a smooth layered model under a 400 m
water column, vs = 0 in the water).  The SU files equal the FP32 oracle's
seismograms for the same inputs."""
import os
import shutil

import numpy as np
import pytest

import dnhip
import dnio
import orc
import synth
import tdfilt_oracle as tf
from cases import rel_l2
from test_gpu_cli import _reference_run, _run

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "marmousi")


def test_shipped_marmousi_example_forward(tmp_path):
    tmp = str(tmp_path)
    for d in ("start", "su", "log", "source", "receiver"):
        os.makedirs(os.path.join(tmp, d), exist_ok=True)
    shutil.copy(os.path.join(GOLD, "DENISE_marm_OBC.inp"), tmp)
    shutil.copy(os.path.join(GOLD, "receiver_OBC.dat"), os.path.join(tmp, "receiver"))
    lines = open(os.path.join(GOLD, "source_OBC_VSP.dat")).read().splitlines()
    keep = [lines[1], lines[50], lines[100]]
    with open(os.path.join(tmp, "source", "source_OBC_VSP.dat"), "w") as f:
        f.write(f"{len(keep)}\n" + "\n".join(keep) + "\n")
    par = dnio.read_par(os.path.join(tmp, "DENISE_marm_OBC.inp"))
    assert (par.nx, par.ny, par.quellart, par.nt) == (500, 174, 6, 3000)
    vp, vs, rho = synth.layered_model(par.nx, par.ny, par.dh, seed=4, sigma=4.0)
    water = 20
    vp[:, :water], vs[:, :water], rho[:, :water] = 1500.0, 0.0, 1000.0
    for ext, a in (("vp", vp), ("vs", vs), ("rho", rho)):
        np.ascontiguousarray(a, np.float32).tofile(os.path.join(tmp, f"start/marmousi_II_marine.{ext}"))
    out = _run(tmp, "DENISE_marm_OBC.inp")
    assert "wrote su/DENISE_MARMOUSI_x.su.shot3" in out

    par, p, m, rec, shots, src = _reference_run(tmp, "DENISE_marm_OBC.inp")
    g = dnhip.Context(p)
    g.set_model(*m)
    for ishot in (1, 3):
        loc, sig = shots[ishot - 1]
        sig = tf.apply_tdfilt(sig, par.dt, par.order_spike, par.fc_spike_2, par.fc_spike_1)
        g.set_sources(loc, sig)
        g.set_receivers(rec[:2])
        g.forward()
        o = orc.Oracle(p, *m, prec="f32")
        orc.set_threads(min(16, os.cpu_count() or 1))
        o.set_sources(loc, sig)
        o.set_receivers(rec[:2])
        o.forward()
        for c, name in ((0, "x"), (1, "y")):
            hdr, data = dnio.parse_su(os.path.join(tmp, f"su/DENISE_MARMOUSI_{name}.su.shot{ishot}"))
            assert np.array_equal(data, g.seismogram(c)), (ishot, name)
            err = rel_l2(data, o.seismogram(c))
            print(f"[marmousi] shot {ishot} {name}: rel-L2 vs FP32 oracle {err:.2e}, "
                  f"max |v| {np.abs(data).max():.3e}", flush=True)
            assert err <= 1e-5 and np.abs(data).max() > 0
        o.close()
    g.close()
