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


def test_shipped_marmousi_fwi_workflow(tmp_path):
    """The shipped FWI workflow (tests/golden/marmousi/FWI_workflow_marmousi.inp:
    four stages, TIME_FILT low-pass 2/5/10/20 Hz, EPRECOND 3, LNORM 2) with the
    shipped parameter file (L-BFGS, horizontal gradient taper GRADT 21..25,
    EPS_SCALE 0.01, DTINV 3), changed only in MODE = 1, ITERMAX = 2 and the
    shot subset (shots 1, 50, 100, all test shots).  Observed data: MODE 0 on
    a true model (the starting model plus a fast lens).  The run ends
    normally, logs one line per iteration in the TIME_FILT format with the
    stage number, writes finite models, the stacked gradients and the taper,
    and its first iteration lowers the misfit or ends the stage."""
    import re
    tmp = str(tmp_path)
    for d in ("start", "su/MARMOUSI_spike", "log", "source", "receiver", "jacobian", "taper", "model"):
        os.makedirs(os.path.join(tmp, d), exist_ok=True)
    shutil.copy(os.path.join(GOLD, "receiver_OBC.dat"), os.path.join(tmp, "receiver"))
    shutil.copy(os.path.join(GOLD, "FWI_workflow_marmousi.inp"), tmp)
    lines = open(os.path.join(GOLD, "source_OBC_VSP.dat")).read().splitlines()
    with open(os.path.join(tmp, "source", "source_OBC_VSP.dat"), "w") as f:
        f.write("3\n" + "\n".join([lines[1], lines[50], lines[100]]) + "\n")
    text = open(os.path.join(GOLD, "DENISE_marm_OBC.inp")).read()

    def setv(t, key, value):
        t2, n = re.subn(r"(\(" + re.escape(key) + r"\)\s*=\s*)[^\n]*", lambda m: m.group(1) + value, t)
        assert n == 1, key
        return t2

    par0 = dnio.read_par(os.path.join(GOLD, "DENISE_marm_OBC.inp"))
    nx, ny, dh = par0.nx, par0.ny, par0.dh
    vp, vs, rho = synth.layered_model(nx, ny, dh, seed=4, sigma=4.0)
    vp[:, :20], vs[:, :20], rho[:, :20] = 1500.0, 0.0, 1000.0
    tvp, tvs = vp.copy(), vs.copy()
    tvp[200:300, 60:80] *= 1.1
    tvs[200:300, 60:80] *= 1.1
    for name, model in (("true", (tvp, tvs, rho)), ("marmousi_II_marine", (vp, vs, rho))):
        for ext, a in zip(("vp", "vs", "rho"), model):
            np.ascontiguousarray(a, np.float32).tofile(os.path.join(tmp, f"start/{name}.{ext}"))
    # observed data: MODE 0 on the true model, moved to DATA_DIR
    t0 = setv(text, "MODE", "0")
    t0 = t0.replace("MFILE = start/marmousi_II_marine", "MFILE = start/true")
    open(os.path.join(tmp, "true.inp"), "w").write(t0)
    _run(tmp, "true.inp")
    for s in (1, 2, 3):
        for c in ("x", "y"):
            os.replace(os.path.join(tmp, f"su/DENISE_MARMOUSI_{c}.su.shot{s}"),
                       os.path.join(tmp, f"su/MARMOUSI_spike/DENISE_MARMOUSI_{c}.su.shot{s}"))
    t1 = setv(text, "MODE", "1")
    t1 = setv(t1, "ITERMAX", "2")
    t1 = setv(t1, "TESTSHOT_START,TESTSHOT_END,TESTSHOT_INCR", "1, 3, 1")
    open(os.path.join(tmp, "fwi.inp"), "w").write(t1)
    out = _run(tmp, "fwi.inp", "FWI_workflow_marmousi.inp")
    log = [ln.split() for ln in open(os.path.join(tmp, "LOG_TEST.dat")).read().splitlines()]
    print(f"[marmousi-fwi] {len(log)} iterations; misfits {[float(r[7]) for r in log]}", flush=True)
    assert 1 <= len(log) <= 8 and all(len(r) == 11 for r in log)
    stages = [int(r[10]) for r in log]
    assert stages == sorted(stages) and stages[0] == 1
    assert [float(r[9]) for r in log if int(r[10]) == 1][0] == 2.0       # FC of stage 1
    assert float(log[0][7]) > 0
    if len(log) > 1 and stages[1] == 1:
        assert float(log[1][7]) < float(log[0][7])
    for ext in ("vp", "vs", "rho"):
        a = np.fromfile(os.path.join(tmp, f"model/modelTest_{ext}.bin"), np.float32)
        assert a.size == nx * ny and np.isfinite(a).all()
        g = np.fromfile(os.path.join(tmp, f"jacobian/jacobian_Test_grad_{ext}"), np.float32)
        assert g.size == nx * ny and np.isfinite(g).all()
    taper = np.fromfile(os.path.join(tmp, "taper/taper_coeff_hor.bin"), np.float32).reshape(nx, ny)
    assert np.all(taper[:, :20] == 0.0) and np.all(taper[:, 24] == 1.0)
    assert "TDFWI ITERATION 1 of 2 (stage 1 of 4)" in out
