"""Wavefield snapshots of the drop-in executable (MODE 0, SNAP = 4): with
SNAP > 0 only shot SNAP_SHOT is modelled (FD_PSV.c:264-268), and the
snapshot files hold, every TSNAPINC from TSNAP1 to TSNAP2, the library's
wavefields at those steps through the restated writer (snap_ssg: vx, vy,
div, curl) -- equal value for value."""
import os

import numpy as np
import pytest

import dnhip
import snap_oracle as so
import synth
import inpfile
from test_gpu_cli import _files, _reference_run, _run

pytestmark = pytest.mark.gpu


def test_cli_snapshots_equal_library_fields(tmp_path):
    tmp = str(tmp_path)
    nx, ny, dh, dt, nt = 80, 60, 10.0, 1e-3, 320
    model = synth.layered_model(nx, ny, dh, seed=9, sigma=3.0)
    sources = [(250.0, 30.0, 15.0, 1.0, 1), (500.0, 40.0, 12.0, 1.0, 1)]
    receivers = [(40.0 + 20.0 * k, 20.0) for k in range(30)]
    _files(tmp, nx, ny, dh, model, sources, receivers, mname="model/m")
    os.makedirs(os.path.join(tmp, "snap"), exist_ok=True)
    inpfile.write_inp(os.path.join(tmp, "snap.inp"), NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt, MFILE="model/m",
                      SNAP=4, SNAP_SHOT=2, TSNAP1=0.05, TSNAP2=0.3, TSNAPINC=0.05, SNAP_FORMAT=3,
                      SNAP_FILE="snap/s")
    _run(tmp, "snap.inp")
    assert os.path.exists(os.path.join(tmp, "su/out_x.su.shot2"))
    assert not os.path.exists(os.path.join(tmp, "su/out_x.su.shot1"))

    par, p, m, rec, shots, src = _reference_run(tmp, "snap.inp")
    g = dnhip.Context(p)
    g.set_model(*m)
    loc, sig = shots[1]
    g.set_sources(loc, sig)
    g.set_receivers(rec[:2])
    hc = g.hc()
    steps = list(range(50, 301, 50))
    want = {k: [] for k in ("x", "y", "div", "rot")}
    for t in steps:
        g.forward(t)
        f = [g.field(fid) for fid in (dnhip.F_VX, dnhip.F_VY, dnhip.F_SXX, dnhip.F_SYY)]
        vals = so.snap_values(4, par.fdorder, par.dh, *f, m[0], m[1], hc)
        for k in want:
            want[k].append(vals[k])
    g.close()
    for k, v in want.items():
        got = np.fromfile(os.path.join(tmp, f"snap/s.bin.{k}.0.0"), np.float32)
        assert got.size == len(steps) * nx * ny, k
        assert np.array_equal(got, np.concatenate(v)), k
        assert np.abs(got).max() > 0, k
