"""The drop-in executable on the GPU: denise_hip reads a DENISE parameter
file, model, source and receiver files, and writes the SU seismograms
(MODE 0) or the stacked FWI gradient (MODE 1).  Its output must equal what
the library gives for the same inputs through the Python binding, and the
FP32 oracle within the parity tolerances."""
import os
import re
import subprocess

import numpy as np
import pytest

import dnhip
import dnio
import inpfile
import orc
import synth
from cases import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "denise-black-edition_amd", "bin", "denise_hip")
SEIS_TOL = 1e-5
GRAD_TOL = 1e-4


def _files(tmp, nx, ny, dh, model, sources, receivers, L=0, mname="model/m"):
    for d in ("model", "su", "obs", "jac"):
        os.makedirs(os.path.join(tmp, d), exist_ok=True)
    vp, vs, rho = model
    for ext, a in (("vp", vp), ("vs", vs), ("rho", rho)):
        np.ascontiguousarray(a, np.float32).tofile(os.path.join(tmp, f"{mname}.{ext}"))
    if L:
        for ext in ("qp", "qs"):
            np.full((nx, ny), 50.0, np.float32).tofile(os.path.join(tmp, f"{mname}.{ext}"))
    with open(os.path.join(tmp, "source.dat"), "w") as f:
        f.write(f"{len(sources)}\n")
        for (x, y, fc, amp, typ) in sources:
            f.write(f"{x:.2f} 0.0 {y:.2f} 0.0 {fc:.2f} {amp:.2f} 0.0 {typ}\n")
    with open(os.path.join(tmp, "receiver.dat"), "w") as f:
        for (x, y) in receivers:
            f.write(f"{x:.3f}\t{y:.3f}\n")


def _run(tmp, *args):
    print(f"[cli] {' '.join(args)} in {tmp}", flush=True)
    r = subprocess.run([BIN, *args], cwd=tmp, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _reference_run(tmp, inp, mode_params=None):
    """The same shots through the Python binding of the library."""
    par = dnio.read_par(os.path.join(tmp, inp))
    cwd = os.getcwd()
    os.chdir(tmp)
    try:
        src, ts = dnio.read_sources(par.get("source_file"))
        rec = dnio.read_receivers(par)
        vp, vs, rho, taus, taup = dnio.read_model(par)
    finally:
        os.chdir(cwd)
    p = dnhip.Params(nx=par.nx, ny=par.ny, nt=par.nt, dh=par.dh, dt=par.dt, fdorder=par.fdorder,
                     maxrelerror=par.maxrelerror, fw=par.fw, free_surf=par.free_surf, L=par.L,
                     seismo=par.seismo, ndt=par.ndt, invmat1=par.invmat1, grad_form=par.grad_form,
                     dtinv=par.dtinv, quelltypb=par.quelltypb, mode=par.mode, damping=par.damping,
                     fpml=par.fpml, npower=par.npower, k_max_pml=par.k_max_pml, ts=ts,
                     fl=list(par.fl)[:par.L])
    shots = []
    for ishot in range(1, src.shape[1] + 1):
        loc = dnio.split_sources(par, src, ishot - 1, 1)
        sig = dnio.wavelet(par, loc, ishot, ts)
        shots.append((loc, sig))
    return par, p, (vp, vs, rho, taus, taup), rec, shots, src


@pytest.mark.parametrize("L,seismo", [(0, 1), (2, 4)])
def test_cli_forward_writes_the_library_seismograms(tmp_path, L, seismo):
    tmp = str(tmp_path)
    nx, ny, dh, dt, nt = 120, 80, 10.0, 1e-3, 300
    model = synth.layered_model(nx, ny, dh, seed=21, sigma=3.0)
    sources = [(400.0, 30.0, 15.0, 1.0, 1), (800.0, 20.0, 12.0, 2.0, 1)]
    receivers = [(50.0 + 20.0 * k, 20.0) for k in range(50)]
    _files(tmp, nx, ny, dh, model, sources, receivers, L=L)
    inpfile.write_inp(os.path.join(tmp, "m.inp"), NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt, L=L,
                      FL="3.0 30.0", SEISMO=seismo, NDT=2 if seismo == 4 else 1)
    out = _run(tmp, "m.inp")
    assert "wrote su/out_x.su.shot2" in out
    par, p, m, rec, shots, src = _reference_run(tmp, "m.inp")
    g = dnhip.Context(p)
    g.set_model(*m)
    comps = [0, 1] if seismo == 1 else [0, 1, 2]
    names = {0: "su/out_x.su", 1: "su/out_y.su", 2: "su/out_p.su"}
    for ishot, (loc, sig) in enumerate(shots, start=1):
        g.set_sources(loc, sig)
        g.set_receivers(rec[:2])
        g.forward()
        for c in comps:
            hdr, data = dnio.parse_su(os.path.join(tmp, f"{names[c]}.shot{ishot}"))
            want = g.seismogram(c)[:, par.ndt - 1::par.ndt]
            assert data.shape == want.shape
            assert np.array_equal(data, want), (ishot, c)
            assert np.all(hdr["sx"] == round(src[0, ishot - 1] * 1000))
        if ishot == 1:
            o = orc.Oracle(p, *m, prec="f32")
            o.set_sources(loc, sig)
            o.set_receivers(rec[:2])
            o.forward()
            for c in comps:
                assert rel_l2(g.seismogram(c), o.seismogram(c)) <= SEIS_TOL


def test_cli_gradient_stacks_the_library_gradients(tmp_path):
    tmp = str(tmp_path)
    nx, ny, dh, dt, nt = 96, 64, 10.0, 1e-3, 400
    bg = synth.layered_model(nx, ny, dh, seed=5, sigma=3.0)
    true = tuple(a.copy() for a in bg)
    true[0][40:56, 10:20] *= 1.08
    true[1][40:56, 10:20] *= 1.08
    sources = [(300.0, 30.0, 15.0, 1.0, 1), (650.0, 30.0, 15.0, 1.0, 1)]
    receivers = [(40.0 + 20.0 * k, 20.0) for k in range(40)]
    # observed data: MODE 0 on the true model, written where DATA_DIR points
    _files(tmp, nx, ny, dh, true, sources, receivers, mname="model/true")
    inpfile.write_inp(os.path.join(tmp, "true.inp"), NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      MFILE="model/true", SEIS_FILE_VX="obs/d_x.su", SEIS_FILE_VY="obs/d_y.su")
    _run(tmp, "true.inp")
    _files(tmp, nx, ny, dh, bg, sources, receivers, mname="model/start")
    inpfile.write_inp(os.path.join(tmp, "fwi.inp"), MODE=1, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      MFILE="model/start", DTINV=2, RTM_SHOT=1, TESTSHOT="2, 2, 1")
    inpfile.write_workflow(os.path.join(tmp, "wf.inp"))
    out = _run(tmp, "fwi.inp", "wf.inp")
    misfit = float(re.search(r"misfit \(L2, test shots\): (\S+)", out).group(1))

    par, p, m, rec, shots, src = _reference_run(tmp, "fwi.inp")
    g = dnhip.Context(p)
    g.set_model(*m)
    o = orc.Oracle(p, *m, prec="f32")
    total = [np.zeros((nx, ny), np.float32) for _ in range(3)]
    ototal = [np.zeros((nx, ny), np.float32) for _ in range(3)]
    l2sum = 0.0
    for ishot, (loc, sig) in enumerate(shots, start=1):
        obs = [dnio.read_su(os.path.join(tmp, f"obs/d_{c}.su.shot{ishot}"), rec.shape[1], nt) for c in "xy"]
        for runner, acc in ((g, total), (o, ototal)):
            runner.set_sources(loc, sig)
            runner.set_receivers(rec[:2])
            runner.forward()
            l2 = runner.residuals(obs[0], obs[1], None, 1 if ishot == 2 else 0)
            runner.adjoint()
            runner.assemble()
            for k in range(3):
                acc[k] += runner.gradient(3 + k).astype(np.float32)
            if runner is g:
                l2sum += l2
                img = np.fromfile(os.path.join(tmp, f"jac/j_P_image_shot_{ishot}"), np.float32).reshape(nx, ny)
                assert np.array_equal(img, g.gradient(3))
    for k, name in enumerate(("vp", "vs", "rho")):
        got = np.fromfile(os.path.join(tmp, f"jac/j_grad_{name}"), np.float32).reshape(nx, ny)
        assert np.array_equal(got, total[k]), name
        assert np.abs(got).max() > 0
        assert rel_l2(got, ototal[k]) <= GRAD_TOL, name
    assert misfit == pytest.approx(l2sum, rel=1e-8) and misfit > 0
