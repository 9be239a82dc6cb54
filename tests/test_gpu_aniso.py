"""VTI (PHYSICS 3) and TTI (PHYSICS 4) elastic forward modelling on the GPU
against the FP32 oracle (src/VTI/VTI.c, src/TTI/TTI.c with
update_s_elastic_PML_VTI.c / _TTI.c): seismograms and final wavefields
bit for bit for every FD order, with and without the free-surface switch,
every scalar stress tile shape and vector velocity shapes; the error
against the FP64 oracle stays within the seismogram tolerance (rel-L2
1e-5).  The drop-in executable writes the library's seismograms for
PHYSICS 3 and 4 parameter files."""
import os
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


def aniso_case(physics, nx=96, ny=64, nt=300, fdorder=8, free_surf=1, fw=10, seismo=4, stype=1, seed=5,
               tilt=30.0):
    rho, c11, c13, c33, c44, theta = synth.anisotropic_model(nx, ny, 10.0, seed, tilt=tilt if physics == 4 else None)
    vmax = float(np.sqrt(np.maximum(c11, c33) / rho).max())
    dt = float(np.float32(synth.stable_dt(vmax, 10.0, fdorder, safety=0.6)))
    p = dnhip.Params(nx=nx, ny=ny, nt=nt, dh=10.0, dt=dt, fdorder=fdorder, maxrelerror=1, fw=fw,
                     free_surf=free_surf, seismo=seismo, ts=0.05, physics=physics)
    src = synth.source_table(10.0 * (nx // 2), 40.0, 10.0, fc=20.0, stype=stype, azimuth=30.0)
    sig = synth.ricker(nt, dt, 20.0, amp=1e3)[None, :]
    rec = synth.receiver_line(30.0, 20.0, nx // 2 - 4, 30.0, 10.0)
    return p, (rho, c11, c13, c33, c44, theta), src, sig, rec


def gpu_run(p, m, src, sig, rec, tiles=None):
    g = dnhip.Context(p)
    g.set_model_aniso(*m)
    g.set_sources(src, sig)
    g.set_receivers(rec)
    if tiles:
        g.set_tile(*tiles[0], sweep=0)
        g.set_tile(*tiles[1], sweep=1)
    g.forward()
    return g


def oracle_run(p, m, src, sig, rec, prec="f32"):
    rho, c11, c13, c33, c44, theta = m
    o = orc.Oracle(p, None, None, rho, prec=prec, aniso=(p.physics, c11, c13, c33, c44, theta))
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    return o


def assert_equal_runs(g, o, seismo=4):
    comps = {1: (0, 1), 2: (2,), 3: (), 4: (0, 1, 2)}[seismo]
    for c in comps:
        a, b = g.seismogram(c), o.seismogram(c)
        assert np.array_equal(a, b), (c, rel_l2(a, b))
        assert np.abs(b).max() > 0, c
    for fid in (dnhip.F_VX, dnhip.F_VY, dnhip.F_SXX, dnhip.F_SYY, dnhip.F_SXY):
        assert np.array_equal(g.field(fid), o.field(fid)), fid


@pytest.mark.parametrize("physics", [3, 4])
@pytest.mark.parametrize("fdorder", [2, 4, 6, 8, 10, 12])
def test_forward_equals_oracle(physics, fdorder):
    print(f"[aniso] physics {physics} FD{fdorder}", flush=True)
    case = aniso_case(physics, fdorder=fdorder)
    g = gpu_run(*case)
    o = oracle_run(*case)
    assert_equal_runs(g, o)
    g.close()


@pytest.mark.parametrize("physics", [3, 4])
@pytest.mark.parametrize("free_surf,seismo", [(0, 1), (0, 2), (1, 3)])
def test_boundaries_and_receiver_modes(physics, free_surf, seismo):
    """CPML on four sides (FREE_SURF 0) or three; SEISMO 1 (vx, vy), 2 (p),
    3 (nothing: the VTI/TTI sampler has no div/curl)."""
    print(f"[aniso] physics {physics} free_surf {free_surf} seismo {seismo}", flush=True)
    case = aniso_case(physics, free_surf=free_surf, seismo=seismo, nt=260)
    g = gpu_run(*case)
    o = oracle_run(*case)
    if seismo == 3:
        for c in range(5):
            assert not g.seismogram(c).any(), c
    assert_equal_runs(g, o, seismo)
    for k in range(9):
        assert np.array_equal(g.psi(k), o.psi(k)), k
    g.close()


@pytest.mark.parametrize("physics", [3, 4])
@pytest.mark.parametrize("stype", [2, 3, 4, 5])
def test_source_types(physics, stype):
    """Forces in x, y and rotated (velocity sweep) and the moment tensor."""
    print(f"[aniso] physics {physics} source type {stype}", flush=True)
    case = aniso_case(physics, stype=stype, nt=240)
    g = gpu_run(*case)
    o = oracle_run(*case)
    assert_equal_runs(g, o)
    g.close()


SCALAR = [(64, 4, 1, 1), (128, 2, 1, 1), (256, 1, 1, 1), (64, 4, 2, 1), (128, 2, 4, 1), (64, 8, 2, 1)]
VECTOR = [(64, 4, 1, 4), (64, 4, 2, 4), (64, 2, 4, 4), (64, 4, 1, 2), (128, 2, 2, 2), (64, 4, 2, 2)]


@pytest.mark.parametrize("physics", [3, 4])
def test_tile_shapes_do_not_change_results(physics):
    case = aniso_case(physics, nx=150, ny=90, nt=160)
    o = oracle_run(*case)
    for k, ts in enumerate(SCALAR):
        tv = VECTOR[k]
        print(f"[aniso] physics {physics} tiles v{tv} s{ts}", flush=True)
        g = gpu_run(*case, tiles=(tv, ts))
        assert_equal_runs(g, o)
        g.close()


def test_vector_stress_tiles_are_refused_and_autotune_keeps_scalar():
    case = aniso_case(4, nt=60)
    g = dnhip.Context(case[0])
    g.set_model_aniso(*case[1])
    g.set_sources(case[2], case[3])
    g.set_receivers(case[4])
    with pytest.raises(dnhip.DnError, match="scalar"):
        g.set_tile(64, 4, 1, 4, sweep=1)
    g.autotune(10)
    assert g.tile(1)[3] == 1
    g.forward()
    o = oracle_run(*case)
    assert_equal_runs(g, o)
    g.close()


def test_zero_tilt_tti_equals_vti_on_the_gpu():
    p, m, src, sig, rec = aniso_case(3)
    nx, ny = p.nx, p.ny
    v = gpu_run(p, m, src, sig, rec)
    pt = dnhip.Params(**{**p.__dict__, "physics": 4})
    t = gpu_run(pt, m[:5] + (np.zeros((nx, ny), np.float32),), src, sig, rec)
    for c in (0, 1, 2):
        assert np.array_equal(v.seismogram(c), t.seismogram(c)), c


@pytest.mark.parametrize("physics", [3, 4])
def test_error_against_fp64_oracle(physics):
    case = aniso_case(physics, nx=120, ny=80, nt=400)
    g = gpu_run(*case)
    o = oracle_run(*case, prec="f64")
    for c in (0, 1, 2):
        err = rel_l2(g.seismogram(c), o.seismogram(c))
        print(f"[aniso] physics {physics} comp {c} rel-L2 vs FP64 oracle {err:.2e}", flush=True)
        assert err <= SEIS_TOL


def test_errors_of_the_anisotropic_entry_points():
    p, m, src, sig, rec = aniso_case(3, nt=40)
    g = dnhip.Context(p)
    with pytest.raises(dnhip.DnError, match="dn_set_model_aniso"):
        g.set_model(m[1], m[2], m[0])
    rho, c11, c13, c33, c44, _ = m
    # DT above the VTI stability limit of checkfd_ssg_VTI
    fast = dnhip.Params(**{**p.__dict__, "dt": p.dt * 4})
    h = dnhip.Context(fast)
    with pytest.raises(dnhip.DnError) as ei:
        h.set_model_aniso(rho, c11, c13, c33, c44)
    assert ei.value.code == -6
    h.close()
    g.set_model_aniso(rho, c11, c13, c33, c44)
    e = dnhip.Engine()
    with pytest.raises(dnhip.DnError, match="VTI/TTI"):
        e.attach(g)
    e.close()
    g.close()
    tti = dnhip.Context(dnhip.Params(**{**p.__dict__, "physics": 4}))
    with pytest.raises(dnhip.DnError, match="theta"):
        tti.set_model_aniso(rho, c11, c13, c33, c44)
    tti.close()
    # TTI.c stores ux/uy/uxy for GRAD_FORM 2, which the TTI update never writes
    with pytest.raises(dnhip.DnError, match="GRAD_FORM"):
        dnhip.Context(dnhip.Params(**{**p.__dict__, "physics": 4, "mode": 1, "grad_form": 2}))
    with pytest.raises(dnhip.DnError, match="L = 0"):
        dnhip.Context(dnhip.Params(**{**p.__dict__, "L": 2}))


@pytest.mark.parametrize("physics", [3, 4])
def test_cli_forward_writes_the_library_seismograms(tmp_path, physics):
    tmp = str(tmp_path)
    nx, ny, dh, nt = 110, 70, 10.0, 280
    rho, c11, c13, c33, c44, theta = synth.anisotropic_model(nx, ny, dh, 17, tilt=25.0 if physics == 4 else None)
    vmax = float(np.sqrt(np.maximum(c11, c33) / rho).max())
    dt = float(np.float32(synth.stable_dt(vmax, dh, 8, safety=0.6)))
    for d in ("model", "su"):
        os.makedirs(os.path.join(tmp, d))
    arrs = dict(c11=c11, c13=c13, c33=c33, c44=c44, rho=rho)
    if physics == 4:
        arrs["theta"] = theta
    for ext, a in arrs.items():
        np.ascontiguousarray(a, np.float32).tofile(os.path.join(tmp, f"model/m.{ext}"))
    with open(os.path.join(tmp, "source.dat"), "w") as f:
        f.write("2\n400.00 0.0 30.00 0.0 15.00 1.00 0.0 1\n800.00 0.0 20.00 0.0 12.00 2.00 0.0 1\n")
    with open(os.path.join(tmp, "receiver.dat"), "w") as f:
        for k in range(50):
            f.write(f"{50.0 + 20.0 * k:.3f}\t20.000\n")
    inpfile.write_inp(os.path.join(tmp, "m.inp"), PHYSICS=physics, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      SEISMO=4, L=2, FL="3.0 30.0")
    print(f"[aniso cli] PHYSICS {physics} in {tmp}", flush=True)
    r = subprocess.run([BIN, "m.inp"], cwd=tmp, capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr
    # the model copies the reference's readers write
    for ext in ("c11", "c13", "c33", "c44"):
        got = np.fromfile(os.path.join(tmp, f"model/m.denise.{ext}"), np.float32)
        assert np.array_equal(got, arrs[ext].ravel()), ext
    if physics == 4:
        rad = np.fromfile(os.path.join(tmp, "model/m.denise.theta"), np.float32)
        assert np.array_equal(rad, (theta.astype(np.float64) * 3.141592653589793 / 180.0).astype(np.float32).ravel())
        assert not os.path.exists(os.path.join(tmp, "model/m.denise.theta.0.0"))
    par = dnio.read_par(os.path.join(tmp, "m.inp"))
    cwd = os.getcwd()
    os.chdir(tmp)
    try:
        src, ts = dnio.read_sources(par.get("source_file"))
        rec = dnio.read_receivers(par)
    finally:
        os.chdir(cwd)
    p = dnhip.Params(nx=nx, ny=ny, nt=par.nt, dh=dh, dt=par.dt, fdorder=8, maxrelerror=par.maxrelerror, fw=par.fw,
                     free_surf=par.free_surf, seismo=4, ts=ts, physics=physics, damping=par.damping,
                     fpml=par.fpml, npower=par.npower, k_max_pml=par.k_max_pml)
    g = dnhip.Context(p)
    g.set_model_aniso(rho, c11, c13, c33, c44, theta)
    names = {0: "su/out_x.su", 1: "su/out_y.su", 2: "su/out_p.su"}
    for ishot in (1, 2):
        loc = dnio.split_sources(par, src, ishot - 1, 1)
        sig = dnio.wavelet(par, loc, ishot, ts)
        g.set_sources(loc, sig)
        g.set_receivers(rec[:2])
        g.forward()
        for c, name in names.items():
            _, data = dnio.parse_su(os.path.join(tmp, f"{name}.shot{ishot}"))
            assert np.array_equal(data, g.seismogram(c)), (ishot, c)
    g.close()


def _images(runner, obs):
    runner.forward()
    runner.residuals(obs[0], obs[1], obs[2], 1)
    runner.adjoint()
    return [runner.gradient(k) for k in range(3)]


@pytest.mark.parametrize("physics,grad_form,dtinv,quelltypb",
                         [(3, 1, 2, 1), (3, 2, 1, 1), (3, 1, 3, 4), (4, 1, 2, 1), (4, 1, 1, 4)])
def test_rtm_images_equal_oracle(physics, grad_form, dtinv, quelltypb):
    """The VTI/TTI adjoint pass (src/VTI/VTI.c:235-368, src/TTI/TTI.c):
    forward with the snapshot store, residual adjoint sources (pressure
    too for QUELLTYPB 4), adjoint with the density and P correlations; the
    raw images equal the FP32 oracle's bit for bit and the shear image
    stays 0."""
    print(f"[aniso rtm] physics {physics} GRAD_FORM {grad_form} DTINV {dtinv} QUELLTYPB {quelltypb}", flush=True)
    p, m, src, sig, rec = aniso_case(physics, nt=240, seismo=4)
    p = dnhip.Params(**{**p.__dict__, "mode": 1, "grad_form": grad_form, "dtinv": dtinv, "quelltypb": quelltypb})
    g = dnhip.Context(p)
    g.set_model_aniso(*m)
    g.set_sources(src, sig)
    g.set_receivers(rec)
    o = orc.Oracle(p, None, None, m[0], prec="f32", aniso=(physics, m[1], m[2], m[3], m[4], m[5]))
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    obs = (0.8 * o.seismogram(0), 0.9 * o.seismogram(1), 1.1 * o.seismogram(2))
    gi = _images(g, obs)
    oi = _images(o, obs)
    for k in range(3):
        assert np.array_equal(gi[k], oi[k]), (k, rel_l2(gi[k], oi[k]))
    assert np.abs(gi[0]).max() > 0 and np.abs(gi[2]).max() > 0
    assert not gi[1].any()
    with pytest.raises(dnhip.DnError, match="raw correlation"):
        g.assemble()
    g.close()


def test_cli_rtm_images_equal_the_library(tmp_path):
    """MODE 2 with PHYSICS 3 (RTM_VTI.c): the executable's P and S images
    are the stacked raw correlations of the library's adjoint passes with
    the time-integrated observed data as adjoint sources (C_vp = C_vs =
    C_rho = 1, no assembly), bit for bit; the S image is 0."""
    import fwi_oracle as fo
    tmp = str(tmp_path)
    nx, ny, dh, nt = 100, 64, 10.0, 360
    rho, c11, c13, c33, c44, _ = synth.anisotropic_model(nx, ny, dh, 23)
    vmax = float(np.sqrt(np.maximum(c11, c33) / rho).max())
    dt = float(np.float32(synth.stable_dt(vmax, dh, 8, safety=0.6)))
    for d in ("model", "su", "obs", "jac"):
        os.makedirs(os.path.join(tmp, d))
    true = dict(c11=c11.copy(), c13=c13, c33=c33.copy(), c44=c44, rho=rho)
    true["c11"][30:70, 30:34] *= 1.2
    true["c33"][30:70, 30:34] *= 1.2
    mig = dict(c11=c11, c13=c13, c33=c33, c44=c44, rho=rho)
    for name, mod in (("true", true), ("mig", mig)):
        for ext, a in mod.items():
            np.ascontiguousarray(a, np.float32).tofile(os.path.join(tmp, f"model/{name}.{ext}"))
    with open(os.path.join(tmp, "source.dat"), "w") as f:
        f.write("2\n300.00 0.0 30.00 0.0 15.00 1.00 0.0 1\n650.00 0.0 30.00 0.0 15.00 1.00 0.0 1\n")
    with open(os.path.join(tmp, "receiver.dat"), "w") as f:
        for k in range(40):
            f.write(f"{40.0 + 20.0 * k:.3f}\t20.000\n")
    inpfile.write_inp(os.path.join(tmp, "true.inp"), PHYSICS=3, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      MFILE="model/true", SEIS_FILE_VX="obs/d_x.su", SEIS_FILE_VY="obs/d_y.su")
    r = subprocess.run([BIN, "true.inp"], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    inpfile.write_inp(os.path.join(tmp, "rtm.inp"), PHYSICS=3, MODE=2, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      MFILE="model/mig", DTINV=2, RTM_SHOT=0, TESTSHOT="1, 2, 1")
    inpfile.write_workflow(os.path.join(tmp, "wf.inp"))
    print(f"[aniso cli rtm] in {tmp}", flush=True)
    r = subprocess.run([BIN, "rtm.inp", "wf.inp"], cwd=tmp, capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr
    par = dnio.read_par(os.path.join(tmp, "rtm.inp"))
    cwd = os.getcwd()
    os.chdir(tmp)
    try:
        src, ts = dnio.read_sources(par.get("source_file"))
        rec = dnio.read_receivers(par)
    finally:
        os.chdir(cwd)
    p = dnhip.Params(nx=nx, ny=ny, nt=par.nt, dh=dh, dt=par.dt, fdorder=8, maxrelerror=par.maxrelerror, fw=par.fw,
                     free_surf=par.free_surf, seismo=1, ts=ts, physics=3, mode=1, dtinv=2, grad_form=par.grad_form,
                     damping=par.damping, fpml=par.fpml, npower=par.npower, k_max_pml=par.k_max_pml)
    g = dnhip.Context(p)
    g.set_model_aniso(rho, c11, c13, c33, c44)
    total = [np.zeros((nx, ny), np.float32) for _ in range(2)]
    for ishot in (1, 2):
        loc = dnio.split_sources(par, src, ishot - 1, 1)
        g.set_sources(loc, dnio.wavelet(par, loc, ishot, ts))
        g.set_receivers(rec[:2])
        g.forward()
        obs = [dnio.read_su(os.path.join(tmp, f"obs/d_{c}.su.shot{ishot}"), rec.shape[1], par.nt) for c in "xy"]
        diffs = [fo.rtm_adjoint_source(obs[c], g.seismogram(c), par.dt, par.grad_form, 1)[0] for c in (0, 1)]
        g.set_adjoint_sources(diffs[0], diffs[1], None)
        g.adjoint()
        for k in range(2):
            total[k] += g.gradient(k)
    g.close()
    for k, name in enumerate(("P", "S")):
        got = np.fromfile(os.path.join(tmp, f"jac/j_{name}_image"), np.float32).reshape(nx, ny)
        assert np.array_equal(got, total[k]), name
    assert np.abs(total[0]).max() > 0 and not total[1].any()
