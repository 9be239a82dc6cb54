"""Acoustic (PHYSICS 2) forward modelling on the GPU against the FP32
oracle (src/AC/ac.c with update_v_PML_AC.c, update_s_acoustic_PML_AC.c,
psource_AC.c, surface_acoustic_PML_AC.c, seismo_AC.c): seismograms, the
final pressure and velocities and the CPML memory bit for bit for every FD
order, with and without the free surface, every SEISMO mode, every source
type, both INVMAT1 parametrisations and every tile shape; the error against
the FP64 oracle stays within the seismogram tolerance (rel-L2 1e-5).  The
drop-in executable writes the library's seismograms and the reference's
model copies for PHYSICS 2 parameter files."""
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


def ac_case(nx=96, ny=64, nt=300, fdorder=8, free_surf=1, fw=10, seismo=4, stype=1, seed=4, invmat1=1,
            sy=40.0):
    vp, _, rho = synth.layered_model(nx, ny, 10.0, seed, sigma=3.0)
    dt = float(np.float32(synth.stable_dt(float(vp.max()), 10.0, fdorder, safety=0.6)))
    p = dnhip.Params(nx=nx, ny=ny, nt=nt, dh=10.0, dt=dt, fdorder=fdorder, maxrelerror=1, fw=fw,
                     free_surf=free_surf, seismo=seismo, ts=0.05, physics=2, invmat1=invmat1)
    if invmat1 == 3:  # the modulus rho vp^2 in place of vp
        vp = (rho.astype(np.float64) * vp.astype(np.float64) ** 2).astype(np.float32)
    src = synth.source_table(10.0 * (nx // 2), sy, 10.0, fc=20.0, stype=stype, azimuth=30.0)
    sig = synth.ricker(nt, dt, 20.0, amp=1e3)[None, :]
    rec = synth.receiver_line(30.0, 20.0, nx // 2 - 4, 30.0, 10.0)
    return p, (vp, rho), src, sig, rec


def gpu_run(p, m, src, sig, rec, tiles=None):
    g = dnhip.Context(p)
    g.set_model(m[0], None, m[1])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    if tiles:
        g.set_tile(*tiles[0], sweep=0)
        g.set_tile(*tiles[1], sweep=1)
    g.forward()
    return g


def oracle_run(p, m, src, sig, rec, prec="f32"):
    o = orc.Oracle(p, m[0], None, m[1], prec=prec, aniso="ac")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    return o


def assert_equal_runs(g, o, seismo=4, nonzero=True):
    comps = {1: (0, 1), 2: (2,), 3: (3, 4), 4: (0, 1, 2, 3, 4)}[seismo]
    for c in comps:
        a, b = g.seismogram(c), o.seismogram(c)
        assert np.array_equal(a, b), (c, rel_l2(a, b))
        if nonzero:
            assert np.abs(b).max() > 0, c
    for fid in (dnhip.F_VX, dnhip.F_VY, dnhip.F_SXX):
        assert np.array_equal(g.field(fid), o.field(fid)), fid
    # the pressure is the sxx plane; syy reads back the same plane
    assert np.array_equal(g.field(dnhip.F_SYY), g.field(dnhip.F_SXX))
    assert not g.field(dnhip.F_SXY).any()
    for k in range(9):
        assert np.array_equal(g.psi(k), o.psi(k)), k


@pytest.mark.parametrize("fdorder", [2, 4, 6, 8, 10, 12])
def test_forward_equals_oracle(fdorder):
    print(f"[acoustic] FD{fdorder}", flush=True)
    case = ac_case(fdorder=fdorder)
    g = gpu_run(*case)
    o = oracle_run(*case)
    assert_equal_runs(g, o)
    g.close()


@pytest.mark.parametrize("free_surf", [0, 1])
@pytest.mark.parametrize("seismo", [1, 2, 3, 4])
def test_boundaries_and_receiver_modes(free_surf, seismo):
    """CPML on four sides (FREE_SURF 0) or three with p = 0 on row 1;
    SEISMO 1 vx, vy; 2 p; 3 div, curl (both scaled by sqrt(rho vp^2),
    ac.c:199-201); 4 all five."""
    print(f"[acoustic] free_surf {free_surf} seismo {seismo}", flush=True)
    case = ac_case(free_surf=free_surf, seismo=seismo, nt=260)
    g = gpu_run(*case)
    o = oracle_run(*case)
    assert_equal_runs(g, o, seismo)
    if free_surf:
        nd = g.nd
        f = g.field(dnhip.F_SXX)
        assert not f[nd, nd:nd + case[0].nx].any()  # row 1
    g.close()


@pytest.mark.parametrize("stype", [1, 2, 3, 4])
def test_source_types(stype):
    """The explosive source on p (psource_AC.c) and the forces in x, y and
    rotated (update_v_PML_AC.c:640-656)."""
    print(f"[acoustic] source type {stype}", flush=True)
    case = ac_case(stype=stype, nt=240)
    g = gpu_run(*case)
    o = oracle_run(*case)
    assert_equal_runs(g, o)
    g.close()


def test_moment_tensor_source_injects_nothing():
    """QUELLTYP 5 in the acoustic loop: ac.c calls neither psource_AC nor a
    moment-tensor routine, so the run stays at rest."""
    case = ac_case(stype=5, nt=120)
    g = gpu_run(*case)
    o = oracle_run(*case)
    assert_equal_runs(g, o, nonzero=False)
    for c in range(5):
        assert not g.seismogram(c).any(), c
    g.close()


def test_modulus_parametrisation():
    """INVMAT1 3: the model holds the modulus rho vp^2 (ac.c:199-201 and
    update_s_acoustic_PML_AC.c use it directly)."""
    case = ac_case(invmat1=3, nt=260)
    g = gpu_run(*case)
    o = oracle_run(*case)
    assert_equal_runs(g, o)
    g.close()


@pytest.mark.parametrize("sy,stype", [(10.0, 1), (20.0, 1), (10.0, 3)])
def test_sources_at_the_free_surface(sy, stype):
    """Sources on row 1 and row 2 under the free surface.  An explosive
    source on row 1 is added to p and cleared by the surface condition in
    the same step (ac.c:160-178), so that run stays at rest; a force on
    row 1 and an explosive source on row 2 radiate."""
    print(f"[acoustic] source row {sy / 10.0:.0f} type {stype}", flush=True)
    case = ac_case(sy=sy, stype=stype, nt=200)
    g = gpu_run(*case)
    o = oracle_run(*case)
    at_rest = sy == 10.0 and stype == 1
    assert_equal_runs(g, o, nonzero=not at_rest)
    if at_rest:
        assert not g.seismogram(2).any()
    g.close()


SCALAR = [(64, 4, 1, 1), (128, 2, 1, 1), (256, 1, 1, 1), (64, 4, 2, 1), (128, 2, 4, 1), (64, 8, 2, 1)]


def test_tile_shapes_do_not_change_results():
    """Every scalar shape in both sweeps (the acoustic sweeps have scalar
    columns only)."""
    case = ac_case(nx=150, ny=90, nt=160)
    o = oracle_run(*case)
    for k, ts in enumerate(SCALAR):
        tv = SCALAR[(k + 2) % len(SCALAR)]
        print(f"[acoustic] tiles v{tv} s{ts}", flush=True)
        g = gpu_run(*case, tiles=(tv, ts))
        assert_equal_runs(g, o)
        g.close()


def test_vector_tiles_are_refused_and_autotune_keeps_scalar():
    case = ac_case(nt=60)
    g = dnhip.Context(case[0])
    g.set_model(case[1][0], None, case[1][1])
    g.set_sources(case[2], case[3])
    g.set_receivers(case[4])
    for sweep in (0, 1):
        with pytest.raises(dnhip.DnError, match="scalar"):
            g.set_tile(64, 4, 1, 4, sweep=sweep)
    g.autotune(10)
    assert g.tile(0)[3] == 1 and g.tile(1)[3] == 1
    g.forward()
    o = oracle_run(*case)
    assert_equal_runs(g, o)
    g.close()


def test_error_against_fp64_oracle():
    case = ac_case(nx=120, ny=80, nt=400)
    g = gpu_run(*case)
    o = oracle_run(*case, prec="f64")
    for c in range(4):
        err = rel_l2(g.seismogram(c), o.seismogram(c))
        print(f"[acoustic] comp {c} rel-L2 vs FP64 oracle {err:.2e}", flush=True)
        assert err <= SEIS_TOL
    # the curl of an acoustic wavefield is a difference of nearly equal
    # terms (vxy - vyx, non-zero only through the density gradient and the
    # staggering): its error is measured against the divergence's scale
    a, b, div = g.seismogram(4), o.seismogram(4), o.seismogram(3)
    err = float(np.linalg.norm((a.astype(np.float64) - b).ravel()) / np.linalg.norm(div.ravel()))
    print(f"[acoustic] curl error / |div| vs FP64 oracle {err:.2e}, "
          f"|curl| / |div| {np.linalg.norm(b) / np.linalg.norm(div):.2e}", flush=True)
    assert err <= SEIS_TOL
    g.close()


def test_errors_of_the_acoustic_entry_points():
    p, m, src, sig, rec = ac_case(nt=40)
    fast = dnhip.Params(**{**p.__dict__, "dt": p.dt * 4})
    h = dnhip.Context(fast)
    with pytest.raises(dnhip.DnError) as ei:  # checkfd_acoustic
        h.set_model(m[0], None, m[1])
    assert ei.value.code == -6
    h.close()
    g = dnhip.Context(p)
    with pytest.raises(dnhip.DnError):
        g.set_model(None, None, m[1])
    g.set_model(m[0], None, m[1])
    e = dnhip.Engine()
    with pytest.raises(dnhip.DnError, match="acoustic"):
        e.attach(g)
    e.close()
    g.close()
    with pytest.raises(dnhip.DnError, match="L = 0"):
        dnhip.Context(dnhip.Params(**{**p.__dict__, "L": 2}))


@pytest.mark.parametrize("invmat1", [1, 3])
def test_cli_forward_writes_the_library_seismograms(tmp_path, invmat1):
    tmp = str(tmp_path)
    nx, ny, dh, nt = 110, 70, 10.0, 280
    vp, _, rho = synth.layered_model(nx, ny, dh, 13, sigma=3.0)
    dt = float(np.float32(synth.stable_dt(float(vp.max()), dh, 8, safety=0.6)))
    if invmat1 == 3:
        vp = (rho.astype(np.float64) * vp.astype(np.float64) ** 2).astype(np.float32)
    for d in ("model", "su"):
        os.makedirs(os.path.join(tmp, d))
    first = "vp" if invmat1 == 1 else "lam"
    np.ascontiguousarray(vp, np.float32).tofile(os.path.join(tmp, f"model/m.{first}"))
    np.ascontiguousarray(rho, np.float32).tofile(os.path.join(tmp, "model/m.rho"))
    with open(os.path.join(tmp, "source.dat"), "w") as f:
        f.write("2\n400.00 0.0 30.00 0.0 15.00 1.00 0.0 1\n800.00 0.0 20.00 0.0 12.00 2.00 0.0 1\n")
    with open(os.path.join(tmp, "receiver.dat"), "w") as f:
        for k in range(50):
            f.write(f"{50.0 + 20.0 * k:.3f}\t20.000\n")
    inpfile.write_inp(os.path.join(tmp, "m.inp"), PHYSICS=2, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      SEISMO=4, L=2, FL="3.0 30.0", INVMAT1=invmat1)
    print(f"[acoustic cli] INVMAT1 {invmat1} in {tmp}", flush=True)
    r = subprocess.run([BIN, "m.inp"], cwd=tmp, capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr
    # readmod_AC.c:82-92 writes the model copies whatever WRITEMOD is
    for ext, a in (("pi", vp), ("rho", rho)):
        for name in (f"model/m.fdveps.{ext}", f"model/m.fdveps.{ext}.0.0"):
            got = np.fromfile(os.path.join(tmp, name), np.float32)
            assert np.array_equal(got, a.ravel()), name
    par = dnio.read_par(os.path.join(tmp, "m.inp"))
    cwd = os.getcwd()
    os.chdir(tmp)
    try:
        src, ts = dnio.read_sources(par.get("source_file"))
        rec = dnio.read_receivers(par)
    finally:
        os.chdir(cwd)
    p = dnhip.Params(nx=nx, ny=ny, nt=par.nt, dh=dh, dt=par.dt, fdorder=8, maxrelerror=par.maxrelerror, fw=par.fw,
                     free_surf=par.free_surf, seismo=4, ts=ts, physics=2, invmat1=invmat1, damping=par.damping,
                     fpml=par.fpml, npower=par.npower, k_max_pml=par.k_max_pml)
    g = dnhip.Context(p)
    g.set_model(vp, None, rho)
    # SEISMO 4 writes vx, vy and p (saveseis_glob.c:74-86)
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


def _adjoint(runner, obs):
    runner.forward()
    runner.residuals(obs[0], obs[1], obs[2], 1)
    runner.adjoint()
    raw = [runner.gradient(k) for k in range(3)]
    runner.assemble()
    return raw, [runner.gradient(k) for k in range(3, 6)]


@pytest.mark.parametrize("grad_form,invmat1,dtinv,quelltypb",
                         [(1, 1, 2, 1), (2, 1, 1, 1), (1, 3, 3, 4), (2, 3, 2, 4), (1, 1, 1, 2)])
def test_rtm_gradients_equal_oracle(grad_form, invmat1, dtinv, quelltypb):
    """The acoustic adjoint pass (src/AC/ac.c mode 1) and ass_gradAC: the
    raw correlations and the assembled vp (or lambda) and density gradients
    equal the FP32 oracle's bit for bit."""
    print(f"[acoustic rtm] GRAD_FORM {grad_form} INVMAT1 {invmat1} DTINV {dtinv} QUELLTYPB {quelltypb}", flush=True)
    p, m, src, sig, rec = ac_case(nt=240, seismo=4, invmat1=invmat1)
    p = dnhip.Params(**{**p.__dict__, "mode": 1, "grad_form": grad_form, "dtinv": dtinv, "quelltypb": quelltypb})
    g = dnhip.Context(p)
    g.set_model(m[0], None, m[1])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    o = orc.Oracle(p, m[0], None, m[1], prec="f32", aniso="ac")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    obs = (0.8 * o.seismogram(0), 0.9 * o.seismogram(1), 1.1 * o.seismogram(2))
    graw, ggrad = _adjoint(g, obs)
    oraw, ograd = _adjoint(o, obs)
    for k in range(3):
        assert np.array_equal(graw[k], oraw[k]), ("raw", k, rel_l2(graw[k], oraw[k]))
        assert np.array_equal(ggrad[k], ograd[k]), ("grad", k, rel_l2(ggrad[k], ograd[k]))
    assert np.abs(graw[0]).max() > 0 and np.abs(ggrad[0]).max() > 0 and np.abs(ggrad[2]).max() > 0
    assert not graw[1].any() and not ggrad[1].any()
    g.close()


def test_cli_rtm_image_equals_the_library(tmp_path):
    """MODE 2 with PHYSICS 2 (RTM_AC.c): the executable's P image is the
    stacked ass_gradAC vp gradient of the library's adjoint passes with the
    time-integrated observed data as adjoint sources, bit for bit; no S
    image is written (RTM_AC_out.c)."""
    import fwi_oracle as fo
    tmp = str(tmp_path)
    nx, ny, dh, nt = 100, 64, 10.0, 360
    vp, _, rho = synth.layered_model(nx, ny, dh, 27, sigma=3.0)
    dt = float(np.float32(synth.stable_dt(float(vp.max()) * 1.2, dh, 8, safety=0.6)))
    for d in ("model", "su", "obs", "jac"):
        os.makedirs(os.path.join(tmp, d))
    vt = vp.copy()
    vt[30:70, 30:34] *= 1.2
    for name, v in (("true", vt), ("mig", vp)):
        np.ascontiguousarray(v, np.float32).tofile(os.path.join(tmp, f"model/{name}.vp"))
        np.ascontiguousarray(rho, np.float32).tofile(os.path.join(tmp, f"model/{name}.rho"))
    with open(os.path.join(tmp, "source.dat"), "w") as f:
        f.write("2\n300.00 0.0 30.00 0.0 15.00 1.00 0.0 1\n650.00 0.0 30.00 0.0 15.00 1.00 0.0 1\n")
    with open(os.path.join(tmp, "receiver.dat"), "w") as f:
        for k in range(40):
            f.write(f"{40.0 + 20.0 * k:.3f}\t20.000\n")
    inpfile.write_inp(os.path.join(tmp, "true.inp"), PHYSICS=2, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      MFILE="model/true", SEIS_FILE_VX="obs/d_x.su", SEIS_FILE_VY="obs/d_y.su")
    r = subprocess.run([BIN, "true.inp"], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    inpfile.write_inp(os.path.join(tmp, "rtm.inp"), PHYSICS=2, MODE=2, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      MFILE="model/mig", DTINV=2, RTM_SHOT=0, TESTSHOT="1, 2, 1")
    inpfile.write_workflow(os.path.join(tmp, "wf.inp"))
    print(f"[acoustic cli rtm] in {tmp}", flush=True)
    r = subprocess.run([BIN, "rtm.inp", "wf.inp"], cwd=tmp, capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr
    assert not os.path.exists(os.path.join(tmp, "jac/j_S_image"))
    par = dnio.read_par(os.path.join(tmp, "rtm.inp"))
    cwd = os.getcwd()
    os.chdir(tmp)
    try:
        src, ts = dnio.read_sources(par.get("source_file"))
        rec = dnio.read_receivers(par)
    finally:
        os.chdir(cwd)
    p = dnhip.Params(nx=nx, ny=ny, nt=par.nt, dh=dh, dt=par.dt, fdorder=8, maxrelerror=par.maxrelerror, fw=par.fw,
                     free_surf=par.free_surf, seismo=1, ts=ts, physics=2, mode=1, dtinv=2, grad_form=par.grad_form,
                     damping=par.damping, fpml=par.fpml, npower=par.npower, k_max_pml=par.k_max_pml)
    g = dnhip.Context(p)
    g.set_model(vp, None, rho)
    total = np.zeros((nx, ny), np.float32)
    for ishot in (1, 2):
        loc = dnio.split_sources(par, src, ishot - 1, 1)
        g.set_sources(loc, dnio.wavelet(par, loc, ishot, ts))
        g.set_receivers(rec[:2])
        g.forward()
        obs = [dnio.read_su(os.path.join(tmp, f"obs/d_{c}.su.shot{ishot}"), rec.shape[1], par.nt) for c in "xy"]
        diffs = [fo.rtm_adjoint_source(obs[c], g.seismogram(c), par.dt, par.grad_form, 1)[0] for c in (0, 1)]
        g.set_adjoint_sources(diffs[0], diffs[1], None)
        g.adjoint()
        g.assemble()
        total += g.gradient(3)
    g.close()
    got = np.fromfile(os.path.join(tmp, "jac/j_P_image"), np.float32).reshape(nx, ny)
    assert np.array_equal(got, total)
    assert np.abs(total).max() > 0
