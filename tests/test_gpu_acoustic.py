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


@pytest.mark.parametrize("variant", ["plain", "precond_pcg", "spatfilter4", "pressure"])
def test_cli_fwi_iterations_equal_python_iterations(tmp_path, variant):
    """MODE 1 with PHYSICS 2 (FWI_AC.c, grad_obj_ac.c, step_length_est_ac.c,
    calc_mat_change_test_AC.c): three iterations of the executable against
    the same iterations in Python, with the library's acoustic gradients
    (ass_gradAC, their parity with the FP32 oracle tested above) and the
    numpy restatement's two-class L-BFGS / PCG, step search and model update
    of vp and rho.  The vp and rho models after every iteration, the final
    INV_MODELFILE_vp/_rho.bin, the misfit log and the L-BFGS files equal the
    Python run's bit for bit; no vs model or vs file is written; the misfit
    falls.  plain: L-BFGS, vx and vy residuals.  precond_pcg: EPRECOND 3
    (energy of vx and vy, ac.c:251-253) and the horizontal taper with PCG
    (PCG_BETA 2).  spatfilter4: the velocity-dependent Gaussian with the vp
    model (FWI_AC.c:676-677).  pressure: QUELLTYPB 4 with SEISMO 2 (the
    pressure residual added to p, psource_AC.c:33-43): on this model the
    reference's direction raises the misfit at every trial step, so the
    search fails and the stage ends with the model unchanged, in both runs
    alike (equal logs and files)."""
    import fwi_oracle as fo
    from test_gpu_cli import _files, _run
    itermax = 3
    pcg = variant == "precond_pcg"
    precond = pcg
    spat4 = variant == "spatfilter4"
    qb, seismo = (4, 2) if variant == "pressure" else (1, 1)
    tmp = str(tmp_path)
    os.makedirs(os.path.join(tmp, "taper"), exist_ok=True)
    nx, ny, dh, dt, nt = 96, 64, 10.0, 1e-3, 400
    vp0, vs0, rho0 = synth.layered_model(nx, ny, dh, seed=5, sigma=3.0)
    vt, rt = vp0.copy(), rho0.copy()
    vt[40:56, 10:20] *= 1.08
    rt[40:56, 10:20] *= 1.05
    sources = [(300.0, 30.0, 15.0, 1.0, 1), (650.0, 30.0, 15.0, 1.0, 1)]
    receivers = [(40.0 + 20.0 * k, 20.0) for k in range(40)]
    _files(tmp, nx, ny, dh, (vt, vs0, rt), sources, receivers, mname="model/true")
    inpfile.write_inp(os.path.join(tmp, "true.inp"), PHYSICS=2, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      MFILE="model/true", SEIS_FILE_VX="obs/d_x.su", SEIS_FILE_VY="obs/d_y.su",
                      SEIS_FILE_P="obs/d_p.su", SEISMO=seismo)
    _run(tmp, "true.inp")
    _files(tmp, nx, ny, dh, (vp0, vs0, rho0), sources, receivers, mname="model/start")
    inpfile.write_inp(os.path.join(tmp, "fwi.inp"), PHYSICS=2, MODE=1, QUELLTYPB=qb, SEISMO=seismo, NX=nx, NY=ny,
                      DH=dh, DT=dt, TIME=nt * dt, MFILE="model/start", DTINV=2, TESTSHOT="1, 2, 1",
                      ITERMAX=itermax, INV_MOD_OUT=1, MISFIT_LOG_FILE="misfit.dat",
                      SWS_TAPER_GRAD_HOR=int(precond), GRADT="3, 12, 60, 62", EXP_TAPER_GRAD_HOR=2.0,
                      GRAD_METHOD=1 if pcg else 2, PCG_BETA=2)
    fce = 10.0 if spat4 else 0.0
    inpfile.write_workflow(os.path.join(tmp, "wf.inp"), fc_end=fce, eprecond=3 if precond else 0,
                           spatfilter=4 if spat4 else 0, wd_damp=0.5, wd_damp1=0.3)
    out = _run(tmp, "fwi.inp", "wf.inp")
    assert "rho_avg" in out and "Vs_avg" not in out

    par = dnio.read_par(os.path.join(tmp, "fwi.inp"))
    cwd = os.getcwd()
    os.chdir(tmp)
    try:
        src, ts = dnio.read_sources(par.get("source_file"))
        rec = dnio.read_receivers(par)
    finally:
        os.chdir(cwd)
    shots = []
    for ishot in (1, 2):
        loc = dnio.split_sources(par, src, ishot - 1, 1)
        shots.append((loc, dnio.wavelet(par, loc, ishot, ts)))
    p = dnhip.Params(nx=nx, ny=ny, nt=par.nt, dh=dh, dt=par.dt, fdorder=par.fdorder, maxrelerror=par.maxrelerror,
                     fw=par.fw, free_surf=par.free_surf, seismo=seismo, ts=ts, physics=2, mode=1, dtinv=2,
                     grad_form=par.grad_form, quelltypb=qb, damping=par.damping, fpml=par.fpml,
                     npower=par.npower, k_max_pml=par.k_max_pml)
    st = dnio.read_stage(os.path.join(tmp, "wf.inp"), 1)
    upd = dict(vpupperlim=par.vpupperlim, vplowerlim=par.vplowerlim, vsupperlim=par.vsupperlim,
               vslowerlim=par.vslowerlim, rhoupperlim=par.rhoupperlim, rholowerlim=par.rholowerlim,
               scalerho=st.scalerho, inv_vp_iter=st.inv_vp_iter, inv_vs_iter=st.inv_vs_iter,
               inv_rho_iter=st.inv_rho_iter, invmat1=par.invmat1)
    comps = "p" if qb == 4 else "xy"
    obs = {i: [dnio.read_su(os.path.join(tmp, f"obs/d_{c}.su.shot{i}"), rec.shape[1], nt) for c in comps]
           for i in (1, 2)}
    g = dnhip.Context(p)
    if precond:
        g.set_energy_weights(3)

    def shot(i, vp, rho):
        g.set_model(vp, None, rho)
        g.set_sources(*shots[i - 1])
        g.set_receivers(rec[:2])
        g.forward()

    def residuals(i):
        if qb == 4:
            return g.residuals(None, None, obs[i][0], 1)
        return g.residuals(obs[i][0], obs[i][1], None, 1)

    def misfit(tv, ts_, tr, itest):
        assert ts_ is None
        l2 = 0.0
        for i in (1, 2):
            shot(i, tv, tr)
            l2 += residuals(i)
        return l2

    vp, rho = vp0.copy(), rho0.copy()
    lb = fo.PCG(nx, ny) if pcg else fo.LBFGS(nx, ny, par.nlbfgs, nclass=2)
    c = (fo.model_mean(vp), np.float32(0.0), fo.model_mean(rho))
    carry = ([0.0, 0.0, 0.0], [0.0, 0.0, 0.0])
    lines, l2s = [], []
    for it in range(1, itermax + 1):
        grads = [np.zeros((nx, ny), np.float32), None, np.zeros((nx, ny), np.float32)]
        l2sum = 0.0
        for i in (1, 2):
            shot(i, vp, rho)
            l2sum += residuals(i)
            g.adjoint()
            g.assemble()
            gs = [g.gradient(3), None, g.gradient(5)]
            if precond:
                _, gv, _, gr = fo.energy_precond(3, par.dh, g.energy(0), None, *c, gs[0], None, gs[2])
                gs = [gv, None, gr]
            for k in (0, 2):
                grads[k] += gs[k]
        if spat4:
            grads = [fo.gauss_filter_var(a, vp, par.dh, 0.5, 0.3, fce) if a is not None else None for a in grads]
        if precond:
            coeff = fo.taper_hor(nx, ny, par.dh, 3, 12, 2.0)
            grads = [(a * coeff).astype(np.float32) if a is not None else None for a in grads]
        l2s.append(l2sum)
        d = lb.direction(2, *grads)[:3] if pcg else lb.direction(it, *c, *grads, vp, None, rho)
        assert d[1] is None
        if not pcg:
            saved = (lb.mold.copy(), lb.gold.copy(), [None if a is None else a.copy() for a in d])
        ls = fo.step_length(upd, it, par.eps_scale, par.scalefac, par.stepmax, l2sum, vp, None, rho, *d, misfit,
                            carry=carry)
        carry = (ls["epst"], ls["L2t"])
        eps = ls["eps_scale"]
        if it > par.min_iter and ls["step1"] == 0:
            eps = np.float32(0.0)
        lines.append("%e \t %e \t %e \t %e \t %e \t %e \t %e \t %e \t %d " % (eps, *ls["epst"], *ls["L2t"], l2sum, 1))
        vp, _, rho, _ = fo.mat_change_test(upd, it, eps, vp, None, rho, *d)
        print(f"[acoustic fwi] iteration {it}: misfit {l2sum:.6e}, step {float(eps):.4e}, "
              f"{ls['nevals']} test forward runs", flush=True)
        if it > par.min_iter and ((it > 2 and abs((l2s[it - 3] - l2s[it - 1]) / l2s[it - 3]) <= st.pro)
                                  or ls["step3"] == 1):
            break
        for name, a in (("vp", vp), ("rho", rho)):
            got = np.fromfile(os.path.join(tmp, f"model/inv_{name}_stage_1_it_{it}.bin"), np.float32)
            assert np.array_equal(got.reshape(nx, ny), a), (it, name)
        assert not os.path.exists(os.path.join(tmp, f"model/inv_vs_stage_1_it_{it}.bin"))
    g.close()
    for name, a in (("vp", vp), ("rho", rho)):
        got = np.fromfile(os.path.join(tmp, f"model/inv_{name}.bin"), np.float32).reshape(nx, ny)
        assert np.array_equal(got, a), name
    assert not os.path.exists(os.path.join(tmp, "model/inv_vs.bin"))
    assert not os.path.exists(os.path.join(tmp, "jac/j_grad_vs"))
    log = open(os.path.join(tmp, "misfit.dat")).read().splitlines()
    assert log == lines
    jac = os.path.join(tmp, "jac/j")
    if pcg:
        assert not os.path.exists(f"{jac}_p.old")
    else:
        mold, gold, dd = saved
        nc = nx * ny
        for names, slot, dk in ((("p_vp", "p", "c"), 1, dd[0]), (("p_mrho", "p_rho", "c_rho"), 0, dd[2])):
            want = (mold[slot * nc:(slot + 1) * nc], gold[slot * nc:(slot + 1) * nc], dk.ravel())
            for n, w in zip(names, want):
                assert np.array_equal(np.fromfile(f"{jac}_{n}.old", np.float32), w), n
        for n in ("p_vs", "p_u", "c_u"):
            assert not os.path.exists(f"{jac}_{n}.old"), n
    if variant == "pressure" and len(l2s) == 1 and ls["step3"] == 1:
        assert float(eps) == 0.0 and np.array_equal(vp, vp0) and ls["nevals"] > 1
        return
    assert l2s[-1] < l2s[0]
    assert not np.array_equal(vp, vp0)
