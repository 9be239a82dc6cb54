"""SH (PHYSICS 5) forward modelling on the GPU against the FP32 oracle
(src/SH/sh.c with update_v_PML_SH.c, update_s_elastic_PML_SH.c,
av_mu_SH.c, inv_rho_SH.c, PML_pro_SH.c): the vz seismograms, the final vz,
sxz, syz and the CPML memory bit for bit for every FD order, with and
without the FREE_SURF switch, every source type, both INVMAT1
parametrisations and every scalar tile shape; the error against the FP64
oracle stays within the seismogram tolerance (rel-L2 1e-5).  The drop-in
executable writes the library's vz seismograms (SEIS_FILE_VY) and the
reference's model copies for PHYSICS 5 parameter files.  The SH gradient
(sh.c mode 1, ass_gradSH) equals the oracle bit for bit, and the
executable's SH FWI (MODE 1) equals a Python re-run of its iterations."""
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


def sh_case(nx=96, ny=64, nt=300, fdorder=8, free_surf=1, fw=10, seismo=1, stype=1, seed=8, invmat1=1,
            sy=40.0, mode=0, dtinv=1, grad_form=1, quelltypb=1):
    _, vs, rho = synth.layered_model(nx, ny, 10.0, seed, sigma=3.0)
    dt = float(np.float32(synth.stable_dt(float(vs.max()), 10.0, fdorder, safety=0.6)))
    p = dnhip.Params(nx=nx, ny=ny, nt=nt, dh=10.0, dt=dt, fdorder=fdorder, maxrelerror=1, fw=fw,
                     free_surf=free_surf, seismo=seismo, ts=0.05, physics=5, invmat1=invmat1, mode=mode,
                     dtinv=dtinv, grad_form=grad_form, quelltypb=quelltypb)
    if invmat1 == 3:  # the modulus rho vs^2 in place of vs
        vs = (rho.astype(np.float64) * vs.astype(np.float64) ** 2).astype(np.float32)
    src = synth.source_table(10.0 * (nx // 2), sy, 10.0, fc=20.0, stype=stype, azimuth=30.0)
    sig = synth.ricker(nt, dt, 20.0, amp=1e3)[None, :]
    rec = synth.receiver_line(30.0, 20.0, nx // 2 - 4, 30.0, 10.0)
    return p, (vs, rho), src, sig, rec


def gpu_run(p, m, src, sig, rec, tiles=None):
    g = dnhip.Context(p)
    g.set_model(None, m[0], m[1])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    if tiles:
        g.set_tile(*tiles[0], sweep=0)
        g.set_tile(*tiles[1], sweep=1)
    g.forward()
    return g


def oracle_run(p, m, src, sig, rec, prec="f32"):
    o = orc.Oracle(p, None, m[0], m[1], prec=prec, aniso="sh")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    return o


def assert_equal_runs(g, o, nonzero=True):
    for c in range(5):
        a, b = g.seismogram(c), o.seismogram(c)
        assert np.array_equal(a, b), (c, rel_l2(a, b))
    if nonzero and g.p.seismo:
        assert np.abs(o.seismogram(1)).max() > 0
    # vz, sxz, syz in the vx, sxx, syy planes
    for fid in (dnhip.F_VX, dnhip.F_SXX, dnhip.F_SYY):
        a, b = g.field(fid), o.field(fid)
        assert np.array_equal(a, b), (fid, rel_l2(a, b))
        if nonzero:
            assert np.abs(b).max() > 0, fid
    for k in range(9):
        assert np.array_equal(g.psi(k), o.psi(k)), k


@pytest.mark.parametrize("fdorder", [2, 4, 6, 8, 10, 12])
def test_forward_equals_oracle(fdorder):
    print(f"[sh] FD{fdorder}", flush=True)
    case = sh_case(fdorder=fdorder)
    g = gpu_run(*case)
    o = oracle_run(*case)
    assert_equal_runs(g, o)
    g.close()


@pytest.mark.parametrize("free_surf,seismo", [(0, 1), (1, 1), (0, 0)])
def test_boundaries_and_receivers(free_surf, seismo):
    """CPML on four sides (FREE_SURF 0) or three (the SH loop has no
    free-surface routine); SEISMO 1 records vz in component 1."""
    print(f"[sh] free_surf {free_surf} seismo {seismo}", flush=True)
    case = sh_case(free_surf=free_surf, seismo=seismo, nt=260)
    g = gpu_run(*case)
    o = oracle_run(*case)
    assert_equal_runs(g, o)
    if free_surf:
        nd = g.nd
        assert not g.field(dnhip.F_SYY)[:nd].any()  # ghost rows stay 0
    g.close()


@pytest.mark.parametrize("stype", [1, 2, 3, 4, 5])
def test_source_types(stype):
    """QUELLTYP 1 is the force on vz (update_v_PML_SH.c:560-568); the other
    types inject nothing in the SH loop."""
    print(f"[sh] source type {stype}", flush=True)
    case = sh_case(stype=stype, nt=200)
    g = gpu_run(*case)
    o = oracle_run(*case)
    assert_equal_runs(g, o, nonzero=stype == 1)
    if stype != 1:
        assert not g.field(dnhip.F_VX).any()
    g.close()


def test_modulus_parametrisation_and_surface_row_source():
    """INVMAT1 3 (the model holds mu) with the source on row 1."""
    case = sh_case(invmat1=3, sy=10.0, nt=260)
    g = gpu_run(*case)
    o = oracle_run(*case)
    assert_equal_runs(g, o)
    g.close()


SCALAR = [(64, 4, 1, 1), (128, 2, 1, 1), (256, 1, 1, 1), (64, 4, 2, 1), (128, 2, 4, 1), (64, 8, 2, 1)]


def test_tile_shapes_do_not_change_results():
    """Every scalar shape in both sweeps (the SH sweeps have scalar columns
    only)."""
    case = sh_case(nx=150, ny=90, nt=160)
    o = oracle_run(*case)
    for k, ts in enumerate(SCALAR):
        tv = SCALAR[(k + 3) % len(SCALAR)]
        print(f"[sh] tiles v{tv} s{ts}", flush=True)
        g = gpu_run(*case, tiles=(tv, ts))
        assert_equal_runs(g, o)
        g.close()


def test_vector_tiles_are_refused_and_autotune_keeps_scalar():
    case = sh_case(nt=60)
    g = dnhip.Context(case[0])
    g.set_model(None, case[1][0], case[1][1])
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
    case = sh_case(nx=120, ny=80, nt=400)
    g = gpu_run(*case)
    o = oracle_run(*case, prec="f64")
    err = rel_l2(g.seismogram(1), o.seismogram(1))
    print(f"[sh] vz rel-L2 vs FP64 oracle {err:.2e}", flush=True)
    assert err <= SEIS_TOL
    g.close()


def test_errors_of_the_sh_entry_points():
    p, m, src, sig, rec = sh_case(nt=40)
    fast = dnhip.Params(**{**p.__dict__, "dt": p.dt * 4})
    h = dnhip.Context(fast)
    with pytest.raises(dnhip.DnError) as ei:  # checkfd_elast_SH
        h.set_model(None, m[0], m[1])
    assert ei.value.code == -6
    h.close()
    g = dnhip.Context(p)
    with pytest.raises(dnhip.DnError):
        g.set_model(None, None, m[1])
    with pytest.raises(dnhip.DnError, match="VTI"):
        g.set_model_aniso(m[1], m[0], m[0], m[0], m[0])
    g.set_model(None, m[0], m[1])
    e = dnhip.Engine()
    with pytest.raises(dnhip.DnError, match="SH"):
        e.attach(g)
    e.close()
    g.close()
    for bad in ({"L": 2}, {"seismo": 2}):
        with pytest.raises(dnhip.DnError):
            dnhip.Context(dnhip.Params(**{**p.__dict__, **bad}))


@pytest.mark.parametrize("invmat1", [1, 3])
def test_cli_forward_writes_the_library_seismograms(tmp_path, invmat1):
    tmp = str(tmp_path)
    nx, ny, dh, nt = 110, 70, 10.0, 280
    _, vs, rho = synth.layered_model(nx, ny, dh, 19, sigma=3.0)
    dt = float(np.float32(synth.stable_dt(float(vs.max()), dh, 8, safety=0.6)))
    if invmat1 == 3:
        vs = (rho.astype(np.float64) * vs.astype(np.float64) ** 2).astype(np.float32)
    for d in ("model", "su"):
        os.makedirs(os.path.join(tmp, d))
    first = "vs" if invmat1 == 1 else "mu"
    np.ascontiguousarray(vs, np.float32).tofile(os.path.join(tmp, f"model/m.{first}"))
    np.ascontiguousarray(rho, np.float32).tofile(os.path.join(tmp, "model/m.rho"))
    with open(os.path.join(tmp, "source.dat"), "w") as f:
        f.write("2\n400.00 0.0 30.00 0.0 15.00 1.00 0.0 1\n800.00 0.0 20.00 0.0 12.00 2.00 0.0 1\n")
    with open(os.path.join(tmp, "receiver.dat"), "w") as f:
        for k in range(50):
            f.write(f"{50.0 + 20.0 * k:.3f}\t20.000\n")
    inpfile.write_inp(os.path.join(tmp, "m.inp"), PHYSICS=5, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      SEISMO=1, L=0, INVMAT1=invmat1, WRITEMOD=1)
    print(f"[sh cli] INVMAT1 {invmat1} in {tmp}", flush=True)
    r = subprocess.run([BIN, "m.inp"], cwd=tmp, capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0, r.stderr
    # readmod_elastic_SH.c:85-98 writes the model copies under WRITEMOD
    for ext, a in (("mu", vs), ("rho", rho)):
        for name in (f"model/m.denise.{ext}", f"model/m.denise.{ext}.0.0"):
            got = np.fromfile(os.path.join(tmp, name), np.float32)
            assert np.array_equal(got, a.ravel()), name
    # saveseis_glob_SH.c: vz goes to SEIS_FILE_VY only
    assert not os.path.exists(os.path.join(tmp, "su/out_x.su.shot1"))
    par = dnio.read_par(os.path.join(tmp, "m.inp"))
    cwd = os.getcwd()
    os.chdir(tmp)
    try:
        src, ts = dnio.read_sources(par.get("source_file"))
        rec = dnio.read_receivers(par)
    finally:
        os.chdir(cwd)
    p = dnhip.Params(nx=nx, ny=ny, nt=par.nt, dh=dh, dt=par.dt, fdorder=8, maxrelerror=par.maxrelerror, fw=par.fw,
                     free_surf=par.free_surf, seismo=1, ts=ts, physics=5, invmat1=invmat1, damping=par.damping,
                     fpml=par.fpml, npower=par.npower, k_max_pml=par.k_max_pml)
    g = dnhip.Context(p)
    g.set_model(None, vs, rho)
    for ishot in (1, 2):
        loc = dnio.split_sources(par, src, ishot - 1, 1)
        sig = dnio.wavelet(par, loc, ishot, ts)
        g.set_sources(loc, sig)
        g.set_receivers(rec[:2])
        g.forward()
        _, data = dnio.parse_su(os.path.join(tmp, f"su/out_y.su.shot{ishot}"))
        assert np.array_equal(data, g.seismogram(1)), ishot
        assert np.abs(data).max() > 0
    g.close()


def _gradient_pair(case, tiles=None, adjoint_sources=None):
    """Forward with the snapshot store, the vz residual against 0.8 x the
    synthetics (or the given adjoint sources), the adjoint run, assembly;
    GPU and FP32 oracle."""
    g = gpu_run(*case, tiles=tiles)
    o = oracle_run(*case)
    assert np.array_equal(g.seismogram(1), o.seismogram(1))
    if adjoint_sources is None:
        obs = (0.8 * o.seismogram(1)).astype(np.float32)
        lg = g.residuals(None, obs, None, 1)
        lo = o.residuals(None, obs, None, 1)
        assert abs(lg - lo) <= 1e-5 * abs(lo)
        assert np.array_equal(g.adjoint_sources(1), o.adjoint_sources(1))
    else:
        g.set_adjoint_sources(None, adjoint_sources, None)
        o.set_adjoint_sources(None, adjoint_sources, None)
    g.adjoint()
    o.adjoint()
    g.assemble()
    o.assemble()
    return g, o


@pytest.mark.parametrize("grad_form,invmat1,dtinv", [(1, 1, 1), (1, 3, 2), (2, 1, 3), (2, 3, 1), (1, 1, 2)])
def test_gradient_equals_oracle(grad_form, invmat1, dtinv):
    """sh.c mode 1 (the adjoint run with the density and shear-modulus
    correlations) and ass_gradSH: the raw sums and the vs / rho gradients
    equal the FP32 oracle bit for bit; the final adjoint fields too."""
    print(f"[sh grad] GRAD_FORM {grad_form} INVMAT1 {invmat1} DTINV {dtinv}", flush=True)
    case = sh_case(nt=320, mode=1, dtinv=dtinv, grad_form=grad_form, invmat1=invmat1, free_surf=0)
    if grad_form == 2:   # not integrated: the residual itself, time-reversed
        g0 = gpu_run(*case)
        adj = np.ascontiguousarray((0.2 * g0.seismogram(1))[:, ::-1], np.float32)
        g0.close()
        g, o = _gradient_pair(case, adjoint_sources=adj)
    else:
        g, o = _gradient_pair(case)
    for k in (1, 2, 4, 5):
        a, b = g.gradient(k), o.gradient(k)
        assert np.array_equal(a, b), (k, rel_l2(a, b))
        assert np.abs(b).max() > 0, k
    for fid in (dnhip.F_VX, dnhip.F_SXX, dnhip.F_SYY):
        assert np.array_equal(g.field(fid), o.field(fid)), fid
    g.close()


def test_gradient_tile_shapes_and_fd_orders():
    """Every scalar shape in both sweeps of the adjoint pass, FD2 and FD12."""
    for fdorder in (2, 12):
        case = sh_case(nx=130, ny=80, nt=220, fdorder=fdorder, mode=1, dtinv=2)
        for k, ts in enumerate(SCALAR):
            tv = SCALAR[(k + 2) % len(SCALAR)]
            print(f"[sh grad] FD{fdorder} tiles v{tv} s{ts}", flush=True)
            g, o = _gradient_pair(case, tiles=(tv, ts))
            for q in (1, 2, 4, 5):
                assert np.array_equal(g.gradient(q), o.gradient(q)), (fdorder, ts, q)
            g.close()


def test_gradient_error_against_fp64_oracle():
    case = sh_case(nt=400, mode=1, dtinv=2)
    g, _ = _gradient_pair(case)
    o = oracle_run(*case, prec="f64")
    obs = (0.8 * oracle_run(*case).seismogram(1)).astype(np.float32)
    o.residuals(None, obs, None, 1)
    o.adjoint()
    o.assemble()
    for k in (4, 5):
        err = rel_l2(g.gradient(k), o.gradient(k))
        print(f"[sh grad] gradient {k} rel-L2 vs FP64 oracle {err:.2e}", flush=True)
        assert err <= 1e-4
    g.close()


@pytest.mark.parametrize("variant", ["plain", "precond", "spatfilter3", "pcg"])
def test_cli_fwi_iterations_equal_python_iterations(tmp_path, variant):
    """MODE 1 with PHYSICS 5 (FWI_SH.c, grad_obj_sh.c, ass_gradSH.c,
    step_length_est_sh.c, calc_mat_change_test_SH.c): three iterations of
    the executable against the same iterations in Python, with the
    library's raw SH sums and the numpy restatement's stacking, ass_gradSH,
    SH L-BFGS ([vs | rho], normalised models), step search and update of vs
    and rho.  The vs and rho models after every iteration, INV_MODELFILE
    _vs/_rho.bin, the misfit log and the L-BFGS files equal the Python
    run's bit for bit; no vp file is written; the misfit falls.  plain:
    L-BFGS.  precond: EPRECOND 3 (energy of vz, eprecond_SH.c) divided out
    per shot, the horizontal taper (precond_SH) and GRAD_FILTER 1 (a 3 x 3
    median of the gradients) after it.  spatfilter3: the Gaussian gradient
    filter with the start model's mean vs.  pcg: conjugate gradients (PCG_BETA
    2) over the vs and rho classes, whose store_PCG_SH writes the descent
    directions -g as JACOBIAN_p_u.old and _p_rho.old."""
    import fwi_oracle as fo
    from test_gpu_cli import _files, _run
    itermax = 3
    precond = variant == "precond"
    spat3 = variant == "spatfilter3"
    pcg = variant == "pcg"
    tmp = str(tmp_path)
    os.makedirs(os.path.join(tmp, "taper"), exist_ok=True)
    nx, ny, dh, nt = 96, 64, 10.0, 400
    _, vs0, rho0 = synth.layered_model(nx, ny, dh, seed=6, sigma=3.0)
    dt = float(np.float32(synth.stable_dt(float(vs0.max()) * 1.2, dh, 8, safety=0.6)))
    vt, rt = vs0.copy(), rho0.copy()
    vt[40:56, 10:20] *= 1.08
    rt[40:56, 10:20] *= 1.05
    sources = [(300.0, 30.0, 15.0, 1.0, 1), (650.0, 30.0, 15.0, 1.0, 1)]
    receivers = [(40.0 + 20.0 * k, 20.0) for k in range(40)]
    _files(tmp, nx, ny, dh, (vt, vt, rt), sources, receivers, mname="model/true")
    inpfile.write_inp(os.path.join(tmp, "true.inp"), PHYSICS=5, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      MFILE="model/true", SEIS_FILE_VX="obs/d_x.su", SEIS_FILE_VY="obs/d_y.su", SEISMO=1)
    _run(tmp, "true.inp")
    _files(tmp, nx, ny, dh, (vs0, vs0, rho0), sources, receivers, mname="model/start")
    inpfile.write_inp(os.path.join(tmp, "fwi.inp"), PHYSICS=5, MODE=1, QUELLTYPB=1, SEISMO=1, NX=nx, NY=ny, DH=dh,
                      DT=dt, TIME=nt * dt, MFILE="model/start", DTINV=2, TESTSHOT="1, 2, 1", ITERMAX=itermax,
                      INV_MOD_OUT=1, MISFIT_LOG_FILE="misfit.dat", SWS_TAPER_GRAD_HOR=int(precond),
                      GRADT="3, 12, 60, 62", EXP_TAPER_GRAD_HOR=2.0, GRAD_METHOD=1 if pcg else 2, PCG_BETA=2,
                      GRAD_FILTER=int(precond),
                      FILT_SIZE_GRAD=3)
    fce = 10.0 if spat3 else 0.0
    inpfile.write_workflow(os.path.join(tmp, "wf.inp"), fc_end=fce, eprecond=3 if precond else 0,
                           spatfilter=3 if spat3 else 0, wd_damp=0.5)
    out = _run(tmp, "fwi.inp", "wf.inp")
    assert "Vs_avg" in out

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
                     fw=par.fw, free_surf=par.free_surf, seismo=1, ts=ts, physics=5, mode=1, dtinv=2,
                     grad_form=par.grad_form, quelltypb=1, damping=par.damping, fpml=par.fpml, npower=par.npower,
                     k_max_pml=par.k_max_pml)
    st = dnio.read_stage(os.path.join(tmp, "wf.inp"), 1)
    upd = dict(vpupperlim=par.vpupperlim, vplowerlim=par.vplowerlim, vsupperlim=par.vsupperlim,
               vslowerlim=par.vslowerlim, rhoupperlim=par.rhoupperlim, rholowerlim=par.rholowerlim,
               scalerho=st.scalerho, inv_vp_iter=st.inv_vp_iter, inv_vs_iter=st.inv_vs_iter,
               inv_rho_iter=st.inv_rho_iter, invmat1=par.invmat1)
    obs = {i: dnio.read_su(os.path.join(tmp, f"obs/d_y.su.shot{i}"), rec.shape[1], nt) for i in (1, 2)}
    g = dnhip.Context(p)
    if precond:
        g.set_energy_weights(3)

    def shot(i, vs, rho):
        g.set_model(None, vs, rho)
        g.set_sources(*shots[i - 1])
        g.set_receivers(rec[:2])
        g.forward()

    def misfit(tv, ts_, tr, itest):
        assert tv is None
        l2 = 0.0
        for i in (1, 2):
            shot(i, ts_, tr)
            l2 += g.residuals(None, obs[i], None, 1)
        return l2

    vs, rho = vs0.copy(), rho0.copy()
    lb = fo.PCG(nx, ny) if pcg else fo.LBFGS(nx, ny, par.nlbfgs, sh=True)
    c = (np.float32(0.0), fo.model_mean(vs), fo.model_mean(rho))
    carry = ([0.0, 0.0, 0.0], [0.0, 0.0, 0.0])
    lines, l2s = [], []
    for it in range(1, itermax + 1):
        gu, gr = np.zeros((nx, ny), np.float32), np.zeros((nx, ny), np.float32)
        l2sum = 0.0
        for i in (1, 2):
            shot(i, vs, rho)
            l2sum += g.residuals(None, obs[i], None, 1)
            g.adjoint()
            su, sr = g.gradient(1), g.gradient(2)
            if precond:
                _, _, su, sr = fo.energy_precond(3, par.dh, g.energy(0), None, *c, None, su, sr)
            gu += su
            gr += sr
        gu, gr = fo.assemble_sh(par.dt, par.invmat1, par.grad_form, it, st.inv_vs_iter, st.inv_rho_iter, vs, rho,
                                gu, gr)
        if precond:
            coeff = fo.taper_hor(nx, ny, par.dh, 3, 12, 2.0)
            gu, gr = (gu * coeff).astype(np.float32), (gr * coeff).astype(np.float32)
            gu, gr = fo.median_filter(gu, 3), fo.median_filter(gr, 3)
        if spat3:
            (gu, wu), (gr, _) = fo.gauss_filter(gu, par.dh, 0.5, float(c[1]), fce), fo.gauss_filter(gr, par.dh, 0.5,
                                                                                               float(c[1]), fce)
            assert wu >= 3
        l2s.append(l2sum)
        if pcg:
            descent = ((-gu).astype(np.float32), (-gr).astype(np.float32))
            d = lb.direction(2, None, gu, gr)[:3]
        else:
            d = lb.direction(it, *c, None, gu, gr, None, vs, rho)
            saved = (lb.mold.copy(), lb.gold.copy(), [None if a is None else a.copy() for a in d])
        ls = fo.step_length(upd, it, par.eps_scale, par.scalefac, par.stepmax, l2sum, None, vs, rho, *d, misfit,
                            carry=carry)
        carry = (ls["epst"], ls["L2t"])
        eps = ls["eps_scale"]
        if it > par.min_iter and ls["step1"] == 0:
            eps = np.float32(0.0)
        lines.append("%e \t %e \t %e \t %e \t %e \t %e \t %e \t %e \t %d " % (eps, *ls["epst"], *ls["L2t"], l2sum, 1))
        _, vs, rho, _ = fo.mat_change_test(upd, it, eps, None, vs, rho, *d)
        print(f"[sh fwi] iteration {it}: misfit {l2sum:.6e}, step {float(eps):.4e}, {ls['nevals']} test runs",
              flush=True)
        if it > par.min_iter and ((it > 2 and abs((l2s[it - 3] - l2s[it - 1]) / l2s[it - 3]) <= st.pro)
                                  or ls["step3"] == 1):
            break
        for name, a in (("vs", vs), ("rho", rho)):
            got = np.fromfile(os.path.join(tmp, f"model/inv_{name}_stage_1_it_{it}.bin"), np.float32)
            assert np.array_equal(got.reshape(nx, ny), a), (it, name)
        assert not os.path.exists(os.path.join(tmp, f"model/inv_vp_stage_1_it_{it}.bin"))
    g.close()
    for name, a in (("vs", vs), ("rho", rho)):
        got = np.fromfile(os.path.join(tmp, f"model/inv_{name}.bin"), np.float32).reshape(nx, ny)
        assert np.array_equal(got, a), name
    assert not os.path.exists(os.path.join(tmp, "model/inv_vp.bin"))
    assert open(os.path.join(tmp, "misfit.dat")).read().splitlines() == lines
    jac = os.path.join(tmp, "jac/j")
    if pcg:
        for n, w in (("p_u", descent[0]), ("p_rho", descent[1])):
            for sfx in (".old", ".old.0.0"):
                assert np.array_equal(np.fromfile(f"{jac}_{n}{sfx}", np.float32), w.ravel()), n + sfx
        assert not os.path.exists(f"{jac}_c_u.old")
    else:
        mold, gold, dd = saved
        nc = nx * ny
        for names, slot, dk in ((("p_vs", "p_u", "c_u"), 0, dd[1]), (("p_mrho", "p_rho", "c_rho"), 1, dd[2])):
            want = (mold[slot * nc:(slot + 1) * nc], gold[slot * nc:(slot + 1) * nc], dk.ravel())
            for n, w in zip(names, want):
                assert np.array_equal(np.fromfile(f"{jac}_{n}.old", np.float32), w), n
    assert not os.path.exists(f"{jac}_p_vp.old")
    assert l2s[-1] < l2s[0]
    assert not np.array_equal(vs, vs0)


@pytest.mark.parametrize("L,fdorder,invmat1", [(1, 8, 1), (2, 4, 3), (3, 12, 1)])
def test_viscoelastic_forward_equals_oracle(L, fdorder, invmat1):
    """FD_SH.c with L > 0 (readmod_visc_SH, av_tau, prepare_update_s_visc_SH,
    update_s_visc_PML_SH): vz seismograms, final fields, CPML memory and the
    memory variables r (sxz) and q (syz) equal the FP32 oracle bit for bit;
    every scalar tile shape for L = 2."""
    print(f"[sh visco] L {L} FD{fdorder} INVMAT1 {invmat1}", flush=True)
    p, m, src, sig, rec = sh_case(fdorder=fdorder, invmat1=invmat1, nt=260)
    p = dnhip.Params(**{**p.__dict__, "L": L, "fl": [3.0, 25.0, 60.0][:L]})
    rng = np.random.default_rng(L)
    taus = rng.uniform(0.02, 0.12, m[0].shape).astype(np.float32)
    o = orc.Oracle(p, None, m[0], m[1], taus, prec="f32", aniso="sh")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    shapes = SCALAR if L == 2 else [None]
    for k, ts in enumerate(shapes):
        g = dnhip.Context(p)
        g.set_model(None, m[0], m[1], taus)
        g.set_sources(src, sig)
        g.set_receivers(rec)
        if ts:
            g.set_tile(*SCALAR[(k + 1) % len(SCALAR)], sweep=0)
            g.set_tile(*ts, sweep=1)
        g.forward()
        assert_equal_runs(g, o)
        for which in (0, 2):       # r at sxz, q at syz
            for l in range(1, L + 1):
                assert np.array_equal(g.memvar(which, l), o.memvar(which, l)), (which, l)
        assert np.abs(o.memvar(0, 1)).max() > 0
        g.close()


def test_viscoelastic_error_against_fp64_oracle_and_refusals():
    p, m, src, sig, rec = sh_case(nt=400)
    p = dnhip.Params(**{**p.__dict__, "L": 2, "fl": [3.0, 25.0]})
    taus = np.full(m[0].shape, 0.08, np.float32)
    g = dnhip.Context(p)
    with pytest.raises(dnhip.DnError):
        g.set_model(None, m[0], m[1])              # taus missing
    g.set_model(None, m[0], m[1], taus)
    g.set_sources(src, sig)
    g.set_receivers(rec)
    g.forward()
    o = orc.Oracle(p, None, m[0], m[1], taus, prec="f64", aniso="sh")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    err = rel_l2(g.seismogram(1), o.seismogram(1))
    print(f"[sh visco] vz rel-L2 vs FP64 oracle {err:.2e}", flush=True)
    assert err <= SEIS_TOL
    g.close()
    for bad in ({"L": 4, "fl": [3.0, 25.0, 60.0, 90.0]}, {"L": 2, "fl": [3.0, 25.0], "mode": 1}):
        with pytest.raises(dnhip.DnError):
            dnhip.Context(dnhip.Params(**{**p.__dict__, **bad}))


def test_cli_viscoelastic_forward_writes_the_library_seismograms(tmp_path):
    """PHYSICS 5 with L = 2 through the executable (readmod_visc_SH: MFILE.vs,
    .rho and .qs, taus = 2 / Qs): its vz seismograms equal the library's."""
    tmp = str(tmp_path)
    nx, ny, dh, nt = 110, 70, 10.0, 280
    _, vs, rho = synth.layered_model(nx, ny, dh, 23, sigma=3.0)
    dt = float(np.float32(synth.stable_dt(float(vs.max()) * 1.3, dh, 8, safety=0.6)))
    qs = np.random.default_rng(4).uniform(20.0, 60.0, (nx, ny)).astype(np.float32)
    for d in ("model", "su"):
        os.makedirs(os.path.join(tmp, d))
    for ext, a in (("vs", vs), ("rho", rho), ("qs", qs)):
        np.ascontiguousarray(a, np.float32).tofile(os.path.join(tmp, f"model/m.{ext}"))
    with open(os.path.join(tmp, "source.dat"), "w") as f:
        f.write("1\n400.00 0.0 30.00 0.0 15.00 1.00 0.0 1\n")
    with open(os.path.join(tmp, "receiver.dat"), "w") as f:
        for k in range(50):
            f.write(f"{50.0 + 20.0 * k:.3f}\t20.000\n")
    inpfile.write_inp(os.path.join(tmp, "m.inp"), PHYSICS=5, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt, SEISMO=1,
                      L=2, FL="3.0 30.0", INVMAT1=1)
    r = subprocess.run([BIN, "m.inp"], cwd=tmp, capture_output=True, text=True, timeout=120)
    print(r.stdout[-800:], r.stderr[-800:], flush=True)
    assert r.returncode == 0, r.stderr
    par = dnio.read_par(os.path.join(tmp, "m.inp"))
    cwd = os.getcwd()
    os.chdir(tmp)
    try:
        src, ts = dnio.read_sources(par.get("source_file"))
        rec = dnio.read_receivers(par)
    finally:
        os.chdir(cwd)
    p = dnhip.Params(nx=nx, ny=ny, nt=par.nt, dh=dh, dt=par.dt, fdorder=8, maxrelerror=par.maxrelerror, fw=par.fw,
                     free_surf=par.free_surf, seismo=1, ts=ts, physics=5, invmat1=1, L=2, fl=[3.0, 30.0],
                     damping=par.damping, fpml=par.fpml, npower=par.npower, k_max_pml=par.k_max_pml)
    g = dnhip.Context(p)
    g.set_model(None, vs, rho, (2.0 / qs.astype(np.float64)).astype(np.float32))
    loc = dnio.split_sources(par, src, 0, 1)
    g.set_sources(loc, dnio.wavelet(par, loc, 1, ts))
    g.set_receivers(rec[:2])
    g.forward()
    _, data = dnio.parse_su(os.path.join(tmp, "su/out_y.su.shot1"))
    assert np.array_equal(data, g.seismogram(1))
    assert np.abs(data).max() > 0
    g.close()


def test_cli_fwi_source_wavelet_inversion(tmp_path):
    """INV_STF 1 in SH FWI (grad_obj_sh.c:118-120 -> stf_sh.c) with a
    band-pass stage (TIME_FILT, FC_START 4 Hz, FC_END 12 Hz, ORDER 4): the
    spike run's signal is the band-limited spike (ORDER_SPIKE 5, FC_SPIKE_2
    15 Hz) filtered again with the stage's filter (stf_sh.c:30-50, unlike
    stf_psv.c), the observed vz data get the stage's filter, and the wavelet
    file SIGNAL_FILE_shot_<N>.dat equals the Wiener deconvolution of the
    restatement (oracle/stf_oracle.py) within its printed precision."""
    import stf_oracle as so
    import tdfilt_oracle as tf
    from test_gpu_cli import _files, _run
    tmp = str(tmp_path)
    os.makedirs(os.path.join(tmp, "wavelet"), exist_ok=True)
    nx, ny, dh, nt = 96, 64, 10.0, 400
    _, vs0, rho0 = synth.layered_model(nx, ny, dh, seed=6, sigma=3.0)
    dt = float(np.float32(synth.stable_dt(float(vs0.max()) * 1.2, dh, 8, safety=0.6)))
    vt = vs0.copy()
    vt[40:56, 10:20] *= 1.08
    sources = [(300.0, 30.0, 15.0, 1.0, 1), (650.0, 30.0, 15.0, 1.0, 1)]
    receivers = [(40.0 + 20.0 * k, 20.0) for k in range(40)]
    _files(tmp, nx, ny, dh, (vt, vt, rho0), sources, receivers, mname="model/true")
    inpfile.write_inp(os.path.join(tmp, "true.inp"), PHYSICS=5, NX=nx, NY=ny, DH=dh, DT=dt, TIME=nt * dt,
                      MFILE="model/true", SEIS_FILE_VX="obs/d_x.su", SEIS_FILE_VY="obs/d_y.su", SEISMO=1)
    _run(tmp, "true.inp")
    _files(tmp, nx, ny, dh, (vs0, vs0, rho0), sources, receivers, mname="model/start")
    inpfile.write_inp(os.path.join(tmp, "fwi.inp"), PHYSICS=5, MODE=1, QUELLTYPB=1, SEISMO=1, NX=nx, NY=ny, DH=dh,
                      DT=dt, TIME=nt * dt, MFILE="model/start", DTINV=2, TESTSHOT="1, 2, 1", ITERMAX=1,
                      MISFIT_LOG_FILE="misfit.dat", GRAD_METHOD=2)
    fcs, fce, order = 4.0, 12.0, 4
    inpfile.write_workflow(os.path.join(tmp, "wf.inp"), time_filt=1, fc_start=fcs, fc_end=fce, order=order,
                           inv_stf=1, normalize=2, eps_stf=0.1)
    _run(tmp, "fwi.inp", "wf.inp")
    par = dnio.read_par(os.path.join(tmp, "fwi.inp"))
    cwd = os.getcwd()
    os.chdir(tmp)
    try:
        src, ts = dnio.read_sources(par.get("source_file"))
        rec = dnio.read_receivers(par)
    finally:
        os.chdir(cwd)
    p = dnhip.Params(nx=nx, ny=ny, nt=par.nt, dh=dh, dt=par.dt, fdorder=par.fdorder, maxrelerror=par.maxrelerror,
                     fw=par.fw, free_surf=par.free_surf, seismo=1, ts=ts, physics=5, damping=par.damping,
                     fpml=par.fpml, npower=par.npower, k_max_pml=par.k_max_pml)
    g = dnhip.Context(p)
    g.set_model(None, vs0, rho0)
    for i in (1, 2):
        loc = dnio.split_sources(par, src, i - 1, 1)
        spike = np.zeros(nt, np.float32)
        spike[int(np.floor(loc[3, 0] / np.float32(par.dt) + 0.5))] = np.float32(1.0) * loc[5, 0]
        spike = tf.apply_tdfilt(spike[None, :], par.dt, 5, 15.0, -5.0)
        spike = tf.apply_tdfilt(spike, par.dt, order, fce, fcs)
        g.set_sources(loc, spike)
        g.set_receivers(rec[:2])
        g.forward()
        obs = tf.apply_tdfilt(dnio.read_su(os.path.join(tmp, f"obs/d_y.su.shot{i}"), rec.shape[1], nt), par.dt,
                              order, fce, fcs)
        want = so.stf(obs, g.seismogram(1), None, 2, 0.1, spike[0])
        got = np.loadtxt(os.path.join(tmp, f"wavelet/w_shot_{i}.dat"), dtype=np.float64).astype(np.float32)
        err = rel_l2(got, want)
        print(f"[sh stf] shot {i}: wavelet file vs restatement rel-L2 {err:.2e}", flush=True)
        assert got.shape == (nt,) and err < 1e-6
        assert np.max(np.abs(got - want)) <= 1e-6 * np.abs(want).max()
    g.close()
