"""Pins of the oracle's SH restatement (PHYSICS 5: src/SH/sh.c,
update_v_PML_SH.c, update_s_elastic_PML_SH.c, av_mu_SH.c, inv_rho_SH.c,
PML_pro_SH.c) against what the mathematics fixes:

- one stress sweep under uniform vz gradients is Hooke's law,
  sxz = DT mu d_x vz, syz = DT mu d_y vz;
- the velocity sweep is exact on polynomial stress fields:
  vz = DT/rho (d_x sxz + d_y syz);
- the SH system is the acoustic system under the exchange vz <-> p,
  (sxz, syz) <-> (vx, vy), 1/rho_SH <-> the acoustic modulus and
  1/mu_SH <-> the acoustic density (same staggering, same stencils; the
  harmonic mean of mu is the reciprocal of the arithmetic mean of 1/mu):
  SH after n+1 steps equals the acoustic run after n steps (FP64);
- the SH CPML profile is the P/SV one with alpha raised by 0.1 alpha_max
  inside the layer;
- the material: uip, ujp harmonic means of mu over two cells (with the
  one-rank periodic ghosts), 1/rho at the vz points;
- no free-surface routine: with FREE_SURF the ghost rows stay 0 and the top
  CPML memory untouched; SEISMO 1 records vz.
"""
import os

import numpy as np
import pytest

import orc
import synth
from cases import rel_l2
from dnhip import Params

orc.set_threads(2 if os.environ.get("PYTEST_XDIST_WORKER") else 8)

VS, RHO = 1800.0, 2000.0


def params(**kw):
    base = dict(nx=48, ny=40, nt=200, dh=10.0, dt=1e-4, fdorder=8, maxrelerror=0, fw=0,
                free_surf=0, seismo=0, ts=0.05, physics=5)
    base.update(kw)
    return Params(**base)


def const(nx, ny, v):
    return np.full((nx, ny), v, np.float32)


def _coords(o):
    nd = o.nd
    j = np.arange(-nd + 1, o.p.ny + nd + 1)[:, None].astype(float)
    i = np.arange(-nd + 1, o.p.nx + nd + 1)[None, :].astype(float)
    return j, i


def _sh(p, vs, rho, prec="f64"):
    return orc.Oracle(p, None, vs, rho, prec=prec, aniso="sh")


def _interior(o, n):
    nd = o.nd
    return (slice(nd - 1 + n + 1, nd - 1 + o.p.ny - n), slice(nd - 1 + n + 1, nd - 1 + o.p.nx - n))


@pytest.mark.parametrize("fdorder", [2, 8, 12])
def test_stress_sweep_is_hookes_law(fdorder):
    p = params(fdorder=fdorder)
    o = _sh(p, const(p.nx, p.ny, VS), const(p.nx, p.ny, RHO))
    j, i = _coords(o)
    a, b = 2.0e-3, -0.7e-3
    o.set_field(orc.ORC_VX, a * i * p.dh + b * j * p.dh)
    o.step(1)  # the stresses start at 0, so vz keeps its values
    sl = _interior(o, fdorder // 2)
    dt = float(np.float32(p.dt))
    mu = RHO * float(np.float32(VS)) ** 2
    assert np.allclose(o.field(orc.ORC_SXX)[sl], dt * mu * a, rtol=1e-7, atol=0)
    assert np.allclose(o.field(orc.ORC_SYY)[sl], dt * mu * b, rtol=1e-7, atol=0)
    assert not o.field(orc.ORC_VY).any() and not o.field(orc.ORC_SXY).any()


@pytest.mark.parametrize("fdorder", [2, 6, 12])
def test_velocity_sweep_is_exact_on_polynomial_stresses(fdorder):
    n = fdorder // 2
    p = params(nx=40, ny=36, fdorder=fdorder, dh=1.0, dt=1e-3)
    o = _sh(p, const(p.nx, p.ny, 1.0), const(p.nx, p.ny, 2.0))
    j, i = _coords(o)
    c = [0.3, -0.2, 0.05, 0.01, -0.002, 3e-4, -2e-5, 1e-6, -3e-8, 1e-9, -2e-11, 1e-12][: 2 * n]
    P = np.polynomial.Polynomial(c)
    dP = P.deriv()
    s = 0.07
    # sxz lives at i + 1/2, syz at j + 1/2
    o.set_field(orc.ORC_SXX, P(s * (i + 0.5)) + 0 * j)
    o.set_field(orc.ORC_SYY, P(s * 1.5 * (j + 0.5)) + 0 * i)
    o.step(1)
    sl = _interior(o, n)
    dt_rho = float(np.float32(p.dt)) / 2.0
    exp_vz = dt_rho * (s * dP(s * i) + 1.5 * s * dP(s * 1.5 * j))
    assert np.allclose(o.field(orc.ORC_VX)[sl], exp_vz[sl], rtol=1e-6, atol=0)


def test_sh_is_the_acoustic_system_with_exchanged_fields():
    """Heterogeneous mu and rho, no CPML (FW 0), no free surface, an
    initial vz bump and no sources (FP64, INVMAT1 3 in both).  The moduli
    and densities are powers of two, so their reciprocals enter both runs
    exactly."""
    nx, ny, nsteps = 60, 50, 80
    rng = np.random.default_rng(3)
    mu = (2.0 ** rng.integers(0, 3, (nx, ny))).astype(np.float32)
    rho = (2.0 ** rng.integers(-1, 2, (nx, ny))).astype(np.float32)
    p = params(nx=nx, ny=ny, dh=1.0, dt=0.1, fdorder=8, maxrelerror=1, invmat1=3)
    sh = _sh(p, mu, rho)
    ac = orc.Oracle(Params(**{**p.__dict__, "physics": 2}), (1.0 / rho.astype(np.float64)).astype(np.float32),
                    None, (1.0 / mu.astype(np.float64)).astype(np.float32), prec="f64", aniso="ac")
    j, i = _coords(sh)
    bump = np.exp(-(((i - 30.0) / 4.0) ** 2 + ((j - 22.0) / 3.0) ** 2))
    nd = sh.nd
    bump[:nd, :] = 0
    bump[-nd:, :] = 0
    bump[:, :nd] = 0
    bump[:, -nd:] = 0
    sh.set_field(orc.ORC_VX, bump)
    ac.set_field(orc.ORC_SXX, bump)
    for nt in range(1, nsteps + 1):
        sh.step(nt)
        ac.step(nt)
    # stresses of SH after n steps = velocities of the acoustic run after n
    scale = np.abs(ac.field(orc.ORC_VX)).max()
    assert scale > 1e-3
    assert np.allclose(sh.field(orc.ORC_SXX), ac.field(orc.ORC_VX), rtol=0, atol=1e-12 * scale)
    assert np.allclose(sh.field(orc.ORC_SYY), ac.field(orc.ORC_VY), rtol=0, atol=1e-12 * scale)
    # vz of SH after n+1 steps = p of the acoustic run after n
    sh.step(nsteps + 1)
    pscale = np.abs(ac.field(orc.ORC_SXX)).max()
    assert np.allclose(sh.field(orc.ORC_VX), ac.field(orc.ORC_SXX), rtol=0, atol=1e-12 * pscale)


def test_sh_cpml_profile_raises_alpha_by_a_tenth_of_its_maximum():
    fw = 12
    kw = dict(nx=60, ny=50, fw=fw, fpml=15.0, damping=1500.0, npower=2.0, k_max_pml=1.0)
    sh = _sh(params(**kw), const(60, 50, VS), const(60, 50, RHO))
    psv = orc.Oracle(params(physics=1, **kw), const(60, 50, 3000.0), const(60, 50, VS), const(60, 50, RHO),
                     prec="f64")
    alpha_max = float(np.float32(2.0 * np.pi * (15.0 / 2.0)))
    for al in (2, 7, 12, 17):  # alpha at the full and half points, x and y
        a_sh, a_psv = sh.pml(al), psv.pml(al)
        inside = a_psv > 0
        assert inside.sum() >= fw
        assert np.allclose(a_sh[inside] - a_psv[inside], 0.1 * alpha_max, rtol=1e-6), al
    # the outermost point (normalised depth 1): alpha = 0.1 alpha_max, and b
    # = exp(-(d/K + alpha) DT) there
    d, K, al, b = sh.pml(0)[0], sh.pml(1)[0], sh.pml(2)[0], sh.pml(4)[0]
    assert al == pytest.approx(0.1 * alpha_max, rel=1e-6)
    dt = float(np.float32(1e-4))
    assert b == pytest.approx(np.exp(-(d / K + al) * dt), rel=1e-12)
    # the damping profile itself is the P/SV one
    for k in (0, 1, 5, 6, 10, 11, 15, 16):
        assert np.array_equal(sh.pml(k), psv.pml(k)), k


def test_material_harmonic_means_and_inverse_density():
    """5 x 4 model of distinct values: uip, ujp = 2/(1/mu1 + 1/mu2) over
    the cell and its right / lower neighbour (the last column / row paired
    with column / row 1), rhoi = 1/rho; INVMAT1 1 forms mu = rho vs^2."""
    nx, ny = 5, 4
    vs = (1000.0 + 37.0 * np.arange(nx * ny)).reshape(nx, ny).astype(np.float32)
    rho = (1500.0 + 11.0 * np.arange(nx * ny)[::-1]).reshape(nx, ny).astype(np.float32)
    p = params(nx=nx, ny=ny, fdorder=2, dh=10.0, dt=1e-5)
    o = _sh(p, vs, rho)
    mu = rho.astype(np.float64) * vs.astype(np.float64) ** 2
    nd = o.nd
    uip, ujp, rhoi = o.field(orc.ORC_UIP), o.field(orc.ORC_UJP), o.field(orc.ORC_RIP)
    for i in range(nx):
        for j in range(ny):
            m1, m2, m3 = mu[i, j], mu[(i + 1) % nx, j], mu[i, (j + 1) % ny]
            r, c = nd + j, nd + i
            assert uip[r, c] == pytest.approx(2.0 / (1.0 / m1 + 1.0 / m2), rel=1e-12)
            assert ujp[r, c] == pytest.approx(2.0 / (1.0 / m1 + 1.0 / m3), rel=1e-12)
            assert rhoi[r, c] == pytest.approx(1.0 / float(rho[i, j]), rel=1e-12)


def test_free_surface_switch_and_vz_receivers():
    nx, ny, nt = 80, 60, 260
    vp, vs, rho = synth.layered_model(nx, ny, 10.0, 6, sigma=3.0)
    dt = float(np.float32(synth.stable_dt(float(vs.max()), 10.0, 8, safety=0.6)))
    p = Params(nx=nx, ny=ny, nt=nt, dh=10.0, dt=dt, fdorder=8, maxrelerror=1, fw=10, free_surf=1, seismo=1,
               ts=0.05, physics=5)
    src = synth.source_table(400.0, 30.0, 10.0, fc=20.0, stype=1)
    sig = synth.ricker(nt, dt, 20.0, amp=1e3)[None, :]
    rec = synth.receiver_line(30.0, 20.0, nx // 2 - 4, 20.0, 10.0)
    o = _sh(p, vs, rho, prec="f32")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    nd = o.nd
    for fid in (orc.ORC_VX, orc.ORC_SXX, orc.ORC_SYY):
        assert not o.field(fid)[:nd].any(), fid  # rows -nd+1..0 stay 0
    assert np.abs(o.field(orc.ORC_VX)[nd]).max() > 0
    # the y-band CPML memory: rows 1..FW (top) untouched, bottom band used
    for k in (5, 7):
        psi = o.psi(k)
        assert not psi[:p.fw].any() and np.abs(psi[p.fw:]).max() > 0, k
    # SEISMO 1 records vz in component 1; the other components stay 0
    f = o.field(orc.ORC_VX)
    jr, ir = rec[1] - 1 + nd, rec[0] - 1 + nd
    assert np.array_equal(o.seismogram(1)[:, -1], f[jr, ir])
    assert np.abs(o.seismogram(1)).max() > 0
    for c in (0, 2, 3, 4):
        assert not o.seismogram(c).any(), c


@pytest.mark.parametrize("stype", [2, 3, 4, 5])
def test_only_quelltyp_1_injects(stype):
    """update_v_PML_SH.c:560-568 injects QUELLTYP 1 (the force on vz) only."""
    p = params(nt=60, seismo=1)
    o = _sh(p, const(p.nx, p.ny, VS), const(p.nx, p.ny, RHO), prec="f32")
    o.set_sources(synth.source_table(240.0, 200.0, 10.0, fc=20.0, stype=stype), synth.ricker(60, p.dt, 20.0)[None, :])
    o.set_receivers(synth.receiver_line(100.0, 10.0, 10, 150.0, 10.0))
    o.forward()
    assert not o.field(orc.ORC_VX).any()


def test_sh_refuses_what_the_loop_does_not_record():
    p = params(seismo=2)
    with pytest.raises(ValueError, match="SEISMO"):
        _sh(p, const(p.nx, p.ny, VS), const(p.nx, p.ny, RHO))


# --------------------------------------------------------------------------
# SH gradient (src/SH/sh.c:150-223, ass_gradSH.c)
# --------------------------------------------------------------------------
def _sh_gradient_setup(dtinv=1, invmat1=3, grad_form=1, nx=64, ny=56, nt=700, fc=8.0, recrow=14, zero_data=False):
    """A force on vz in a model homogeneous around the source and receiver
    row and smoothly stiffer below; observed vz from the same model with
    +5% Gaussian anomalies of vs and rho.  misfit(model, grad) runs the
    FP64 oracle: forward, the vz residual (calc_res_SH.c), and with grad the
    adjoint run and ass_gradSH; the model is (vs or mu, rho)."""
    from test_oracle_pins import _bump
    p = params(nx=nx, ny=ny, nt=nt, dt=1e-3, fw=10, seismo=1, mode=1, dtinv=dtinv, fpml=fc, damping=2500.0,
               ts=1.0 / fc, invmat1=invmat1, grad_form=grad_form, quelltypb=1)
    jj = np.arange(ny)[None, :] * np.ones((nx, 1))
    f = np.clip((jj - recrow - 4) / 20.0, 0.0, 1.0)
    vs, rho = 1800.0 * (1 + 0.15 * f), 2000.0 * (1 + 0.1 * f)

    def model(vs, rho):
        first = vs if invmat1 == 1 else rho * vs * vs
        return [first.astype(np.float32), rho.astype(np.float32)]

    m0 = model(vs, rho)
    mt = model(vs * (1 + 0.05 * _bump(nx, ny, nx // 2 - 6, ny - 18, 5)),
               rho * (1 + 0.05 * _bump(nx, ny, nx // 2 + 6, ny - 20, 5)))
    src = np.zeros((8, 1), np.float32)
    src[0, 0], src[1, 0], src[4, 0], src[5, 0], src[7, 0] = nx // 2, recrow, fc, 1.0, 1
    sig = synth.ricker(nt, p.dt, fc, amp=1e3)[None, :]
    xs = list(range(14, nx - 13, 2))
    rec = np.array([xs, [recrow] * len(xs)], np.int32)

    def runner(m):
        o = _sh(p, *m)
        o.set_sources(src, sig)
        o.set_receivers(rec)
        return o

    t = runner(mt)
    t.forward()
    obs = t.seismogram(1).astype(np.float32)
    if zero_data:
        obs = np.zeros_like(obs)

    def misfit(m, grad=False):
        o = runner([np.asarray(a, np.float32) for a in m])
        o.forward()
        l2 = o.residuals(None, obs, None, 1)
        if not grad:
            return l2, None
        o.adjoint()
        o.assemble()
        return l2, [o.gradient(k).astype(np.float64) for k in (4, 5, 1, 2)]

    return p, m0, misfit, float(rho[0, recrow])


def _fd(misfit, m0, w, d):
    mp = [a.astype(np.float64) for a in m0]
    mm = [a.astype(np.float64) for a in m0]
    mp[w] = mp[w] + d
    mm[w] = mm[w] - d
    return (misfit(mp)[0] - misfit(mm)[0]) / 2.0


# The scale of the P/SV gradient pins (test_oracle_pins.py) holds here for
# the same reasons: the vz residual R is added to vz (update_v_PML_SH.c:
# 571-583, where the transposed update carries 1/rho), the adjoint vz is
# integrated into vzp1 (DT vz) and every DTINV-th step is correlated, so a
# gradient equals (rho_rec / DTINV) dJ/dm.  The shear term correlates
# forward and adjoint stresses with no 1/mu^2 (sh.c:195-205) and
# ass_gradSH.c:28 takes mu' = -DT times that sum, so mu' / mu^2 is the
# scaled derivative of the misfit with respect to mu.
@pytest.mark.parametrize("dtinv", [1, 2])
def test_sh_mu_and_rho_gradients_equal_scaled_misfit_derivative(dtinv):
    """INVMAT1 3, GRAD_FORM 1: along smooth directions (mu and rho enter
    through averages at staggered points) the misfit's central differences
    over mu'/mu^2 . dmu and over rho' . drho are DTINV / rho_rec within 8%."""
    from synth import gaussian_smooth as gaussian_filter
    p, m0, misfit, rho_rec = _sh_gradient_setup(dtinv=dtinv)
    nx, ny = p.nx, p.ny
    _, g = misfit(m0, grad=True)
    mu = m0[0].astype(np.float64)
    gm = [g[0] / (mu * mu), g[1]]
    for w, name in ((0, "mu"), (1, "rho")):
        ratios = []
        for s in (2.0, 3.0):
            for region in ((slice(12, 52), slice(24, 44)), (slice(14, 34), slice(22, 46)),
                           (slice(30, 52), slice(22, 46))):
                d = np.zeros((nx, ny))
                d[region] = gaussian_filter(gm[w], s)[region]
                d = gaussian_filter(d, s)
                d *= 1e-3 * float(np.abs(m0[w]).max()) / np.abs(d).max()
                ratios.append(_fd(misfit, m0, w, d) / float(np.sum(gm[w] * d)) * rho_rec / dtinv)
        ratios = np.array(ratios)
        print(f"SH {name}: FD / gradient x rho_rec / DTINV =", np.round(ratios, 4))
        assert np.all(np.abs(ratios - 1.0) < 0.08), name


@pytest.mark.parametrize("grad_form", [1, 2])
def test_sh_vs_parameterisation(grad_form):
    """INVMAT1 1 against INVMAT1 3 on the same physical model (zero data,
    FP64): the raw correlations are equal, the density gradient is the
    chain rule d/drho|vs = vs^2 d/dmu + d/drho|mu with d/dmu = mu'/mu^2
    (GRAD_FORM 1, whose mu' lacks the 1/mu^2) or mu' (GRAD_FORM 2), and the
    vs gradient is 2 rho vs mu' for GRAD_FORM 2 (the chain rule) and, for
    GRAD_FORM 1, (2/rho) vs^3 mu' as ass_gradSH.c:37 writes it -- not the
    chain rule's 2/(rho vs^3) mu', a precedence slip of the reference that
    the drop-in keeps."""
    _, m1, mis1, _ = _sh_gradient_setup(invmat1=1, grad_form=grad_form, zero_data=True, nt=500)
    _, m3, mis3, _ = _sh_gradient_setup(invmat1=3, grad_form=grad_form, zero_data=True, nt=500)
    _, g1 = mis1(m1, grad=True)
    _, g3 = mis3(m3, grad=True)
    vs, rho = (a.astype(np.float64) for a in m1)
    mu = m3[0].astype(np.float64)
    for k in (2, 3):                     # raw mu and rho correlations
        assert rel_l2(g1[k], g3[k]) < 1e-5, k
    dmu = g3[0] / (mu * mu) if grad_form == 1 else g3[0]
    want_rho = vs * vs * dmu + g3[1]
    want_vs = (2.0 / rho) * vs ** 3 * g3[0] if grad_form == 1 else 2.0 * rho * vs * g3[0]
    for got, want, name in ((g1[1], want_rho, "rho"), (g1[0], want_vs, "vs")):
        err = rel_l2(got, want)
        print(f"SH GRAD_FORM {grad_form} {name}: relative L2 {err:.2e}")
        assert np.abs(want).max() > 0
        assert err < 1e-5, name


def test_sh_adjoint_adds_the_residual_to_vz_only_for_quelltypb_1():
    """update_v_PML_SH.c:571-583: QUELLTYPB 1 adds the residual to vz at the
    receivers; other values inject nothing (the gradient stays 0)."""
    for qb, nonzero in ((1, True), (2, False)):
        p = params(nx=40, ny=36, nt=120, dt=1e-3, fw=8, seismo=1, mode=1, quelltypb=qb)
        o = _sh(p, const(p.nx, p.ny, VS), const(p.nx, p.ny, RHO))
        rec = np.array([[10, 20, 30], [5, 5, 5]], np.int32)
        o.set_receivers(rec)
        o.set_sources(np.array([[20], [10], [0], [0], [8.0], [1.0], [0], [1]], np.float32),
                      synth.ricker(p.nt, p.dt, 8.0, amp=1e3)[None, :])
        o.forward()
        l2 = o.residuals(None, np.zeros((3, p.nt), np.float32), None, 1)
        assert l2 > 0
        o.adjoint()
        assert bool(np.abs(o.gradient(2)).max() > 0) == nonzero


# --------------------------------------------------------------------------
# viscoelastic SH (src/SH/update_s_visc_PML_SH.c, prepare_update_s_visc_SH.c)
# --------------------------------------------------------------------------
def _sh_visc(p, vs, rho, taus, prec="f64"):
    return orc.Oracle(p, None, vs, rho, taus, prec=prec, aniso="sh")


def test_sh_visco_constant_strain_rate_matches_gsls():
    """Under uniform vz gradients held fixed (vz reset every step) each shear
    stress integrates the GSLS relaxation ODE: with mu_R = mu / (1 + tau
    sum_l w0^2 tau_l^2 / (1 + w0^2 tau_l^2)), w0 = 2 pi FL[1],
    sxz(t) = mu_R a t + mu_R tau a sum_l (1 - exp(-w_l t)) / w_l and syz
    likewise with b; trapezoidal steps leave O((w DT)^2)."""
    import math
    nx = ny = 40
    dh, dt, nsteps = 10.0, 1e-4, 400
    fl = [5.0, 40.0]
    tau = 0.08
    p = params(nx=nx, ny=ny, nt=nsteps, dh=dh, dt=dt, L=2, fl=fl)
    o = _sh_visc(p, const(nx, ny, VS), const(nx, ny, RHO), const(nx, ny, tau))
    a, b = 2.0e-3, -0.7e-3
    j, i = _coords(o)
    vz = a * (i - 1) * dh + b * (j - 1) * dh
    for nt_ in range(1, nsteps + 1):
        o.set_field(orc.ORC_VX, vz)
        o.step(nt_)
    t = nsteps * float(np.float32(dt))
    w = [2 * math.pi * f for f in fl]
    S = sum((w[0] / wl) ** 2 / (1 + (w[0] / wl) ** 2) for wl in w)
    tau32 = float(np.float32(tau))
    mu = RHO * float(np.float32(VS)) ** 2 / (1 + tau32 * S)
    relax = sum((1 - math.exp(-wl * t)) / wl for wl in w)
    nd = o.nd
    jc, ic = ny // 2 + nd - 1, nx // 2 + nd - 1
    for fid, rate in ((orc.ORC_SXX, a), (orc.ORC_SYY, b)):
        got = o.field(fid)[jc, ic]
        want = mu * rate * t + mu * tau32 * rate * relax
        assert got == pytest.approx(want, rel=2e-4), (fid, got, want)
        assert abs(mu * tau32 * rate * relax) > 50 * 2e-4 * abs(want)


def test_sh_visco_with_zero_tau_is_the_elastic_update():
    """taus = 0: the memory variables stay 0 and fipjp = uip DT, f = ujp DT,
    so the run equals the elastic SH run (FP64, to rounding)."""
    nx, ny = 60, 50
    _, vs, rho = synth.layered_model(nx, ny, 10.0, 3, sigma=3.0)
    base = dict(nx=nx, ny=ny, nt=300, dt=1e-3, fw=10, seismo=1)
    pv = params(L=2, fl=[4.0, 30.0], **base)
    pe = params(**base)
    src = np.array([[30], [10], [0], [0], [10.0], [1.0], [0], [1]], np.float32)
    sig = synth.ricker(pv.nt, pv.dt, 10.0, amp=1e3)[None, :]
    rec = np.array([[10, 30, 50], [5, 5, 5]], np.int32)
    runs = []
    for o in (_sh_visc(pv, vs, rho, np.zeros((nx, ny), np.float32)), _sh(pe, vs, rho)):
        o.set_sources(src, sig)
        o.set_receivers(rec)
        o.forward()
        runs.append(o)
    assert rel_l2(runs[0].seismogram(1), runs[1].seismogram(1)) < 1e-12
    assert np.abs(runs[1].seismogram(1)).max() > 0
    for fid in (orc.ORC_VX, orc.ORC_SXX, orc.ORC_SYY):
        assert rel_l2(runs[0].field(fid), runs[1].field(fid)) < 1e-12
