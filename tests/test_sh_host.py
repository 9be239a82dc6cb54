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
