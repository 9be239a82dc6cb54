"""Pins of the oracle's acoustic restatement (PHYSICS 2: src/AC/ac.c,
update_v_PML_AC.c, update_s_acoustic_PML_AC.c, surface_acoustic_PML_AC.c,
psource_AC.c, seismo_AC.c) against what the mathematics fixes:

- one pressure sweep under uniform velocity gradients is p = DT rho vp^2
  (d_x vx + d_y vy);
- the velocity sweep is exact on polynomial pressure fields:
  vx = DT/rho d_x p, vy = DT/rho d_y p;
- the acoustic run equals the P/SV run of the same model with vs = 0
  (sxx = syy = p, sxy = 0) without a free surface (FP64);
- the free surface: p = 0 at row 1, ghost rows antisymmetric.
"""
import os

import numpy as np
import pytest

import orc
import synth
from cases import rel_l2
from dnhip import Params

orc.set_threads(2 if os.environ.get("PYTEST_XDIST_WORKER") else 8)

VP, RHO = 2500.0, 1800.0


def params(**kw):
    base = dict(nx=48, ny=40, nt=200, dh=10.0, dt=1e-4, fdorder=8, maxrelerror=0, fw=0,
                free_surf=0, seismo=0, ts=0.05, physics=2)
    base.update(kw)
    return Params(**base)


def const(nx, ny, v):
    return np.full((nx, ny), v, np.float32)


def _coords(o):
    nd = o.nd
    j = np.arange(-nd + 1, o.p.ny + nd + 1)[:, None].astype(float)
    i = np.arange(-nd + 1, o.p.nx + nd + 1)[None, :].astype(float)
    return j, i


def _ac(p, vp, rho, prec="f64"):
    return orc.Oracle(p, vp, None, rho, prec=prec, aniso="ac")


@pytest.mark.parametrize("fdorder", [2, 8, 12])
def test_pressure_sweep_is_the_bulk_modulus_times_the_divergence(fdorder):
    p = params(fdorder=fdorder)
    o = _ac(p, const(p.nx, p.ny, VP), const(p.nx, p.ny, RHO))
    j, i = _coords(o)
    a, b = 2.0e-3, -0.7e-3
    o.set_field(orc.ORC_VX, a * (i + 0.5) * p.dh + 0.3e-3 * j * p.dh)
    o.set_field(orc.ORC_VY, b * (j + 0.5) * p.dh + 0.2e-3 * i * p.dh)
    o.step(1)
    nd, n = o.nd, fdorder // 2
    sl = (slice(nd - 1 + n + 1, nd - 1 + p.ny - n), slice(nd - 1 + n + 1, nd - 1 + p.nx - n))
    dt = float(np.float32(p.dt))
    k = RHO * float(np.float32(VP)) ** 2
    assert np.allclose(o.field(orc.ORC_SXX)[sl], dt * k * (a + b), rtol=1e-7, atol=0)


@pytest.mark.parametrize("fdorder", [2, 6, 12])
def test_velocity_sweep_is_exact_on_polynomial_pressure(fdorder):
    n = fdorder // 2
    p = params(nx=40, ny=36, fdorder=fdorder, dh=1.0, dt=1e-3)
    o = _ac(p, const(p.nx, p.ny, 1.0), const(p.nx, p.ny, 2.0))
    j, i = _coords(o)
    c = [0.3, -0.2, 0.05, 0.01, -0.002, 3e-4, -2e-5, 1e-6, -3e-8, 1e-9, -2e-11, 1e-12][: 2 * n]
    P = np.polynomial.Polynomial(c)
    dP = P.deriv()
    s = 0.07
    o.set_field(orc.ORC_SXX, P(s * i) + P(s * 1.5 * j))
    o.step(1)
    nd = o.nd
    sl = (slice(nd - 1 + n + 1, nd - 1 + p.ny - n), slice(nd - 1 + n + 1, nd - 1 + p.nx - n))
    dt_rho = float(np.float32(p.dt)) / 2.0
    exp_vx = dt_rho * s * dP(s * (i + 0.5)) + 0 * j
    exp_vy = dt_rho * 1.5 * s * dP(s * 1.5 * (j + 0.5)) + 0 * i
    assert np.allclose(o.field(orc.ORC_VX)[sl], exp_vx[sl], rtol=1e-6, atol=0)
    assert np.allclose(o.field(orc.ORC_VY)[sl], exp_vy[sl], rtol=1e-6, atol=0)


def _case(nx=80, ny=60, nt=260, free_surf=1, seismo=4, stype=1, fw=10):
    vp, _, rho = synth.layered_model(nx, ny, 10.0, 6, sigma=3.0)
    dt = float(np.float32(synth.stable_dt(float(vp.max()), 10.0, 8, safety=0.6)))
    p = Params(nx=nx, ny=ny, nt=nt, dh=10.0, dt=dt, fdorder=8, maxrelerror=1, fw=fw, free_surf=free_surf,
               seismo=seismo, ts=0.05, physics=2)
    src = synth.source_table(10.0 * (nx // 2), 100.0, 10.0, fc=20.0, stype=stype, azimuth=30.0)
    sig = synth.ricker(nt, dt, 20.0, amp=1e3)[None, :]
    rec = synth.receiver_line(30.0, 20.0, nx // 2 - 4, 150.0, 10.0)
    return p, vp, rho, src, sig, rec


def test_acoustic_run_is_the_elastic_run_without_shear():
    """vs = 0 in the P/SV path: sxy stays 0 and sxx = syy = p, so the
    velocities equal and the P/SV pressure -sxx-syy is -2p (FP64)."""
    p, vp, rho, src, sig, rec = _case(free_surf=0, seismo=4)
    ac = _ac(p, vp, rho)
    ac.set_sources(src, sig)
    ac.set_receivers(rec)
    ac.forward()
    pe = Params(**{**p.__dict__, "physics": 1})
    el = orc.Oracle(pe, vp, np.zeros_like(vp), rho, prec="f64")
    el.set_sources(src, sig)
    el.set_receivers(rec)
    el.forward()
    for c in (0, 1):
        assert rel_l2(ac.seismogram(c), el.seismogram(c)) < 1e-9, c
    assert rel_l2(ac.seismogram(2), -0.5 * el.seismogram(2)) < 1e-9
    assert np.abs(ac.seismogram(2)).max() > 0


def test_free_surface_sets_p_to_zero_and_mirrors_the_ghost_rows():
    p, vp, rho, src, sig, rec = _case(free_surf=1, seismo=2)
    o = _ac(p, vp, rho, prec="f32")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    f = o.field(orc.ORC_SXX)
    nd = o.nd
    row1 = nd  # index of j = 1 in the padded layout (rows -nd+1 ..)
    assert not f[row1, nd:nd + p.nx].any()
    for m in range(1, p.fdorder // 2 + 1):
        assert np.array_equal(f[row1 - m, nd:nd + p.nx], -f[row1 + m, nd:nd + p.nx]), m
    assert np.abs(f[row1 + 1]).max() > 0
    # SEISMO 2 samples p itself (seismo_AC.c: sectionp = p)
    j, i = rec[1] - 1 + nd, rec[0] - 1 + nd
    assert np.array_equal(o.seismogram(2)[:, -1], f[j, i])


def test_seismo_modes_of_the_acoustic_loop():
    """SEISMO 4 gives vx, vy, p, div and curl; curl uses vp as the
    reference passes ppi for u (ac.c:199-201), so both factors are
    sqrt(rho vp^2)."""
    p, vp, rho, src, sig, rec = _case(seismo=4)
    o = _ac(p, vp, rho, prec="f32")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    for c in range(5):
        assert np.abs(o.seismogram(c)).max() > 0, c
    p3 = Params(**{**p.__dict__, "seismo": 3})
    o3 = _ac(p3, vp, rho, prec="f32")
    o3.set_sources(src, sig)
    o3.set_receivers(rec)
    o3.forward()
    assert not o3.seismogram(0).any() and not o3.seismogram(2).any()
    assert np.array_equal(o3.seismogram(3), o.seismogram(3))


def _adjoint(o, obs):
    o.forward()
    o.residuals(obs[0], obs[1], None, 1)
    o.adjoint()
    raw = [o.gradient(k) for k in range(3)]
    o.assemble()
    return raw, [o.gradient(k) for k in range(3, 6)]


@pytest.mark.parametrize("grad_form", [1, 2])
def test_acoustic_adjoint_is_the_elastic_adjoint_without_shear(grad_form):
    """The acoustic adjoint pass (ac.c mode 1: adjoint forces, vxp1 time
    integral, p correlation) against the P/SV one on the same model with
    mu = 0 (INVMAT1 3, FP64, no free surface): there sxx = syy = p, so the
    P/SV P correlation (fx + fy)(sxx + syy) is 4 times the acoustic fx p
    and the density correlations agree; assembled (GRAD_FORM 1), the
    acoustic lambda gradient -DT fx p / (4 lambda^2) is a quarter of the
    P/SV one and the density gradients -DT w_rho agree."""
    nx, ny, nt = 72, 56, 200
    vp, _, rho = synth.layered_model(nx, ny, 10.0, 9, sigma=3.0)
    lam = (rho.astype(np.float64) * vp.astype(np.float64) ** 2).astype(np.float32)
    dt = float(np.float32(synth.stable_dt(float(vp.max()), 10.0, 8, safety=0.6)))
    base = dict(nx=nx, ny=ny, nt=nt, dh=10.0, dt=dt, fdorder=8, maxrelerror=1, fw=10, free_surf=0, seismo=1,
                ts=0.05, mode=1, dtinv=2, grad_form=grad_form, invmat1=3)
    src = synth.source_table(10.0 * (nx // 2), 100.0, 10.0, fc=20.0)
    sig = synth.ricker(nt, dt, 20.0, amp=1e3)[None, :]
    rec = synth.receiver_line(30.0, 20.0, 30, 150.0, 10.0)
    el = orc.Oracle(Params(**base, physics=1), lam, np.zeros_like(lam), rho, prec="f64")
    ac = orc.Oracle(Params(**base, physics=2), lam, None, rho, prec="f64", aniso="ac")
    for o in (el, ac):
        o.set_sources(src, sig)
        o.set_receivers(rec)
    el.forward()
    obs = (0.8 * el.seismogram(0), 1.1 * el.seismogram(1))
    raw_el, grad_el = _adjoint(el, obs)
    raw_ac, grad_ac = _adjoint(ac, obs)
    assert np.abs(raw_ac[0]).max() > 0 and np.abs(raw_ac[2]).max() > 0
    assert rel_l2(raw_ac[0], 0.25 * raw_el[0]) < 1e-9
    assert rel_l2(raw_ac[2], raw_el[2]) < 1e-9
    assert not raw_ac[1].any() and not grad_ac[1].any()
    if grad_form == 1:
        assert rel_l2(grad_ac[0], 0.25 * grad_el[0]) < 1e-9
    assert rel_l2(grad_ac[2], grad_el[2]) < 1e-9


def test_acoustic_adjoint_pressure_source_is_added_once():
    """QUELLTYPB 4 (psource_AC.c:36-47): the residual of the pressure
    seismogram enters p once per receiver and step; with the adjoint
    sources set to one unit impulse at step 1, p at that receiver after one
    adjoint step is exactly 1 (the velocities are still 0, so the pressure
    update adds nothing)."""
    p = params(mode=1, quelltypb=4, seismo=2, nt=20)
    o = _ac(p, const(p.nx, p.ny, VP), const(p.nx, p.ny, RHO))
    o.set_sources(synth.source_table(200.0, 200.0, 10.0, fc=20.0), synth.ricker(20, p.dt, 20.0)[None, :])
    rec = np.array([[20], [15]], np.int32)
    o.set_receivers(rec)
    dp = np.zeros((1, 20), np.float32)
    dp[0, 0] = 1.0
    o.set_adjoint_sources(None, None, dp)
    o.zero()
    o.step(1, mode=1)
    nd = o.nd
    assert o.field(orc.ORC_SXX)[15 - 1 + nd, 20 - 1 + nd] == 1.0
