"""Pins of the oracle's anisotropic P/SV restatement (PHYSICS 3 VTI,
PHYSICS 4 TTI: src/VTI/update_s_elastic_PML_VTI.c, src/TTI/
update_s_elastic_PML_TTI.c, src/TTI/rot_el_tensor_TTI.c, src/av_harm.c,
src/VTI/seismo_ssg_VTI.c) against what the mathematics fixes:

- one stress sweep under uniform strain rates is Hooke's law with the
  stiffness the update names (any swapped or misplaced coefficient fails);
- the rotated TTI stiffness is the tensor rotation of the VTI stiffness;
- a zero tilt gives the VTI run exactly, an isotropic stiffness the
  isotropic P/SV run;
- the harmonic means of av_harm on a small model with distinct values;
- no free-surface mirror: ghost rows stay zero.
"""
import math
import os

import numpy as np
import pytest

import orc
import synth
from cases import rel_l2
from dnhip import Params

orc.set_threads(2 if os.environ.get("PYTEST_XDIST_WORKER") else 8)


def params(**kw):
    base = dict(nx=48, ny=40, nt=200, dh=10.0, dt=1e-3, fdorder=8, maxrelerror=1, fw=0,
                free_surf=0, seismo=0, ts=0.05)
    base.update(kw)
    return Params(**base)


def const(nx, ny, v):
    return np.full((nx, ny), v, np.float32)


def _coords(o):
    nd = o.nd
    j = np.arange(-nd + 1, o.p.ny + nd + 1)[:, None].astype(float)
    i = np.arange(-nd + 1, o.p.nx + nd + 1)[None, :].astype(float)
    return j, i


def _linear_velocity(o, a, b, c, d, dh):
    """vx at (i+1/2, j), vy at (i, j+1/2) with uniform strain rates
    d_x vx = a, d_y vy = b, d_y vx = c, d_x vy = d."""
    j, i = _coords(o)
    return a * (i + 0.5) * dh + c * j * dh, b * (j + 0.5) * dh + d * i * dh


def bond_rotated(c11, c13, c33, c44, theta_deg):
    """The 2D (x, z) stiffness of a VTI medium rotated as a fourth-order
    tensor, C'_abcd = R_ai R_bj R_ck R_dl C_ijkl with R the rotation by
    -theta, in Voigt order (xx, zz, xz)."""
    V = np.array([[c11, c13, 0.0], [c13, c33, 0.0], [0.0, 0.0, c44]], np.float64)
    vo = {(0, 0): 0, (1, 1): 1, (0, 1): 2, (1, 0): 2}
    C = np.zeros((2, 2, 2, 2))
    for i in range(2):
        for j in range(2):
            for k in range(2):
                for l in range(2):
                    C[i, j, k, l] = V[vo[i, j], vo[k, l]]
    t = -math.radians(theta_deg)
    R = np.array([[math.cos(t), -math.sin(t)], [math.sin(t), math.cos(t)]])
    Cr = np.einsum("ai,bj,ck,dl,ijkl->abcd", R, R, R, R, C)
    idx = [(0, 0), (1, 1), (0, 1)]
    return np.array([[Cr[i, j, k, l] for (k, l) in idx] for (i, j) in idx])


# moduli exactly representable in float32 (the model files are float32)
C11, C13, C33, C44 = (float(np.float32(v)) for v in (3.0e10, 0.9e10, 2.2e10, 0.7e10))


@pytest.mark.parametrize("fdorder", [2, 8, 12])
def test_vti_stress_sweep_is_hookes_law(fdorder):
    """Uniform strain rates (a, b, c, d): after one sweep from rest
    sxx = DT (c11 a + c13 b), syy = DT (c13 a + c33 b),
    sxy = DT c44 (c + d)."""
    p = params(fdorder=fdorder, maxrelerror=0, dt=1e-4)
    nx, ny = p.nx, p.ny
    o = orc.Oracle(p, None, None, const(nx, ny, 2000.0), prec="f64",
                   aniso=(3, const(nx, ny, C11), const(nx, ny, C13), const(nx, ny, C33), const(nx, ny, C44), None))
    a, b, c, d = 2.0e-3, -0.7e-3, 1.1e-3, 0.4e-3
    vx, vy = _linear_velocity(o, a, b, c, d, p.dh)
    o.set_field(orc.ORC_VX, vx)
    o.set_field(orc.ORC_VY, vy)
    o.step(1)
    # the velocity sweep of the step saw zero stress: the velocities are unchanged
    nd, n = o.nd, fdorder // 2
    sl = (slice(nd - 1 + n + 1, nd - 1 + ny - n), slice(nd - 1 + n + 1, nd - 1 + nx - n))
    dt = float(np.float32(p.dt))
    for fid, ref in ((orc.ORC_SXX, dt * (C11 * a + C13 * b)), (orc.ORC_SYY, dt * (C13 * a + C33 * b)),
                     (orc.ORC_SXY, dt * C44 * (c + d))):
        got = o.field(fid)[sl]
        # float32 coefficient table: exact up to its rounding (about 1e-8 relative)
        assert np.allclose(got, ref, rtol=1e-7, atol=0), (fid, got.flat[0], ref)


@pytest.mark.parametrize("theta", [0.0, 17.0, -35.0, 90.0])
def test_tti_stiffness_is_the_rotated_vti_tensor(theta):
    """rot_el_tensor_TTI equals the tensor rotation of the VTI stiffness
    (FP64 oracle, 1e-12 relative to the largest modulus)."""
    p = params()
    nx, ny = p.nx, p.ny
    o = orc.Oracle(p, None, None, const(nx, ny, 2000.0), prec="f64",
                   aniso=(4, const(nx, ny, C11), const(nx, ny, C13), const(nx, ny, C33), const(nx, ny, C44),
                          const(nx, ny, theta)))
    want = bond_rotated(C11, C13, C33, C44, float(np.float32(theta)))
    nd = o.nd
    jc, ic = ny // 2 + nd - 1, nx // 2 + nd - 1
    got = {f: o.field(f)[jc, ic] for f in (orc.ORC_D11, orc.ORC_D13, orc.ORC_D15, orc.ORC_D33, orc.ORC_D35)}
    tol = 1e-12 * C11
    assert got[orc.ORC_D11] == pytest.approx(want[0, 0], abs=tol)
    assert got[orc.ORC_D13] == pytest.approx(want[0, 1], abs=tol)
    assert got[orc.ORC_D15] == pytest.approx(want[0, 2], abs=tol)
    assert got[orc.ORC_D33] == pytest.approx(want[1, 1], abs=tol)
    assert got[orc.ORC_D35] == pytest.approx(want[1, 2], abs=tol)
    # homogeneous: the harmonic means at the shear points equal the point values
    assert o.field(orc.ORC_D55H)[jc, ic] == pytest.approx(want[2, 2], abs=tol)
    if theta != 0.0 and theta != 90.0:
        assert abs(want[0, 2]) > 1e-3 * C11   # the coupling terms are exercised


def test_tti_stress_sweep_is_hookes_law_with_coupling():
    """Tilted medium under uniform strain rates: the three stresses follow
    the full rotated stiffness, including the normal-shear coupling
    d15, d35 (sxx = DT (d11 a + d13 b + d15 (c + d)), ...)."""
    p = params(fdorder=8, maxrelerror=0, dt=1e-4)
    nx, ny, theta = p.nx, p.ny, 28.0
    o = orc.Oracle(p, None, None, const(nx, ny, 2000.0), prec="f64",
                   aniso=(4, const(nx, ny, C11), const(nx, ny, C13), const(nx, ny, C33), const(nx, ny, C44),
                          const(nx, ny, theta)))
    D = bond_rotated(C11, C13, C33, C44, theta)
    a, b, c, d = 2.0e-3, -0.7e-3, 1.1e-3, 0.4e-3
    vx, vy = _linear_velocity(o, a, b, c, d, p.dh)
    o.set_field(orc.ORC_VX, vx)
    o.set_field(orc.ORC_VY, vy)
    o.step(1)
    nd, n = o.nd, 4
    sl = (slice(nd - 1 + n + 1, nd - 1 + ny - n), slice(nd - 1 + n + 1, nd - 1 + nx - n))
    dt = float(np.float32(p.dt))
    e = np.array([a, b, c + d])
    for row, fid in enumerate((orc.ORC_SXX, orc.ORC_SYY, orc.ORC_SXY)):
        ref = dt * float(D[row] @ e)
        assert np.allclose(o.field(fid)[sl], ref, rtol=1e-7, atol=0), fid


def _aniso_case(physics, nx=80, ny=60, nt=240, free_surf=1, fw=10, fdorder=8, seismo=4, tilt=None, seed=5):
    rho, c11, c13, c33, c44, theta = synth.anisotropic_model(nx, ny, 10.0, seed, tilt=tilt)
    vmax = float(np.sqrt(np.maximum(c11, c33) / rho).max())
    dt = float(np.float32(synth.stable_dt(vmax, 10.0, fdorder, safety=0.6)))
    p = Params(nx=nx, ny=ny, nt=nt, dh=10.0, dt=dt, fdorder=fdorder, maxrelerror=1, fw=fw,
               free_surf=free_surf, seismo=seismo, ts=0.05)
    src = synth.source_table(10.0 * (nx // 2), 40.0, 10.0, fc=20.0)
    sig = synth.ricker(nt, dt, 20.0, amp=1e3)[None, :]
    rec = synth.receiver_line(30.0, 20.0, nx // 2 - 4, 30.0, 10.0)
    return p, (rho, c11, c13, c33, c44, theta), src, sig, rec


def _run(p, model, src, sig, rec, physics, prec="f32", theta=None):
    rho, c11, c13, c33, c44, th = model
    o = orc.Oracle(p, None, None, rho, prec=prec, aniso=(physics, c11, c13, c33, c44, theta if physics == 4 else None))
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    return o


def test_tti_with_zero_tilt_is_the_vti_run():
    """theta = 0: d11 = c11, d13 = c13, d33 = c33, d15 = d35 = 0 (so d15h =
    d35h = 0) and d55h = c44h exactly, so the FP32 runs are equal."""
    p, m, src, sig, rec = _aniso_case(3)
    nx, ny = p.nx, p.ny
    v = _run(p, m, src, sig, rec, 3)
    t = _run(p, m, src, sig, rec, 4, theta=np.zeros((nx, ny), np.float32))
    for comp in (0, 1, 2):
        assert np.array_equal(v.seismogram(comp), t.seismogram(comp)), comp
    assert np.abs(v.seismogram(1)).max() > 0
    assert np.array_equal(t.field(orc.ORC_D55H), v.field(orc.ORC_C44H))


def test_vti_isotropic_limit_is_the_isotropic_run():
    """c11 = c33 = lambda + 2 mu, c13 = lambda, c44 = mu: the VTI run equals
    the isotropic P/SV run without a free surface (FP64, rounding only)."""
    nx, ny, nt = 72, 56, 220
    vp, vs, rho = synth.layered_model(nx, ny, 10.0, 9, sigma=3.0)
    dt = float(np.float32(synth.stable_dt(float(vp.max()), 10.0, 8, safety=0.6)))
    p = Params(nx=nx, ny=ny, nt=nt, dh=10.0, dt=dt, fdorder=8, fw=10, free_surf=0, seismo=4, ts=0.05)
    r, a, b = rho.astype(np.float64), vp.astype(np.float64), vs.astype(np.float64)
    mu, lam = r * b * b, r * (a * a - 2 * b * b)
    src = synth.source_table(10.0 * (nx // 2), 100.0, 10.0, fc=20.0)
    sig = synth.ricker(nt, dt, 20.0, amp=1e3)[None, :]
    rec = synth.receiver_line(30.0, 20.0, 30, 150.0, 10.0)
    iso = orc.Oracle(p, vp, vs, rho, prec="f64")
    iso.set_sources(src, sig)
    iso.set_receivers(rec)
    iso.forward()
    # the moduli the isotropic update forms itself, handed over as the VTI tensor
    f = (r * b * b).astype(np.float64)
    ani = orc.Oracle(p, None, None, rho, prec="f64",
                     aniso=(3, (lam + 2 * mu).astype(np.float32), lam.astype(np.float32),
                            (lam + 2 * mu).astype(np.float32), f.astype(np.float32), None))
    ani.set_sources(src, sig)
    ani.set_receivers(rec)
    ani.forward()
    for comp in (0, 1, 2):
        err = rel_l2(ani.seismogram(comp), iso.seismogram(comp))
        assert err < 1e-6, (comp, err)


def test_av_harm_known_answer():
    """c44h = 4 / sum(1/c44) over the four cells around (i+1/2, j+1/2) with
    the one-rank periodic ghosts (last row/column averaged with the first),
    0 next to a zero cell (src/av_harm.c:13-19)."""
    nx, ny = 5, 4
    p = params(nx=nx, ny=ny, nt=4)
    rng = np.random.default_rng(4)
    c44 = (1e9 * (1.0 + rng.random((nx, ny)))).astype(np.float32)
    c44[2, 1] = 0.0
    o = orc.Oracle(p, None, None, const(nx, ny, 2000.0), prec="f64",
                   aniso=(3, const(nx, ny, C11), const(nx, ny, C13), const(nx, ny, C33), c44, None))
    g = o.field(orc.ORC_C44H)
    nd = o.nd
    m = c44.astype(np.float64).T  # [j][i]
    zeros = 0
    for j in range(ny):
        for i in range(nx):
            four = [m[j, i], m[j, (i + 1) % nx], m[(j + 1) % ny, i], m[(j + 1) % ny, (i + 1) % nx]]
            want = 0.0 if min(four) == 0.0 else 4.0 / sum(1.0 / v for v in four)
            zeros += want == 0.0
            assert g[j + nd, i + nd] == pytest.approx(want, rel=1e-12), (j, i)
    assert zeros == 4


def test_no_free_surface_mirror_and_no_top_cpml():
    """VTI/TTI with FREE_SURF: the loop calls no surface routine, so the
    stress ghost rows above row 1 stay 0, and the top CPML is off."""
    p, m, src, sig, rec = _aniso_case(4, tilt=25.0)
    o = _run(p, m, src, sig, rec, 4, theta=m[5])
    nd = o.nd
    for fid in (orc.ORC_SXX, orc.ORC_SYY, orc.ORC_SXY):
        f = o.field(fid)
        assert not f[:nd].any(), fid
        assert np.abs(f[nd:nd + 3]).max() > 0, fid
    # y-band CPML memory of the top rows (h = 1..FW) is untouched
    assert not o.psi(7)[: p.fw].any() and not o.psi(8)[: p.fw].any()
    assert np.abs(o.psi(7)[p.fw:]).max() > 0


def test_seismo_modes_of_the_anisotropic_loop():
    """seismo_ssg_VTI: SEISMO 4 samples vx, vy and p = -sxx - syy and no
    div/curl; SEISMO 3 samples nothing."""
    p, m, src, sig, rec = _aniso_case(3, seismo=4)
    o = _run(p, m, src, sig, rec, 3)
    assert np.abs(o.seismogram(2)).max() > 0
    assert not o.seismogram(3).any() and not o.seismogram(4).any()
    p3, *_ = _aniso_case(3, seismo=3)
    o3 = _run(p3, m, src, sig, rec, 3)
    for comp in range(5):
        assert not o3.seismogram(comp).any(), comp


def test_oracle_refuses_tti_adjoint_with_grad_form_2():
    """TTI.c stores ux/uy/uxy for GRAD_FORM 2, which the TTI stress update
    never writes; the oracle refuses that combination (GRAD_FORM 1 runs)."""
    p = params(mode=1, grad_form=2)
    nx, ny = p.nx, p.ny
    th = np.zeros((nx, ny), np.float32)
    model = (4, const(nx, ny, C11), const(nx, ny, C13), const(nx, ny, C33), const(nx, ny, C44), th)
    with pytest.raises(ValueError, match="GRAD_FORM"):
        orc.Oracle(p, None, None, const(nx, ny, 2000.0), prec="f32", aniso=model)
    orc.Oracle(params(mode=1, grad_form=1), None, None, const(nx, ny, 2000.0), prec="f32", aniso=model).close()


def _adjoint(o, obs):
    o.forward()
    o.residuals(obs[0], obs[1], None, 1)
    o.adjoint()
    return [o.gradient(k) for k in range(3)]


@pytest.mark.parametrize("grad_form", [1, 2])
def test_vti_rtm_images_in_the_isotropic_limit(grad_form):
    """The VTI loop's adjoint pass (src/VTI/VTI.c:235-368): with the
    isotropic tensor its P image (sum of (fx + fy)(sxx + syy)) and density
    image equal the isotropic P/SV run's raw correlations, and its shear
    image stays 0 (the reference comments that term out); FP64, GRAD_FORM 1
    (stored stresses) and 2 (stored ux, uy, uxy)."""
    nx, ny, nt = 72, 56, 200
    vp, vs, rho = synth.layered_model(nx, ny, 10.0, 9, sigma=3.0)
    dt = float(np.float32(synth.stable_dt(float(vp.max()), 10.0, 8, safety=0.6)))
    p = Params(nx=nx, ny=ny, nt=nt, dh=10.0, dt=dt, fdorder=8, fw=10, free_surf=0, seismo=1, ts=0.05,
               mode=1, dtinv=2, grad_form=grad_form)
    r, a, b = rho.astype(np.float64), vp.astype(np.float64), vs.astype(np.float64)
    mu, lam = r * b * b, r * (a * a - 2 * b * b)
    src = synth.source_table(10.0 * (nx // 2), 100.0, 10.0, fc=20.0)
    sig = synth.ricker(nt, dt, 20.0, amp=1e3)[None, :]
    rec = synth.receiver_line(30.0, 20.0, 30, 150.0, 10.0)
    iso = orc.Oracle(p, vp, vs, rho, prec="f64")
    iso.set_sources(src, sig)
    iso.set_receivers(rec)
    iso.forward()
    obs = (0.8 * iso.seismogram(0), 1.1 * iso.seismogram(1))
    g_iso = _adjoint(iso, obs)
    ani = orc.Oracle(p, None, None, rho, prec="f64",
                     aniso=(3, (lam + 2 * mu).astype(np.float32), lam.astype(np.float32),
                            (lam + 2 * mu).astype(np.float32), mu.astype(np.float32), None))
    ani.set_sources(src, sig)
    ani.set_receivers(rec)
    g_ani = _adjoint(ani, obs)
    for k in (0, 2):
        err = rel_l2(g_ani[k], g_iso[k])
        assert err < 1e-6 and np.abs(g_iso[k]).max() > 0, (k, err)
    assert not g_ani[1].any() and np.abs(g_iso[1]).max() > 0


def test_tti_rtm_with_zero_tilt_is_the_vti_rtm():
    p, m, src, sig, rec = _aniso_case(3, seismo=1)
    p = Params(**{**p.__dict__, "mode": 1, "dtinv": 2})
    nx, ny = p.nx, p.ny
    v = _run(p, m, src, sig, rec, 3)
    obs = (0.7 * v.seismogram(0), 0.7 * v.seismogram(1))
    gv = _adjoint(v, obs)
    t = orc.Oracle(p, None, None, m[0], prec="f32",
                   aniso=(4, m[1], m[2], m[3], m[4], np.zeros((nx, ny), np.float32)))
    t.set_sources(src, sig)
    t.set_receivers(rec)
    gt = _adjoint(t, obs)
    for k in range(3):
        assert np.array_equal(gv[k], gt[k]), k
    assert np.abs(gv[0]).max() > 0 and np.abs(gv[2]).max() > 0
