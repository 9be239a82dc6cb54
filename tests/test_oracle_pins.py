"""Pins of the CPU oracle against what the mathematics fixes (CPU only).

The reference ships no tests, golden vectors or fixtures for this path, and
its sources need mpi.h, so the oracle is pinned here by checks that are
independent of its own formulas: closed forms, exact invariants of the
discrete operators, and physics it must reproduce.  A plausible mistake (a
wrong stencil offset, a swapped operand, a dropped term, a sign) breaks at
least one of them.
"""
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import orc
import synth
from cases import rel_l2
from dnhip import Params

# OpenMP threads of the oracle: parallel test workers share the cores, and
# oversubscribed spin barriers of the many small time steps here crawl
orc.set_threads(2 if os.environ.get("PYTEST_XDIST_WORKER") else 8)


def params(**kw):
    base = dict(nx=64, ny=48, nt=200, dh=10.0, dt=1e-3, fdorder=8, maxrelerror=1, fw=0,
                free_surf=0, seismo=0, ts=0.05)
    base.update(kw)
    return Params(**base)


def homogeneous(nx, ny, vp=3000.0, vs=1732.0, rho=2000.0):
    return (np.full((nx, ny), vp, np.float32), np.full((nx, ny), vs, np.float32),
            np.full((nx, ny), rho, np.float32))


# --------------------------------------------------------------------------
# holbergcoeff
# --------------------------------------------------------------------------
def taylor_staggered(n):
    """Coefficients c_1..c_n of the staggered first derivative
    (1/h) sum c_m [f(x+(m-1/2)h) - f(x-(m-1/2)h)] exact to order 2n:
    sum_m c_m (2m-1)^(2k-1) = delta_k1, k = 1..n (exact rationals)."""
    A = [[Fraction(2 * m - 1) ** (2 * k - 1) for m in range(1, n + 1)] for k in range(1, n + 1)]
    b = [Fraction(1 if k == 0 else 0) for k in range(n)]
    # Gauss-Jordan in rationals
    M = [row[:] + [b[r]] for r, row in enumerate(A)]
    for col in range(n):
        piv = next(r for r in range(col, n) if M[r][col] != 0)
        M[col], M[piv] = M[piv], M[col]
        pv = M[col][col]
        M[col] = [x / pv for x in M[col]]
        for r in range(n):
            if r != col and M[r][col] != 0:
                f = M[r][col]
                M[r] = [x - f * y for x, y in zip(M[r], M[col])]
    return [M[r][n] for r in range(n)]


@pytest.mark.parametrize("fdorder", [2, 4, 6, 8, 10, 12])
def test_taylor_row_is_the_staggered_taylor_stencil(fdorder):
    n = fdorder // 2
    exact = taylor_staggered(n)
    hc = orc.holberg(fdorder, 0)
    for m in range(1, n + 1):
        assert hc[m] == np.float32(float(exact[m - 1])), (fdorder, m)
    for m in range(n + 1, 7):
        assert hc[m] == 0.0


def phase_velocity_error(hc, n, khmax, samples=4000):
    """max |k_num/k - 1| of the staggered operator, k_num h = 2 sum c_m sin((2m-1) kh/2)."""
    kh = np.linspace(1e-6, khmax, samples)
    kn = sum(float(hc[m]) * 2.0 * np.sin((2 * m - 1) * kh / 2) for m in range(1, n + 1)) / kh
    return float(np.max(np.abs(kn - 1.0)))


@pytest.mark.parametrize("mre,E", [(1, 0.001), (2, 0.005), (3, 0.01), (4, 0.03)])
@pytest.mark.parametrize("fdorder", [2, 4, 6, 8, 10, 12])
def test_holberg_rows_meet_their_error_bound(mre, E, fdorder):
    """Every Holberg row keeps the relative phase-velocity error of the
    staggered derivative below its design value E (0.1, 0.5, 1, 3 %) up to
    kh = 2*pi/G, G = hc[0] grid points per shortest wavelength.  The published
    coefficients are rounded, hence 2 % slack on E; a wrong digit in any
    coefficient breaks the bound."""
    n = fdorder // 2
    hc = orc.holberg(fdorder, mre)
    khmax = 2 * math.pi / float(hc[0])
    err = phase_velocity_error(hc, n, khmax)
    assert err <= 1.02 * E + 5e-5, (fdorder, mre, err)
    # the bound is reached, not vacuous
    assert err >= 0.9 * E, (fdorder, mre, err)


def test_holberg_rejects_bad_order():
    with pytest.raises(ValueError):
        orc.holberg(7, 1)
    with pytest.raises(ValueError):
        orc.holberg(8, 5)


# --------------------------------------------------------------------------
# PML_pro (Komatitsch & Martin 2007 C-PML profiles)
# --------------------------------------------------------------------------
PML_NAMES = ["d_x", "K_x", "alpha_x", "a_x", "b_x", "d_x_half", "K_x_half", "alpha_x_half", "a_x_half",
             "b_x_half", "d_y", "K_y", "alpha_y", "a_y", "b_y", "d_y_half", "K_y_half", "alpha_y_half",
             "a_y_half", "b_y_half"]


def cpml_closed_form(n, fw, dh, dt, damping, fpml, npower, kmax, clamp_alpha):
    """d(x) = d0 (x/L)^N, K = 1 + (kmax-1)(x/L)^N, alpha = alpha_max (1-x/L),
    d0 = -(N+1) V log(R) / (2L), R = 1e-3, alpha_max = pi fpml,
    b = exp(-(d/K + alpha) dt), a = d (b-1) / (K (d + K alpha)) -- evaluated
    in float64 at the grid points and half points of both bands."""
    L = fw * dh
    d0 = -(npower + 1) * damping * math.log(0.001) / (2 * L)
    amax = math.pi * fpml
    out = {k: np.zeros(2 * fw) for k in ("d", "K", "al", "dh", "Kh", "alh")}
    for h in range(1, 2 * fw + 1):
        if h <= fw:
            x = (h - 1) * dh
            absc, absch = L - x, L - (x + dh / 2)
        else:
            i = n - 2 * fw + h
            x = (i - 1) * dh
            absc, absch = x - ((n - 1) * dh - L), x + dh / 2 - ((n - 1) * dh - L)
        for key, a in (("", absc), ("h", absch)):
            if a >= 0:
                xn = a / L
                out["d" + key][h - 1] = d0 * xn ** npower
                out["K" + key][h - 1] = 1 + (kmax - 1) * xn ** npower
                out["al" + key][h - 1] = amax * (1 - xn)
            else:
                out["K" + key][h - 1] = 1.0
    if clamp_alpha:
        out["al"] = np.maximum(out["al"], 0)
        out["alh"] = np.maximum(out["alh"], 0)
    for key in ("", "h"):
        d, K, al = out["d" + key], out["K" + key], out["al" + key]
        b = np.exp(-(d / K + al) * dt)
        # the damping gate: a is set only where the integer part of d is non-zero
        gate = np.abs(np.trunc(d)) > 0
        a = np.where(gate, d * (b - 1) / (K * (d + K * al) + (~gate)), 0.0)
        out["b" + key], out["a" + key] = b, a
    return out


@pytest.mark.parametrize("fw,kmax", [(20, 1.0), (10, 2.0), (7, 1.5)])
def test_cpml_profiles_closed_form(fw, kmax):
    p = params(nx=90, ny=70, fw=fw, damping=1500.0, fpml=10.0, npower=4.0, k_max_pml=kmax, dt=1e-3)
    vp, vs, rho = homogeneous(p.nx, p.ny)
    o = orc.Oracle(p, vp, vs, rho, prec="f64")
    for axis, n, clamp in (("x", p.nx, True), ("y", p.ny, False)):
        ref = cpml_closed_form(n, fw, p.dh, p.dt, p.damping, p.fpml, p.npower, kmax, clamp)
        base = 0 if axis == "x" else 10
        mapping = {"d": 0, "K": 1, "al": 2, "a": 3, "b": 4, "dh": 5, "Kh": 6, "alh": 7, "ah": 8, "bh": 9}
        for key, k in mapping.items():
            got = o.pml(base + k)
            # float32 inputs (dh, dt, damping...) enter the oracle's double math
            assert np.allclose(got, ref[key], rtol=2e-6, atol=1e-9), (axis, key)
        # shape invariants: damping grows toward the outer edge, b in (0, 1], a <= 0
        d = o.pml(base + 0)
        assert np.all(np.diff(d[:fw]) <= 0) and np.all(np.diff(d[fw:]) >= 0)
        b = o.pml(base + 4)
        assert np.all((b > 0) & (b <= 1))
        assert np.all(o.pml(base + 3) <= 0)


def test_cpml_alpha_clamped_only_in_x():
    p = params(nx=60, ny=60, fw=10, fpml=10.0)
    vp, vs, rho = homogeneous(p.nx, p.ny)
    o = orc.Oracle(p, vp, vs, rho, prec="f64")
    assert o.pml(7).min() >= 0.0          # alpha_x_half clamped
    assert o.pml(17)[-1] < 0.0            # alpha_y_half at the outermost half point is not


# --------------------------------------------------------------------------
# update_v / update_s stencils: exact on polynomials (Taylor coefficients)
# --------------------------------------------------------------------------
def _padded(o):
    nd = o.nd
    return np.zeros((o.p.ny + 2 * nd, o.p.nx + 2 * nd))


def _coords(o):
    nd = o.nd
    j = np.arange(-nd + 1, o.p.ny + nd + 1)[:, None].astype(float)
    i = np.arange(-nd + 1, o.p.nx + nd + 1)[None, :].astype(float)
    return j, i


@pytest.mark.parametrize("fdorder", [2, 4, 6, 8, 10, 12])
def test_velocity_stencils_exact_on_polynomials(fdorder):
    """sxx at (i, j), sxy at (i+1/2, j+1/2), syy at (i, j); vx at (i+1/2, j),
    vy at (i, j+1/2).  With polynomials of degree 2n-1 the Taylor stencils are
    exact: vx = DT/rho (d_x sxx + d_y sxy), vy = DT/rho (d_x sxy + d_y syy)."""
    n = fdorder // 2
    p = params(nx=40, ny=36, nt=5, fdorder=fdorder, maxrelerror=0, dh=1.0, dt=1e-3)
    vp, vs, rho = homogeneous(p.nx, p.ny, vp=1.0, vs=0.5, rho=2.0)
    o = orc.Oracle(p, vp, vs, rho, prec="f64")
    j, i = _coords(o)
    deg = 2 * n - 1
    c = [0.3, -0.2, 0.05, 0.01, -0.002, 3e-4, -2e-5, 1e-6, -3e-8, 1e-9, -2e-11, 1e-12][: deg + 1]
    P = np.polynomial.Polynomial(c)
    dP = P.deriv()
    s = 0.07  # keep the polynomial values moderate on the grid
    f = lambda x: P(s * x)
    df = lambda x: s * dP(s * x)
    x_int, y_int = i, j
    o.set_field(orc.ORC_SXX, f(x_int) + 0 * y_int)
    o.set_field(orc.ORC_SYY, f(2.0 * y_int) + 0 * x_int)
    o.set_field(orc.ORC_SXY, f(x_int + 0.5) + f(1.5 * (y_int + 0.5)))
    o.step(1)
    vx = o.field(orc.ORC_VX)
    vy = o.field(orc.ORC_VY)
    nd = o.nd
    sl = (slice(nd - 1 + n + 1, nd - 1 + p.ny - n), slice(nd - 1 + n + 1, nd - 1 + p.nx - n))
    dt_rho = float(np.float32(p.dt)) / 2.0  # DT enters as the float32 parameter
    exp_vx = dt_rho * (df(x_int + 0.5) + 1.5 * df(1.5 * (y_int + 0.5 - 0.5)))
    exp_vy = dt_rho * (df(x_int + 0.5 - 0.5) * 1.0 + 2.0 * df(2.0 * (y_int + 0.5)))
    # the coefficient table is float32 (as in the reference), so the stencil is
    # exact up to the rounding of its coefficients (1e-8..1e-7 relative after cancellation)
    assert np.allclose(vx[sl], (exp_vx + 0 * vx)[sl], rtol=1e-6, atol=0)
    assert np.allclose(vy[sl], (exp_vy + 0 * vy)[sl], rtol=1e-6, atol=0)


@pytest.mark.parametrize("fdorder", [2, 6, 8, 12])
def test_stress_stencils_exact_on_polynomials(fdorder):
    """vx at (i+1/2, j), vy at (i, j+1/2) polynomial; one stress sweep gives
    sxx = DT[(lam+2mu) d_x vx + lam d_y vy], syy = DT[lam d_x vx + (lam+2mu) d_y vy],
    sxy = DT mu (d_y vx + d_x vy) at (i+1/2, j+1/2)."""
    n = fdorder // 2
    p = params(nx=40, ny=36, nt=5, fdorder=fdorder, maxrelerror=0, dh=1.0, dt=1e-3)
    vpv, vsv, rhov = 2.0, 1.0, 3.0
    vp, vs, rho = homogeneous(p.nx, p.ny, vp=vpv, vs=vsv, rho=rhov)
    mu, lam = rhov * vsv ** 2, rhov * (vpv ** 2 - 2 * vsv ** 2)
    o = orc.Oracle(p, vp, vs, rho, prec="f64")
    j, i = _coords(o)
    deg = 2 * n - 1
    c = [0.1, 0.2, -0.05, 0.01, -0.002, 3e-4, -2e-5, 1e-6, -3e-8, 1e-9, -2e-11, 1e-12][: deg + 1]
    P = np.polynomial.Polynomial(c)
    dP = P.deriv()
    s = 0.06
    f = lambda x: P(s * x)
    df = lambda x: s * dP(s * x)
    o.set_field(orc.ORC_VX, f(i + 0.5) + f(0.8 * j))
    o.set_field(orc.ORC_VY, f(1.3 * i) + f(j + 0.5))
    o.step(1)
    nd = o.nd
    sl = (slice(nd - 1 + n + 1, nd - 1 + p.ny - n), slice(nd - 1 + n + 1, nd - 1 + p.nx - n))
    dvx_dx = df(i + 0.0 + 0 * j)            # at integer points
    dvy_dy = df(j + 0.0 + 0 * i)
    dvx_dy = 0.8 * df(0.8 * (j + 0.5)) + 0 * i   # at (i+1/2, j+1/2)
    dvy_dx = 1.3 * df(1.3 * (i + 0.5)) + 0 * j
    dt = float(np.float32(p.dt))
    for fid, ref in ((orc.ORC_SXX, dt * ((lam + 2 * mu) * dvx_dx + lam * dvy_dy)),
                     (orc.ORC_SYY, dt * (lam * dvx_dx + (lam + 2 * mu) * dvy_dy)),
                     (orc.ORC_SXY, dt * mu * (dvx_dy + dvy_dx))):
        assert np.allclose(o.field(fid)[sl], ref[sl], rtol=1e-7, atol=1e-10 * np.abs(ref[sl]).max()), fid


# --------------------------------------------------------------------------
# exact symmetries of the discrete scheme (FP32, bit for bit)
# --------------------------------------------------------------------------
def explosive_at(nx, ny, i0, j0, nt, dt, fc=25.0, dh=10.0):
    src = np.zeros((8, 1), np.float32)
    src[0, 0], src[1, 0], src[4, 0], src[5, 0], src[7, 0] = i0, j0, fc, 1.0, 1
    sig = synth.ricker(nt, dt, fc, amp=1e3)[None, :]
    return src, sig


def test_mirror_symmetry_about_the_source():
    """Homogeneous medium, explosive source on the centre column: the wavefield
    is mirror symmetric in x exactly (vx, sxy antisymmetric about the half
    points, vy, sxx, syy symmetric), as long as the edges have not been
    reached (stencil light cone: 2 FDORDER/2 cells per step)."""
    nx, ny, nt = 161, 101, 16
    p = params(nx=nx, ny=ny, nt=nt, dt=1.5e-3)
    vp, vs, rho = homogeneous(nx, ny)
    i0, j0 = (nx + 1) // 2, 45
    o = orc.Oracle(p, vp, vs, rho, prec="f32")
    o.set_sources(*explosive_at(nx, ny, i0, j0, nt, p.dt))
    o.forward()
    nd = o.nd
    col = lambda i: i + nd - 1
    for fid, sign, half in ((orc.ORC_VX, -1, True), (orc.ORC_SXY, -1, True), (orc.ORC_VY, 1, False),
                            (orc.ORC_SXX, 1, False), (orc.ORC_SYY, 1, False)):
        a = o.field(fid)
        assert np.abs(a).max() > 0
        for i in range(1, nx + 1):
            im = (2 * i0 - 1 - i) if half else (2 * i0 - i)
            if 1 <= im <= nx:
                assert np.array_equal(a[:, col(i)], sign * a[:, col(im)]), (fid, i)


def test_transpose_symmetry_square_grid():
    """Square homogeneous grid, source on the diagonal, no free surface or
    CPML: swapping x and y maps vx<->vy, sxx<->syy, sxy<->sxy exactly."""
    n, nt = 101, 14
    p = params(nx=n, ny=n, nt=nt, dt=1.5e-3)
    vp, vs, rho = homogeneous(n, n)
    o = orc.Oracle(p, vp, vs, rho, prec="f32")
    o.set_sources(*explosive_at(n, n, 51, 51, nt, p.dt))
    o.forward()
    assert np.array_equal(o.field(orc.ORC_VX), o.field(orc.ORC_VY).T)
    assert np.array_equal(o.field(orc.ORC_SXX), o.field(orc.ORC_SYY).T)
    assert np.array_equal(o.field(orc.ORC_SXY), o.field(orc.ORC_SXY).T)


# --------------------------------------------------------------------------
# discrete reciprocity (FP64): rho_x(B) vx_B[force x at A] = rho_x(A) vx_A[force x at B]
# --------------------------------------------------------------------------
def force_run(p, model, at, rec, ftype, sig, prec="f64"):
    o = orc.Oracle(p, *model, prec=prec)
    src = np.zeros((8, 1), np.float32)
    src[0, 0], src[1, 0], src[4, 0], src[5, 0], src[7, 0] = at[0], at[1], 20.0, 1.0, ftype
    o.set_sources(src, sig)
    o.set_receivers(np.array([[rec[0]], [rec[1]]], np.int32))
    o.forward()
    return o


@pytest.mark.parametrize("fdorder", [4, 8])
def test_reciprocity_heterogeneous(fdorder):
    """The velocity-stress operator is skew-adjoint in the energy inner
    product (the backward stencils are the negative transposes of the forward
    ones, the stiffness is symmetric), so force-to-velocity Green's functions
    are reciprocal with the density weights of the velocity points.  Rigid
    edges, heterogeneous model, FP64."""
    nx, ny, nt = 70, 60, 260
    p = params(nx=nx, ny=ny, nt=nt, fdorder=fdorder, dt=8e-4, seismo=1)
    vp, vs, rho = synth.layered_model(nx, ny, 10.0, seed=5, sigma=3.0)
    model = (vp, vs, rho)
    sig = synth.ricker(nt, p.dt, 20.0, amp=1.0)[None, :]
    A, B = (20, 18), (47, 40)
    o = orc.Oracle(p, *model, prec="f64")
    nd = o.nd
    rip = o.field(orc.ORC_RIP)
    rjp = o.field(orc.ORC_RJP)
    rx = lambda c: 1.0 / rip[c[1] + nd - 1, c[0] + nd - 1]
    ry = lambda c: 1.0 / rjp[c[1] + nd - 1, c[0] + nd - 1]
    # x force at A -> vx at B  vs  x force at B -> vx at A
    ab = force_run(p, model, A, B, 2, sig).seismogram(0)[0]
    ba = force_run(p, model, B, A, 2, sig).seismogram(0)[0]
    assert np.abs(ab).max() > 0
    assert rel_l2(rx(B) * ab, rx(A) * ba) < 1e-10
    # y force at A -> vx at B  vs  x force at B -> vy at A
    ab = force_run(p, model, A, B, 3, sig).seismogram(0)[0]
    ba = force_run(p, model, B, A, 2, sig).seismogram(1)[0]
    assert rel_l2(rx(B) * ab, ry(A) * ba) < 1e-10


# --------------------------------------------------------------------------
# free surface
# --------------------------------------------------------------------------
def test_free_surface_invariants():
    """After every step: syy = 0 on row 1 and the ghost rows are the odd
    images syy[1-m] = -syy[1+m], sxy[1-m] = -sxy[m]."""
    nx, ny, nt = 60, 40, 60
    p = params(nx=nx, ny=ny, nt=nt, free_surf=1, fw=8, dt=1e-3)
    vp, vs, rho = synth.layered_model(nx, ny, 10.0, seed=2, sigma=3.0)
    o = orc.Oracle(p, vp, vs, rho, prec="f32")
    o.set_sources(*explosive_at(nx, ny, 30, 3, nt, p.dt))
    nd = o.nd
    r = lambda jj: jj + nd - 1
    for nt_ in range(1, 41):
        o.step(nt_)
        syy, sxy = o.field(orc.ORC_SYY), o.field(orc.ORC_SXY)
        assert np.all(syy[r(1), nd:nd + nx] == 0)
        for m in range(1, p.fdorder // 2 + 1):
            assert np.array_equal(syy[r(1 - m), nd:nd + nx], -syy[r(1 + m), nd:nd + nx])
            assert np.array_equal(sxy[r(1 - m), nd:nd + nx], -sxy[r(m), nd:nd + nx])
    assert np.abs(o.field(orc.ORC_SXX)).max() > 0


def test_rayleigh_wave_speed():
    """Vertical force at the free surface of a Poisson solid (vs = vp/sqrt 3):
    the surface wave moves at c_R = 0.9194 vs (Lamb's problem).  Measured from
    the arrival of the peak |vy| at two surface receivers."""
    nx, ny, dh = 320, 120, 5.0
    vpv = 3000.0
    vsv = vpv / math.sqrt(3.0)
    dt = 6e-4
    nt = 1250
    p = params(nx=nx, ny=ny, nt=nt, dh=dh, dt=dt, free_surf=1, fw=20, seismo=1, fpml=15.0,
               damping=3000.0)
    vp, vs, rho = homogeneous(nx, ny, vp=vpv, vs=vsv, rho=2000.0)
    o = orc.Oracle(p, vp, vs, rho, prec="f64")
    src = np.zeros((8, 1), np.float32)
    src[0, 0], src[1, 0], src[4, 0], src[5, 0], src[7, 0] = 60, 1, 15.0, 1.0, 3
    o.set_sources(src, synth.ricker(nt, dt, 15.0, amp=1.0)[None, :])
    r1, r2 = 110, 230
    o.set_receivers(np.array([[r1, r2], [1, 1]], np.int32))
    o.forward()
    s = np.abs(o.seismogram(1))

    def peak_time(a):
        k = int(np.argmax(a))
        y0, y1, y2 = a[k - 1], a[k], a[k + 1]
        return (k + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)) * dt

    c = (r2 - r1) * dh / (peak_time(s[1]) - peak_time(s[0]))
    cr = 0.9194 * vsv
    assert abs(c / cr - 1.0) < 0.02, (c, cr)


# --------------------------------------------------------------------------
# CPML absorbs; the scheme is stable
# --------------------------------------------------------------------------
def interior_energy(o, margin):
    nd = o.nd
    sl = (slice(nd - 1 + margin + 1, nd - 1 + o.p.ny - margin), slice(nd - 1 + margin + 1, nd - 1 + o.p.nx - margin))
    return sum(float(np.sum(o.field(f)[sl].astype(np.float64) ** 2)) for f in (orc.ORC_VX, orc.ORC_VY))


def test_cpml_absorbs_outgoing_waves():
    """A pulse leaving a homogeneous box: with CPML on all sides the kinetic
    energy left in the interior after the wave has crossed the boundary is a
    small fraction of what rigid edges keep."""
    nx = ny = 120
    nt = 700
    res = {}
    for fw in (0, 20):
        p = params(nx=nx, ny=ny, nt=nt, fw=fw, dt=1e-3, fpml=25.0, damping=3000.0)
        vp, vs, rho = homogeneous(nx, ny)
        o = orc.Oracle(p, vp, vs, rho, prec="f64")
        o.set_sources(*explosive_at(nx, ny, 60, 60, nt, p.dt, fc=25.0))
        o.forward(120)
        e_mid = interior_energy(o, 20)
        o.forward(nt)
        res[fw] = (e_mid, interior_energy(o, 20))
    e_rigid, e_pml = res[0][1], res[20][1]
    assert res[20][0] > 0
    assert e_pml < 0.02 * e_rigid, (e_pml, e_rigid)
    assert e_pml < 0.02 * res[20][0]


def test_long_run_stays_bounded():
    nx, ny, nt = 50, 40, 4000
    p = params(nx=nx, ny=ny, nt=nt, dt=1.2e-3, fw=0)
    vp, vs, rho = synth.layered_model(nx, ny, 10.0, seed=9, sigma=2.0)
    o = orc.Oracle(p, vp, vs, rho, prec="f64")
    o.set_sources(*explosive_at(nx, ny, 25, 20, nt, p.dt))
    o.forward(300)
    e1 = interior_energy(o, 0)
    o.forward(nt)
    e2 = interior_energy(o, 0)
    assert np.isfinite(e2) and e2 < 10 * e1


# --------------------------------------------------------------------------
# sources, receivers, wavelet, residuals: known answers
# --------------------------------------------------------------------------
def test_ricker_closed_form():
    nt, dt, fc, ts = 400, 1e-3, 12.0, 0.02
    got = orc.ricker(nt, dt, ts, fc, 2.5)
    t = np.arange(1, nt + 1) * np.float32(dt)
    tau = math.pi * (t - 1.5 / fc - ts) / (1.5 / fc)
    ref = (1 - 4 * tau ** 2) * np.exp(-2 * tau ** 2) * 2.5
    assert np.allclose(got, ref, rtol=0, atol=5e-6)
    k = int(np.argmax(got))
    assert abs((k + 1) * dt - (1.5 / fc + ts)) <= dt
    # the product-side generator follows the same wavelet.c arithmetic
    assert np.array_equal(synth.ricker(nt, dt, fc, tshift=ts, amp=2.5), got)


def test_explosive_source_injects_the_time_derivative():
    """psource: sxx and syy gain (s[nt+1]-s[nt-1])/DT at the source cell
    (one-sided at nt=1 and nt=NT)."""
    nx, ny, nt = 30, 30, 6
    p = params(nx=nx, ny=ny, nt=nt, dt=1e-3, seismo=0)
    vp, vs, rho = homogeneous(nx, ny)
    o = orc.Oracle(p, vp, vs, rho, prec="f32")
    sig = np.array([[0.5, 2.0, -1.0, 4.0, 8.0, 16.0]], np.float32)
    src = np.zeros((8, 1), np.float32)
    src[0, 0], src[1, 0], src[4, 0], src[7, 0] = 15, 15, 10.0, 1
    o.set_sources(src, sig)
    nd = o.nd
    o.step(1)
    assert o.field(orc.ORC_SXX)[15 + nd - 1, 15 + nd - 1] == np.float32(2.0) / np.float32(1e-3)
    o.zero()
    o.step(3)
    assert o.field(orc.ORC_SYY)[15 + nd - 1, 15 + nd - 1] == (np.float32(4.0) - np.float32(2.0)) / np.float32(1e-3)
    o.zero()
    o.step(6)
    assert o.field(orc.ORC_SXX)[15 + nd - 1, 15 + nd - 1] == -np.float32(8.0) / np.float32(1e-3)


def test_receivers_sample_the_fields():
    nx, ny, nt = 40, 30, 8
    p = params(nx=nx, ny=ny, nt=nt, seismo=4, dt=1e-3)
    vp, vs, rho = synth.layered_model(nx, ny, 10.0, seed=4, sigma=2.0)
    o = orc.Oracle(p, vp, vs, rho, prec="f32")
    o.set_sources(*explosive_at(nx, ny, 20, 15, nt, p.dt))
    rec = np.array([[18, 21, 25], [15, 12, 17]], np.int32)
    o.set_receivers(rec)
    nd = o.nd
    for nt_ in range(1, nt + 1):
        o.step(nt_)
        vx, vy = o.field(orc.ORC_VX), o.field(orc.ORC_VY)
        sxx, syy = o.field(orc.ORC_SXX), o.field(orc.ORC_SYY)
        for t in range(3):
            jj, ii = rec[1, t] + nd - 1, rec[0, t] + nd - 1
            assert o.seismogram(0)[t, nt_ - 1] == vx[jj, ii]
            assert o.seismogram(1)[t, nt_ - 1] == vy[jj, ii]
            assert o.seismogram(2)[t, nt_ - 1] == -sxx[jj, ii] - syy[jj, ii]


def test_l2_residual_known_answer():
    """Time-integrated, time-reversed residual: syn = 1, obs = 0 except that
    the first observed sample is replaced by the synthetic one, so the
    integral after k samples is DT*(k-1), stored at index NT-k."""
    nx, ny, nt = 20, 20, 50
    p = params(nx=nx, ny=ny, nt=nt, dt=2e-3, seismo=1, mode=1)
    vp, vs, rho = homogeneous(nx, ny)
    o = orc.Oracle(p, vp, vs, rho, prec="f64")
    o.set_receivers(np.array([[5, 9], [5, 5]], np.int32))
    # set synthetic seismograms through a zero run, then residual against obs
    o.set_sources(np.zeros((8, 0), np.float32), np.zeros((0, nt), np.float32))
    o.forward()
    syn = o.seismogram(0)
    obs = (syn - 1.0).astype(np.float32)   # residual 1 on every sample but the first
    l2 = o.residuals(obs, obs, None, 1)
    d = o.adjoint_sources(0)[0]
    k = np.arange(1, nt + 1)
    expect = np.float32(2e-3) * (k - 1)
    assert np.allclose(d[nt - k], expect, rtol=1e-6, atol=1e-9)
    assert l2 == pytest.approx(2 * 2 * 0.5 * float(np.sum(expect.astype(np.float64) ** 2)), rel=1e-6)


# --------------------------------------------------------------------------
# adjoint gradient vs finite differences of the misfit
# --------------------------------------------------------------------------
def _bump(nx, ny, ci, cj, s):
    i = np.arange(nx)[:, None]
    j = np.arange(ny)[None, :]
    return np.exp(-((i - ci) ** 2 + (j - cj) ** 2) / (2.0 * s * s))


def _gradient_setup(dtinv=1, free_surf=0, recrow=14, invmat1=3, nx=64, ny=56, nt=700, fc=8.0, zero_data=False):
    """A shot in a model that is homogeneous (vp 2000, vs 2000/sqrt(3),
    rho 2000) around the source and receiver line at row `recrow` and
    smoothly stiffer below; observed data from the same model with three
    +5% Gaussian anomalies (vp, vs, rho).  Returns the parameters, the
    start model in the INVMAT1 parameterisation, and misfit(model, grad)
    that runs the FP64 oracle (forward, L2 residual of calc_res.c LNORM 2,
    and with grad the adjoint run and ass_gradPSV)."""
    p = params(nx=nx, ny=ny, nt=nt, dh=10.0, dt=1e-3, fdorder=8, maxrelerror=0, fw=10, free_surf=free_surf,
               seismo=1, mode=1, dtinv=dtinv, fpml=fc, damping=2500.0, ts=1.0 / fc, invmat1=invmat1)
    jj = np.arange(ny)[None, :] * np.ones((nx, 1))
    f = np.clip((jj - recrow - 4) / 20.0, 0.0, 1.0)
    vp, vs, rho = 2000.0 * (1 + 0.2 * f), 2000.0 / np.sqrt(3.0) * (1 + 0.15 * f), 2000.0 * (1 + 0.1 * f)

    def model(vp, vs, rho):
        if invmat1 == 1:
            return [a.astype(np.float32) for a in (vp, vs, rho)]
        mu = rho * vs * vs
        return [(rho * vp * vp - 2 * mu).astype(np.float32), mu.astype(np.float32), rho.astype(np.float32)]

    m0 = model(vp, vs, rho)
    mt = model(vp * (1 + 0.05 * _bump(nx, ny, nx // 2, ny - 18, 5)),
               vs * (1 + 0.05 * _bump(nx, ny, nx // 2 - 6, ny - 18, 5)),
               rho * (1 + 0.05 * _bump(nx, ny, nx // 2 + 6, ny - 20, 5)))
    src = np.zeros((8, 1), np.float32)
    src[0, 0], src[1, 0], src[4, 0], src[5, 0], src[7, 0] = nx // 2, recrow, fc, 1.0, 1
    sig = synth.ricker(nt, p.dt, fc, amp=1e3)[None, :]
    xs = list(range(14, nx - 13, 2))
    rec = np.array([xs, [recrow] * len(xs)], np.int32)

    def runner(m):
        o = orc.Oracle(p, *m, prec="f64")
        o.set_sources(src, sig)
        o.set_receivers(rec)
        return o

    t = runner(mt)
    t.forward()
    obs = (t.seismogram(0).astype(np.float32), t.seismogram(1).astype(np.float32))
    if zero_data:
        obs = tuple(np.zeros_like(o) for o in obs)

    def misfit(m, grad=False):
        o = runner([np.asarray(a, np.float32) for a in m])
        o.forward()
        l2 = o.residuals(obs[0], obs[1], None, 1)
        if not grad:
            return l2, None
        o.adjoint()
        o.assemble()
        return l2, [o.gradient(k).astype(np.float64) for k in (3, 4, 5)]

    return p, m0, misfit, float(rho[0, recrow])


def _directional_fd(misfit, m0, w, d):
    mp = [a.astype(np.float64) for a in m0]
    mm = [a.astype(np.float64) for a in m0]
    mp[w] = mp[w] + d
    mm[w] = mm[w] - d
    return (misfit(mp)[0] - misfit(mm)[0]) / 2.0


# The scale the discretisation fixes.  With data = particle velocity, the
# exact gradient of J = 1/2 sum R^2 (R the time-integrated residual,
# calc_res.c LNORM 2) takes the adjoint field of the velocity-stress
# scheme driven by the force DT*sum_{k>=n} R_k / rho_rec at each receiver
# (the transposed velocity update carries b = 1/rho there).  DENISE adds R
# itself to vx/vy (update_v_PML_PSV.c:755-768) and integrates the adjoint
# velocity into vxp1 (DT*vx, :455-457), which gives the field driven by
# DT*sum R: rho_rec times the exact adjoint.  It correlates every DTINV-th
# step without a DTINV weight (psv.c:313-387), 1/DTINV of the full sum for
# these band-limited fields.  Hence gradient = (rho_rec / DTINV) * dJ/dm,
# i.e. finite difference / gradient = DTINV / rho_rec, the same for every
# parameter.  A wrong DT, 2 or sign in ass_gradPSV, or in the correlation,
# moves this ratio far outside the tolerances below.
@pytest.mark.parametrize("dtinv,free_surf", [(1, 0), (2, 0), (1, 1)])
def test_lambda_gradient_equals_scaled_misfit_derivative(dtinv, free_surf):
    """INVMAT1 3: the lambda gradient at single cells (lambda enters only
    the cell's own normal-stress update, where DENISE correlates, so no
    staggering shift) against central differences of the misfit: each
    ratio within 6% and their mean within 2.5% of DTINV / rho_rec."""
    recrow = 3 if free_surf else 14
    p, m0, misfit, rho_rec = _gradient_setup(dtinv=dtinv, free_surf=free_surf, recrow=recrow)
    _, g = misfit(m0, grad=True)
    gl = g[0].copy()
    gl[:12], gl[-12:], gl[:, :recrow + 8], gl[:, -12:] = 0, 0, 0, 0
    ratios = []
    for k in np.argsort(-np.abs(gl).ravel())[:40:5]:
        ci, cj = np.unravel_index(k, gl.shape)
        d = np.zeros(gl.shape)
        d[ci, cj] = 1e-3 * float(m0[0][ci, cj])
        ratios.append(_directional_fd(misfit, m0, 0, d) / (g[0][ci, cj] * d[ci, cj]) * rho_rec / dtinv)
    ratios = np.array(ratios)
    print("lambda: FD / gradient x rho_rec / DTINV =", np.round(ratios, 4))
    assert np.all(np.abs(ratios - 1.0) < 0.06)
    assert abs(ratios.mean() - 1.0) < 0.025


@pytest.mark.parametrize("dtinv,free_surf", [(1, 0), (2, 0), (1, 1)])
def test_mu_and_rho_gradients_equal_scaled_misfit_derivative(dtinv, free_surf):
    """INVMAT1 3: the mu and rho gradients along smooth directions (the
    gradient itself smoothed over 2-3 cells in three windows: mu and rho
    enter averages over neighbouring cells, av_mue.c, av_rho.c, and DENISE
    correlates them at staggered points, so single cells are off by the
    half-cell shift) against central differences of the misfit: every
    ratio within 8% of the same DTINV / rho_rec as lambda."""
    from synth import gaussian_smooth as gaussian_filter
    recrow = 3 if free_surf else 14
    p, m0, misfit, rho_rec = _gradient_setup(dtinv=dtinv, free_surf=free_surf, recrow=recrow)
    nx, ny = p.nx, p.ny
    _, g = misfit(m0, grad=True)
    for w, name in ((1, "mu"), (2, "rho")):
        ratios = []
        for s in (2.0, 3.0):
            for region in ((slice(12, 52), slice(24, 44)), (slice(14, 34), slice(22, 46)),
                           (slice(30, 52), slice(22, 46))):
                d = np.zeros((nx, ny))
                d[region] = gaussian_filter(g[w], s)[region]
                d = gaussian_filter(d, s)
                d *= 1e-3 * float(np.abs(m0[w]).max()) / np.abs(d).max()
                ratios.append(_directional_fd(misfit, m0, w, d) / float(np.sum(g[w] * d)) * rho_rec / dtinv)
        ratios = np.array(ratios)
        print(f"{name}: FD / gradient x rho_rec / DTINV =", np.round(ratios, 4))
        assert np.all(np.abs(ratios - 1.0) < 0.08), name


def test_vp_vs_rho_gradients_follow_the_chain_rule():
    """ass_gradPSV's INVMAT1 1 gradients are the lambda / mu / rho ones
    (pinned above) transformed by the chain rule of lambda = rho (vp^2 -
    2 vs^2), mu = rho vs^2: d/dvp = 2 rho vp d/dlambda, d/dvs = -4 rho vs
    d/dlambda + 2 rho vs d/dmu, d/drho|vp,vs = (vp^2 - 2 vs^2) d/dlambda +
    vs^2 d/dmu + d/drho|lambda,mu.  Same physical model in both
    parameterisations, FP64 oracle; the parameter files differ only by the
    float32 rounding of lambda and mu, which the data residual would
    amplify where synthetics and data nearly cancel, so the misfit here is
    against zero data (residual = the time-integrated synthetics)."""
    _, m1, misfit1, _ = _gradient_setup(invmat1=1, zero_data=True)
    _, m3, misfit3, _ = _gradient_setup(invmat1=3, zero_data=True)
    _, g1 = misfit1(m1, grad=True)
    _, g3 = misfit3(m3, grad=True)
    vp, vs, rho = (a.astype(np.float64) for a in m1)
    want = [2 * rho * vp * g3[0],
            -4 * rho * vs * g3[0] + 2 * rho * vs * g3[1],
            (vp * vp - 2 * vs * vs) * g3[0] + vs * vs * g3[1] + g3[2]]
    for k, name in enumerate(("vp", "vs", "rho")):
        err = rel_l2(g1[k], want[k])
        print(f"{name}: chain rule relative L2 {err:.2e}")
        assert np.abs(want[k]).max() > 0
        assert err < 1e-5, name


# --------------------------------------------------------------------------
# viscoelastic update: the generalized standard linear solid
# --------------------------------------------------------------------------
def _linear_velocity(o, a, b, c, d, dh):
    """vx = a x + c y at (i+1/2, j), vy = b y + d x at (i, j+1/2): uniform
    strain rates vxx = a, vyy = b, vxy = c, vyx = d."""
    j, i = _coords(o)
    vx = a * (i + 0.5) * dh + c * j * dh
    vy = b * (j + 0.5) * dh + d * i * dh
    return vx, vy


def test_visco_constant_strain_rate_matches_gsls():
    """Held at constant strain rates, the memory-variable update integrates
    the GSLS relaxation ODE  r' = -w_l (r + M tau edot): with relaxed modulus
    M_R = M_ref / (1 + tau sum_l w0^2 tau_l^2 / (1 + w0^2 tau_l^2))
    (the tau-method's real modulus at the reference frequency w0 = 2 pi FL[1])
    the stress is
        sigma(t) = M_R edot t + M_R tau edot sum_l (1 - exp(-w_l t)) / w_l,
    applied per component (sxx: (lambda+2mu)(a+b) - 2 mu b, syy: ... - 2 mu a,
    sxy: mu (c+d)).  The velocity field is reset every step, so only the
    local update is exercised; trapezoidal time stepping leaves an
    O((w DT)^2) error."""
    nx = ny = 40
    dh, dt, nsteps = 10.0, 1e-4, 400
    fl = [5.0, 40.0]
    tau_s, tau_p = 0.08, 0.05
    p = params(nx=nx, ny=ny, nt=nsteps, dh=dh, dt=dt, L=2, fl=fl, fdorder=8, maxrelerror=0)
    vpv, vsv, rhov = 3000.0, 1700.0, 2000.0
    vp, vs, rho = homogeneous(nx, ny, vp=vpv, vs=vsv, rho=rhov)
    taus = np.full((nx, ny), tau_s, np.float32)
    taup = np.full((nx, ny), tau_p, np.float32)
    o = orc.Oracle(p, vp, vs, rho, taus, taup, prec="f64")
    a, b, c, d = 2.0e-3, -0.7e-3, 1.1e-3, 0.4e-3
    vx, vy = _linear_velocity(o, a, b, c, d, dh)
    for nt_ in range(1, nsteps + 1):
        o.set_field(orc.ORC_VX, vx)
        o.set_field(orc.ORC_VY, vy)
        o.step(nt_)
    t = nsteps * float(np.float32(dt))
    w = [2 * math.pi * f for f in fl]
    w0 = w[0]
    S = sum((w0 / wl) ** 2 / (1 + (w0 / wl) ** 2) for wl in w)
    mu = rhov * float(np.float32(vsv)) ** 2 / (1 + float(np.float32(tau_s)) * S)
    pm = rhov * float(np.float32(vpv)) ** 2 / (1 + float(np.float32(tau_p)) * S)
    relax = sum((1 - math.exp(-wl * t)) / wl for wl in w)
    ts_, tp_ = float(np.float32(tau_s)), float(np.float32(tau_p))
    exp_sxy = mu * (c + d) * t + mu * ts_ * (c + d) * relax
    exp_sxx = (pm * (a + b) - 2 * mu * b) * t + (pm * tp_ * (a + b) - 2 * mu * ts_ * b) * relax
    exp_syy = (pm * (a + b) - 2 * mu * a) * t + (pm * tp_ * (a + b) - 2 * mu * ts_ * a) * relax
    nd = o.nd
    jc, ic = ny // 2 + nd - 1, nx // 2 + nd - 1
    got = {f: o.field(f)[jc, ic] for f in (orc.ORC_SXX, orc.ORC_SYY, orc.ORC_SXY)}
    # uniform away from the zero ghosts
    core = (slice(jc - 8, jc + 8), slice(ic - 8, ic + 8))
    for f in got:
        blk = o.field(f)[core]
        assert np.allclose(blk, got[f], rtol=1e-9, atol=0)
    for f, e in ((orc.ORC_SXX, exp_sxx), (orc.ORC_SYY, exp_syy), (orc.ORC_SXY, exp_sxy)):
        assert got[f] == pytest.approx(e, rel=2e-4), (f, got[f], e)
    # the memory terms matter at this tolerance: dropping them misses by far more
    assert abs(mu * ts_ * (c + d) * relax) > 50 * 2e-4 * abs(exp_sxy)


def _surface_visco_run(tau_s, tau_p, fl, ts, a, b, c, d, nsteps=400, dt=1e-4, dh=10.0):
    """Row 1 (free surface) of a homogeneous viscoelastic model held at
    uniform strain rates (velocities reset every step, so only the local
    stress update acts): returns (sxx, syy) at the surface centre after
    nsteps and the time."""
    nx, ny = 40, 40
    p = params(nx=nx, ny=ny, nt=nsteps, dh=dh, dt=dt, L=len(fl), fl=fl, fdorder=8, maxrelerror=0,
               free_surf=1, ts=ts)
    vp, vs, rho = homogeneous(nx, ny, vp=3000.0, vs=1700.0, rho=2000.0)
    taus = np.full((nx, ny), tau_s, np.float32)
    taup = np.full((nx, ny), tau_p, np.float32)
    o = orc.Oracle(p, vp, vs, rho, taus, taup, prec="f64")
    vx, vy = _linear_velocity(o, a, b, c, d, dh)
    for nt_ in range(1, nsteps + 1):
        o.set_field(orc.ORC_VX, vx)
        o.set_field(orc.ORC_VY, vy)
        o.step(nt_)
    nd = o.nd
    js, ic = nd, nx // 2 + nd - 1
    sxx, syy = o.field(orc.ORC_SXX), o.field(orc.ORC_SYY)
    # uniform along the surface away from the zero ghost columns
    assert np.allclose(sxx[js, ic - 8:ic + 8], sxx[js, ic], rtol=1e-9, atol=0)
    return sxx[js, ic], syy[js, ic], nsteps * float(np.float32(dt))


def _relaxed(m_ref, tau, w_ref, fl):
    """tau-method relaxed modulus whose real part at w_ref is m_ref
    (prepare_update_s_visc_PSV.c, surface_visc_PML_PSV.c:41-52)."""
    return m_ref / (1.0 + tau * sum((w_ref / (2 * math.pi * f)) ** 2 / (1 + (w_ref / (2 * math.pi * f)) ** 2)
                                    for f in fl))


def test_visco_free_surface_is_plane_stress_gsls():
    """Equal relaxation times (tau_s = tau_p = tau: every modulus relaxes
    with the same function psi(t) = 1 + tau sum_l exp(-t/tau_l)) and
    TS = 1/FL[1], so the surface's relaxed moduli (ws = 2 pi/TS,
    surface_visc_PML_PSV.c:42) are the interior's (ws = 2 pi FL[1]): by the
    correspondence principle the free-surface row (syy = 0) is a
    plane-stress GSLS, and under a constant horizontal strain rate a
        sxx(t) = E_R a (t + tau sum_l tau_l (1 - exp(-t/tau_l))),
        E_R = M_R - lambda_R^2 / M_R  (relaxed plane-stress modulus),
    whatever the vertical strain rate b (syy = 0 removes it).  Trapezoidal
    time stepping leaves an O((DT/tau_l)^2) error."""
    fl = [5.0, 40.0]
    tau = 0.06
    a, c, d = 2.0e-3, 1.1e-3, 0.4e-3
    runs = [_surface_visco_run(tau, tau, fl, 1.0 / fl[0], a, b, c, d) for b in (0.0, -0.7e-3, 1.3e-3)]
    sxx, syy, t = runs[0]
    w0 = 2 * math.pi * fl[0]
    tf = float(np.float32(tau))
    mu = _relaxed(2000.0 * float(np.float32(1700.0)) ** 2, tf, w0, fl)
    pm = _relaxed(2000.0 * float(np.float32(3000.0)) ** 2, tf, w0, fl)
    lam = pm - 2 * mu
    e_r = pm - lam * lam / pm
    relax = sum((1 - math.exp(-2 * math.pi * f * t)) / (2 * math.pi * f) for f in fl)
    expect = e_r * a * (t + tf * relax)
    print(f"surface sxx {sxx:.9e} plane-stress GSLS {expect:.9e}")
    assert sxx == pytest.approx(expect, rel=2e-4)
    for sxx_b, syy_b, _ in runs:
        assert syy_b == 0.0
        assert sxx_b == pytest.approx(sxx, rel=1e-9)
    # the relaxation matters at this tolerance: the elastic relaxed response misses by far more
    assert abs(e_r * a * tf * relax) > 50 * 2e-4 * abs(expect)


def test_visco_free_surface_general_tau_and_reference_frequency():
    """tau_s != tau_p and TS != 1/FL[1] (the surface's relaxed moduli at
    ws = 2 pi/TS differ from the interior's at 2 pi FL[1]).  Interior update
    plus surface correction (surface_visc_PML_PSV.c:136-170) integrate, per
    mechanism with tau_l = 1/(2 pi FL[l]),
        sxx' = M_i (a+b) - 2 mu_i b - l_s^2/M_s a - l_s b + sum_l p_l,
        p_l' = -(p_l + S)/tau_l,
        S = P_i tau_p (a+b) - 2 m_i tau_s b + (2 m_s tau_s - P_s tau_p)(l_s/M_s a + b),
    with unrelaxed moduli M = P (1 + L tau_p), mu = m (1 + L tau_s),
    l = M - 2 mu from the relaxed P, m of the interior (i) and surface (s)
    reference frequencies, so that
        sxx(t) = E t - S_a sum_l (t - tau_l (1 - exp(-t/tau_l))).
    Equal moduli (the previous test) cancel b; here a b-term remains."""
    fl = [5.0, 40.0]
    tau_s, tau_p = 0.08, 0.05
    ts = 0.1
    a, b, c, d = 2.0e-3, -0.7e-3, 1.1e-3, 0.4e-3
    sxx, syy, t = _surface_visco_run(tau_s, tau_p, fl, ts, a, b, c, d)
    L = len(fl)
    tsf, tpf = float(np.float32(tau_s)), float(np.float32(tau_p))
    rv, rp = 2000.0 * float(np.float32(1700.0)) ** 2, 2000.0 * float(np.float32(3000.0)) ** 2
    wi, ws = 2 * math.pi * fl[0], 2 * math.pi / float(np.float32(ts))
    m_i, p_i = _relaxed(rv, tsf, wi, fl), _relaxed(rp, tpf, wi, fl)
    m_s, p_s = _relaxed(rv, tsf, ws, fl), _relaxed(rp, tpf, ws, fl)
    M_i, mu_i = p_i * (1 + L * tpf), m_i * (1 + L * tsf)
    M_s, mu_s = p_s * (1 + L * tpf), m_s * (1 + L * tsf)
    l_s = M_s - 2 * mu_s
    rate = M_i * (a + b) - 2 * mu_i * b - l_s * l_s / M_s * a - l_s * b
    S = p_i * tpf * (a + b) - 2 * m_i * tsf * b + (2 * m_s * tsf - p_s * tpf) * (l_s / M_s * a + b)
    mem = sum(t - (1 - math.exp(-2 * math.pi * f * t)) / (2 * math.pi * f) for f in fl)
    expect = rate * t - S * mem
    print(f"surface sxx {sxx:.9e} closed form {expect:.9e}")
    assert syy == 0.0
    assert sxx == pytest.approx(expect, rel=2e-4)
    assert abs(S * mem) > 50 * 2e-4 * abs(expect)


def test_visco_with_zero_tau_is_the_elastic_update():
    """tau = 0 switches the relaxation off: the viscoelastic update, free
    surface included, reduces to the elastic one (FP64: rounding only)."""
    nx, ny, nt = 80, 56, 160
    vp, vs, rho = synth.layered_model(nx, ny, 10.0, seed=5, sigma=3.0)
    zero = np.zeros((nx, ny), np.float32)
    common = dict(nx=nx, ny=ny, nt=nt, dt=8e-4, free_surf=1, fw=10, seismo=4)
    pe = params(**common)
    pv = params(L=2, fl=[5.0, 40.0], **common)
    src, sig = explosive_at(nx, ny, 40, 4, nt, pe.dt)
    rec = synth.receiver_line(50.0, 20.0, 20, 20.0, 10.0)
    runs = []
    for p, taus in ((pe, None), (pv, zero)):
        o = orc.Oracle(p, vp, vs, rho, taus, taus, prec="f64")
        o.set_sources(src, sig)
        o.set_receivers(rec)
        o.forward()
        runs.append(o)
    for comp in range(5):
        assert rel_l2(runs[1].seismogram(comp), runs[0].seismogram(comp)) < 1e-10, comp
    for f in (orc.ORC_VX, orc.ORC_VY, orc.ORC_SXX, orc.ORC_SYY, orc.ORC_SXY):
        assert rel_l2(runs[1].field(f), runs[0].field(f)) < 1e-10, f


# --------------------------------------------------------------------------
# material averages and the one-rank periodic ghosts
# --------------------------------------------------------------------------
def _periodic(a_xmajor):
    """(NY+2, NX+2) array [j][i], j, i = 0..N+1, of an x-major (NX, NY)
    model extended periodically: row 0 = row NY, row NY+1 = row 1, and
    likewise the columns, corners included (what matcopy_elastic_PSV.c:32-126
    leaves on one rank, where every send goes to the process itself)."""
    return np.pad(np.asarray(a_xmajor, np.float64).T, 1, mode="wrap")


@pytest.mark.parametrize("invmat1", [1, 3])
def test_material_averages_known_answer(invmat1):
    """rip, rjp: 1 / arithmetic mean of rho at (i, i+1) and (j, j+1)
    (av_rho.c:11-12); uipjp: harmonic mean of mu = rho vs^2 (INVMAT1 1) or
    of mu itself (3) over the four cells around (i+1/2, j+1/2), 0 when one
    of them is fluid (av_mue.c:12-45); the last row and column average
    with row 1 and column 1 (periodic ghosts).  A 5 x 4 model of distinct
    values, so a swapped or shifted neighbour fails; FP64 oracle."""
    nx, ny = 5, 4
    rng = np.random.default_rng(17)
    vp = rng.uniform(2500, 4000, (nx, ny)).astype(np.float32)
    vs = rng.uniform(1200, 2000, (nx, ny)).astype(np.float32)
    rho = rng.uniform(1800, 2600, (nx, ny)).astype(np.float32)
    vs[2, 1] = 0.0                      # a fluid cell
    if invmat1 == 3:
        mu = (rho.astype(np.float64) * vs * vs).astype(np.float32)
        lam = (rho.astype(np.float64) * vp * vp).astype(np.float32) - 2 * mu
        m = (lam, mu, rho)
    else:
        m = (vp, vs, rho)
    p = params(nx=nx, ny=ny, nt=2, fw=0, free_surf=0, dt=1e-5, dh=50.0, invmat1=invmat1)
    o = orc.Oracle(p, *m, prec="f64")
    R = _periodic(rho)
    U = _periodic(m[1])
    if invmat1 == 1:
        U = R * U * U
    J, I = np.meshgrid(np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    rip = 1.0 / (0.5 * (R[J, I] + R[J, I + 1]))
    rjp = 1.0 / (0.5 * (R[J, I] + R[J + 1, I]))
    with np.errstate(divide="ignore"):
        uip = 4.0 / ((1.0 / U[J, I]) + (1.0 / U[J, I + 1]) + (1.0 / U[J + 1, I]) + (1.0 / U[J + 1, I + 1]))
    fluid = (U[J, I] == 0) | (U[J, I + 1] == 0) | (U[J + 1, I] == 0) | (U[J + 1, I + 1] == 0)
    uip[fluid] = 0.0
    assert fluid.sum() == 4             # the four half-points around the fluid cell
    nd = o.nd
    inner = (slice(nd, nd + ny), slice(nd, nd + nx))
    for fid, want in ((orc.ORC_RIP, rip), (orc.ORC_RJP, rjp), (orc.ORC_UIPJP, uip)):
        got = o.field(fid)[inner]
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=0, err_msg=str(fid))
    # the wrap is what is asked for: the last column / row differ from a
    # zero or a clamped ghost
    assert rip[0, nx - 1] != pytest.approx(1.0 / (0.5 * R[1, nx]), rel=1e-6)


def test_visco_material_averages_known_answer():
    """Viscoelastic: tausipjp = arithmetic mean of tau_s over the four cells
    (av_tau.c:11), periodic like mu (matcopy_PSV), enters
    fipjp = uipjp / (1 + S tausipjp) DT (1 + L tausipjp) with S =
    sum_l w^2 tau_l^2 / (1 + w^2 tau_l^2), w = 2 pi FL[1]
    (prepare_update_s_visc_PSV.c:29-50); f and g the same with the cell's
    own tau_s, tau_p and rho vs^2, rho vp^2."""
    nx, ny = 5, 4
    rng = np.random.default_rng(23)
    vp = rng.uniform(2500, 4000, (nx, ny)).astype(np.float32)
    vs = rng.uniform(1200, 2000, (nx, ny)).astype(np.float32)
    rho = rng.uniform(1800, 2600, (nx, ny)).astype(np.float32)
    taus = rng.uniform(0.02, 0.1, (nx, ny)).astype(np.float32)
    taup = rng.uniform(0.02, 0.1, (nx, ny)).astype(np.float32)
    fl = [5.0, 40.0]
    dt = 1e-5
    p = params(nx=nx, ny=ny, nt=2, fw=0, free_surf=0, dt=dt, dh=50.0, L=2, fl=fl)
    o = orc.Oracle(p, vp, vs, rho, taus, taup, prec="f64")
    R, V, P, T, TP = (_periodic(a) for a in (rho, vs, vp, taus, taup))
    U = R * V * V
    J, I = np.meshgrid(np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    uip = 4.0 / ((1.0 / U[J, I]) + (1.0 / U[J, I + 1]) + (1.0 / U[J + 1, I]) + (1.0 / U[J + 1, I + 1]))
    tip = 0.25 * (T[J, I] + T[J, I + 1] + T[J + 1, I] + T[J + 1, I + 1])
    w = 2 * np.pi * np.float32(fl[0])
    pts = [1.0 / (2 * np.pi * np.float32(f)) for f in fl]
    S = sum((w * w * t * t) / (1.0 + w * w * t * t) for t in pts)
    dtf, L = float(np.float32(dt)), len(fl)
    fipjp = uip / (1.0 + S * tip) * dtf * (1.0 + L * tip)
    f = U[J, I] / (1.0 + S * T[J, I]) * dtf * (1.0 + L * T[J, I])
    g = R[J, I] * P[J, I] ** 2 / (1.0 + S * TP[J, I]) * dtf * (1.0 + L * TP[J, I])
    nd = o.nd
    inner = (slice(nd, nd + ny), slice(nd, nd + nx))
    for fid, want in ((orc.ORC_FIPJP, fipjp), (orc.ORC_F, f), (orc.ORC_G, g)):
        np.testing.assert_allclose(o.field(fid)[inner], want, rtol=1e-6, atol=0, err_msg=str(fid))


# --------------------------------------------------------------------------
# seismo_ssg div / curl, msource
# --------------------------------------------------------------------------
def test_div_curl_receivers_on_linear_fields():
    """SEISMO 3: div = (vxx+vyy) DT sqrt(rho vp^2), curl = (vxy-vyx) DT
    sqrt(rho vs^2) (src/seismo_ssg.c:25-65).  On linear velocity fields the
    strain rates are exact, so the samples are known."""
    nx = ny = 36
    dh, dt = 10.0, 1e-3
    p = params(nx=nx, ny=ny, nt=3, dh=dh, dt=dt, seismo=3, maxrelerror=0)
    vpv, vsv, rhov = 3000.0, 1700.0, 2000.0
    vp, vs, rho = homogeneous(nx, ny, vp=vpv, vs=vsv, rho=rhov)
    o = orc.Oracle(p, vp, vs, rho, prec="f64")
    o.set_receivers(np.array([[14, 18, 22], [15, 18, 20]], np.int32))
    a, b, c, d = 2.0e-3, -0.7e-3, 1.1e-3, 0.4e-3
    vx, vy = _linear_velocity(o, a, b, c, d, dh)
    o.set_field(orc.ORC_VX, vx)
    o.set_field(orc.ORC_VY, vy)
    o.step(1)   # zero stress: the velocity update leaves the interior unchanged
    dt32 = float(np.float32(dt))
    div = (a + b) * dt32 * math.sqrt(rhov * vpv ** 2)
    curl = (c - d) * dt32 * math.sqrt(rhov * float(np.float32(vsv)) ** 2)
    assert np.allclose(o.seismogram(3)[:, 0], div, rtol=1e-6)
    assert np.allclose(o.seismogram(4)[:, 0], curl, rtol=1e-6)


def test_moment_tensor_source_pattern():
    """QUELLTYP 5 (M12 = 1): the four shear-stress points around the source
    cell (i, j), at (i +- 1/2, j +- 1/2), each gain amp/4; normal stresses are
    untouched (src/msource.c:22-40)."""
    nx, ny, nt = 30, 30, 6
    p = params(nx=nx, ny=ny, nt=nt, dt=1e-3, seismo=0)
    vp, vs, rho = homogeneous(nx, ny)
    o = orc.Oracle(p, vp, vs, rho, prec="f64")
    sig = np.array([[0.5, 2.0, -1.0, 4.0, 8.0, 16.0]], np.float32)
    src = np.zeros((8, 1), np.float32)
    src[0, 0], src[1, 0], src[4, 0], src[7, 0] = 12, 17, 10.0, 5
    o.set_sources(src, sig)
    o.step(3)
    amp = (4.0 - 2.0) / float(np.float32(1e-3))
    nd = o.nd
    sxy = o.field(orc.ORC_SXY)
    # sxy[j][i] sits at (i+1/2, j+1/2): the cells (i-1..i, j-1..j) surround (i, j)
    expect = np.zeros_like(sxy)
    for jj in (16, 17):
        for ii in (11, 12):
            expect[jj + nd - 1, ii + nd - 1] = 0.25 * amp
    assert np.allclose(sxy, expect, rtol=1e-12, atol=0)
    assert not np.any(o.field(orc.ORC_SXX)) and not np.any(o.field(orc.ORC_SYY))
    # the pattern's centre is the source cell
    jw, iw = np.nonzero(sxy)
    assert np.mean(iw - (nd - 1) + 0.5) == 12 and np.mean(jw - (nd - 1) + 0.5) == 17
