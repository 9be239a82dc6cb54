"""Host FWI iteration functions of libdnio (include/denise_fwi.h) on the
CPU: the L-BFGS direction, the model averages, the bounded model update,
the parabola fit and the step-length search equal the numpy restatement of
the reference (oracle/fwi_oracle.py) bit for bit, and both are pinned
against what the mathematics fixes: the two-loop recursion equals the
explicit BFGS inverse-Hessian recursion over the stored pairs (in the
reference's slot order, so a wrapped cyclic buffer is covered); the
parabola fit returns the vertex of an exact parabola and its clamps; the
model update moves every free cell by exactly eps_scale * max(model) for a
unit direction and keeps cells that would leave the limits; the step
search on an exactly parabolic misfit returns its minimiser."""
import numpy as np
import pytest

import dnfwi
import fwi_oracle as fo
from cases import rel_l2

F32 = np.float32
NX, NY = 9, 7


def rnd(rng, lo=-1.0, hi=1.0, shape=(NX, NY)):
    return rng.uniform(lo, hi, shape).astype(F32)


def models(rng):
    vp = rnd(rng, 1500.0, 4000.0)
    vs = (vp / F32(np.sqrt(3.0))).astype(F32)
    rho = rnd(rng, 1800.0, 2600.0)
    return vp, vs, rho


def run_iterations(nl, iters, seed):
    """The same gradient/model sequence through the product and the oracle."""
    rng = np.random.default_rng(seed)
    lb, ob = dnfwi.LBFGS(NX, NY, nl), fo.LBFGS(NX, NY, nl)
    vp, vs, rho = models(rng)
    c = (fo.model_mean(vp), fo.model_mean(vs), fo.model_mean(rho))
    out = []
    for it in range(1, iters + 1):
        g = [rnd(rng, -1e-3, 1e-3) for _ in range(3)]
        d_prod = lb.direction(it, *c, *g, vp, vs, rho)
        d_orc = ob.direction(it, *c, *g, vp, vs, rho)
        out.append((it, g, (vp, vs, rho), d_prod, d_orc, lb.pointer(), ob.pointer))
        # next model: a small step along the direction
        vp = (vp - F32(0.5) * d_prod[0] / np.abs(d_prod[0]).max()).astype(F32)
        vs = (vs - F32(0.3) * d_prod[1] / np.abs(d_prod[1]).max()).astype(F32)
        rho = (rho - F32(0.2) * d_prod[2] / np.abs(d_prod[2]).max()).astype(F32)
    lb.close()
    return c, out


def test_lbfgs_direction_matches_oracle():
    _, out = run_iterations(nl=3, iters=8, seed=5)
    pointers = []
    for it, g, m, dp, do, pp, po in out:
        pointers.append(pp)
        assert pp == po
        for a, b in zip(dp, do):
            assert np.array_equal(a, b), it
    # FWI_PSV.c:455-464: the pointer stays at 1 for iterations 1 and 2, then cycles
    assert pointers == [1, 1, 2, 3, 1, 2, 3, 1]


def test_lbfgs_files_hold_model_gradient_and_direction(tmp_path):
    """dnfwi_lbfgs_save after each direction writes the nine files of
    extract_LBFGS_PSV.c:80-249 (model, normalised gradient C*g, direction
    per parameter), each also as the one rank's .old.0.0, in model-file
    order; their values equal the oracle's L-BFGS state ([rho | vs | vp]
    vectors) at iterations 1 (direction = C*g) and 3."""
    rng = np.random.default_rng(9)
    lb, ob = dnfwi.LBFGS(NX, NY, 2), fo.LBFGS(NX, NY, 2)
    vp, vs, rho = models(rng)
    c = (fo.model_mean(vp), fo.model_mean(vs), fo.model_mean(rho))
    jac = str(tmp_path / "jac")
    with pytest.raises(RuntimeError):       # nothing to save before the first direction
        lb.save(jac, vp, vs, rho)
    nc = NX * NY
    names = (("p_vp", "p", "c"), ("p_vs", "p_u", "c_u"), ("p_mrho", "p_rho", "c_rho"))
    for it in (1, 2, 3):
        g = [rnd(rng, -1e-3, 1e-3) for _ in range(3)]
        d = lb.direction(it, *c, *g, vp, vs, rho)
        do = ob.direction(it, *c, *g, vp, vs, rho)
        lb.save(jac, *d)
        for k, slot in enumerate((2, 1, 0)):
            want = (ob.mold[slot * nc:(slot + 1) * nc], ob.gold[slot * nc:(slot + 1) * nc], do[k].ravel())
            for n, w in zip(names[k], want):
                for sfx in (".old", ".old.0.0"):
                    got = np.fromfile(f"{jac}_{n}{sfx}", np.float32)
                    assert np.array_equal(got, w), (it, n, sfx)
        # the model and normalised gradient are the inputs, C * g in float32
        assert np.array_equal(np.fromfile(f"{jac}_p_vs.old", np.float32), vs.ravel())
        assert np.array_equal(np.fromfile(f"{jac}_p_rho.old", np.float32), (F32(c[2]) * g[2]).ravel())
        if it == 1:
            assert np.array_equal(np.fromfile(f"{jac}_c.old", np.float32), (F32(c[0]) * g[0]).ravel())
        vp = (vp - F32(0.5) * d[0] / np.abs(d[0]).max()).astype(F32)
        vs = (vs - F32(0.3) * d[1] / np.abs(d[1]).max()).astype(F32)
        rho = (rho - F32(0.2) * d[2] / np.abs(d[2]).max()).astype(F32)
    lb.close()


def test_lbfgs_equals_bfgs_matrix_recursion():
    """H_k = V_k^T H_{k-1} V_k + rho_k s_k s_k^T, V_k = I - rho_k y_k s_k^T,
    H_0 = gamma I, over the slots 1..NLBFGS in index order; the direction is
    C * H (C g).  FP64 dense matrices against the product's FP32 result."""
    nl = 3
    c, out = run_iterations(nl=nl, iters=7, seed=9)
    cv = np.array(c, np.float64)
    nc = NX * NY
    n = 3 * nc
    scale = np.concatenate([np.full(nc, cv[2]), np.full(nc, cv[1]), np.full(nc, cv[0])])   # rho | vs | vp
    slots = {k: None for k in range(1, nl + 1)}
    prev = None
    for it, g, m, dp, do, ptr, _ in out:
        gn = scale * np.concatenate([g[2].ravel(), g[1].ravel(), g[0].ravel()]).astype(np.float64)
        mv = np.concatenate([m[2].ravel(), m[1].ravel(), m[0].ravel()]).astype(np.float64)
        if it > 1:
            slots[ptr] = (mv - prev[1], gn - prev[0])            # (s, y)
            s_c, y_c = slots[ptr]
            H = np.eye(n) * (y_c @ s_c) / (y_c @ y_c)
            for k in range(1, nl + 1):
                if slots[k] is None:
                    continue
                s, y = slots[k]
                r = 1.0 / (y @ s)
                V = np.eye(n) - r * np.outer(y, s)
                H = V.T @ H @ V + r * np.outer(s, s)
            d = scale * (H @ gn)
            got = np.concatenate([dp[2].ravel(), dp[1].ravel(), dp[0].ravel()]).astype(np.float64)
            rel = np.linalg.norm(got - d) / np.linalg.norm(d)
            print(f"iteration {it}: rel-L2 vs FP64 BFGS matrix {rel:.2e}")
            assert rel < 1e-5, (it, rel)
        else:
            got = np.concatenate([dp[2].ravel(), dp[1].ravel(), dp[0].ravel()]).astype(np.float64)
            assert np.allclose(got, gn, rtol=1e-6)
        prev = (gn, mv)


def test_lbfgs_reset_and_misuse():
    lb = dnfwi.LBFGS(NX, NY, 2)
    rng = np.random.default_rng(1)
    vp, vs, rho = models(rng)
    g = [rnd(rng) for _ in range(3)]
    with pytest.raises(dnfwi.dnio.IOError_):
        lb.direction(2, 1.0, 1.0, 1.0, *g, vp, vs, rho)      # no previous iteration
    d1 = lb.direction(1, 1.0, 1.0, 1.0, *g, vp, vs, rho)
    assert np.array_equal(d1[0], g[0])                        # iter 1: the (C = 1) gradient
    lb.direction(2, 1.0, 1.0, 1.0, *g, vp, vs, rho)
    lb.direction(3, 1.0, 1.0, 1.0, *g, vp, vs, rho)
    assert lb.pointer() == 2
    lb.reset()
    assert lb.pointer() == 1
    lb.close()


def test_model_mean():
    rng = np.random.default_rng(3)
    a = rnd(rng, 1000.0, 5000.0, (37, 23))
    assert dnfwi.model_mean(a) == float(fo.model_mean(a))
    assert dnfwi.model_mean(np.full((NX, NY), 1500.0, F32)) == 1500.0
    assert abs(dnfwi.model_mean(a) / float(a.astype(np.float64).mean()) - 1.0) < 1e-5


PAR = dict(vpupperlim=5000.0, vplowerlim=1000.0, vsupperlim=3000.0, vslowerlim=500.0, rhoupperlim=3000.0,
           rholowerlim=1000.0, scalerho=0.5, inv_vp_iter=0, inv_vs_iter=0, inv_rho_iter=0, invmat1=1)


def test_model_update_matches_oracle():
    rng = np.random.default_rng(11)
    for trial in range(6):
        vp, vs, rho = models(rng)
        vs[0, :3] = 0.0                                       # fluid cells
        d = [rnd(rng) for _ in range(3)]
        par = dict(PAR, inv_vs_iter=2 if trial == 3 else 0, invmat1=3 if trial == 4 else 1)
        eps = [0.01, 0.5, 2.0, 0.05, 0.3, 0.0][trial]
        got = dnfwi.model_update(dnfwi.update_par(**par), 1, eps, vp, vs, rho, *d)
        ref = fo.mat_change_test(par, 1, eps, vp, vs, rho, *d)
        for a, b in zip(got[:3], ref[:3]):
            assert np.array_equal(a, b), trial
        assert got[3] == float(ref[3])


def test_model_update_pins():
    rng = np.random.default_rng(12)
    vp, vs, rho = models(rng)
    one = np.ones((NX, NY), F32)
    par = dnfwi.update_par(**PAR)
    # unit direction: every cell moves by eps_scale * max(model) (no limit reached)
    nv, ns, nr, eps_true = dnfwi.model_update(par, 1, 0.01, vp, vs, rho, one, one, one)
    assert np.array_equal(nv, (vp - F32(F32(0.01) * vp.max())).astype(F32))
    assert np.array_equal(ns, (vs - F32(F32(0.01) * vs.max())).astype(F32))
    # rho's step is scaled by SCALERHO
    assert np.array_equal(nr, (rho - F32(F32(F32(0.01) * rho.max()) * F32(0.5))).astype(F32))
    assert eps_true == float(F32(F32(0.01) * vs.max()))
    # limits: a step that leaves [lower, upper] keeps the old value
    nv, ns, nr, _ = dnfwi.model_update(par, 1, 0.9, vp, vs, rho, one, -one, one)
    assert np.all((nv == vp) | (nv >= 1000.0))
    assert np.array_equal(nv, np.where((vp - F32(F32(0.9) * vp.max())) < 1000.0, vp, nv))
    assert np.all(ns <= 3000.0) and np.all(nr >= 1000.0)
    # a parameter before its INV_*_ITER does not move
    par2 = dnfwi.update_par(**dict(PAR, inv_vp_iter=3))
    nv, _, _, _ = dnfwi.model_update(par2, 2, 0.01, vp, vs, rho, one, one, one)
    assert np.array_equal(nv, vp)
    # fluid cells (vs = 0) stay at 0 below VSLOWERLIM: the lower limit only applies above 1e-6
    vs0 = vs.copy()
    vs0[:, 0] = 0.0
    dz = one.copy()
    dz[:, 0] = 0.0
    _, ns, _, _ = dnfwi.model_update(par, 1, 0.01, vp, vs0, rho, one, dz, one)
    assert np.all(ns[:, 0] == 0.0)


def test_opt_step_pins_and_oracle():
    e = np.array([0.0, 0.01, 0.02], F32)

    def para(a, b, c):
        return [a + b * float(x) + c * float(x) * float(x) for x in e]

    # vertex inside the bracket
    assert abs(dnfwi.opt_step(para(1.0, -20.0, 1000.0), e) - 0.01) < 1e-6
    assert abs(dnfwi.opt_step(para(5.0, -30.0, 1000.0), e) - 0.015) < 1e-6
    # vertex beyond the last trial -> the last trial
    assert dnfwi.opt_step(para(1.0, -100.0, 1000.0), e) == float(e[2])
    # concave with L2[2] > L2[3] -> the last trial
    assert dnfwi.opt_step(para(1.0, 10.0, -1000.0), e) == float(e[2])
    # vertex at negative step -> the first trial (0)
    assert dnfwi.opt_step(para(1.0, 20.0, 1000.0), e) == 0.0
    rng = np.random.default_rng(4)
    for _ in range(50):
        ee = np.sort(rng.uniform(0.0, 0.05, 3)).astype(F32)
        ee[0] = 0.0
        L = list(rng.uniform(0.5, 2.0, 3))
        assert dnfwi.opt_step(L, ee) == float(fo.opt_step(L, ee))


def _search_both(misfit, L2t0, eps0=0.01, scalefac=2.0, stepmax=4, carry=None, seed=21):
    rng = np.random.default_rng(seed)
    vp, vs, rho = models(rng)
    d = [rnd(rng, 0.5, 1.0) for _ in range(3)]
    par = dnfwi.update_par(**dict(PAR, vplowerlim=0.0, vslowerlim=0.0, rholowerlim=0.0))
    calls_p, calls_o = [], []

    def mp(a, b, c, itest):
        v = misfit(a, b, c, vp, d[0])
        calls_p.append((itest, v))
        return v

    def mo(a, b, c, itest):
        v = misfit(a, b, c, vp, d[0])
        calls_o.append((itest, v))
        return v

    st = dnfwi.LineSearch()
    if carry is not None:
        st.epst[:] = carry[0]
        st.L2t[:] = carry[1]
    got = dnfwi.step_length(par, 1, eps0, scalefac, stepmax, L2t0, vp, vs, rho, *d, mp, state=st)
    ref = fo.step_length(dict(PAR, vplowerlim=0.0, vslowerlim=0.0, rholowerlim=0.0), 1, eps0, scalefac, stepmax,
                         L2t0, vp, vs, rho, *d, mo, carry=carry)
    assert calls_p == calls_o
    assert got.eps_scale == float(ref["eps_scale"])
    assert list(got.epst) == [float(v) for v in ref["epst"]]
    assert list(got.L2t) == ref["L2t"]
    assert (got.step1, got.step3, got.nevals) == (ref["step1"], ref["step3"], ref["nevals"])
    return got, ref, vp, d


def test_step_length_brackets_parabola_minimum():
    """Misfit exactly quadratic in the step: min at eps_scale = 0.03."""
    def misfit(a, b, c, vp, dvp):
        eps = float(vp.max()) / float(np.abs(dvp).max())
        t = float(np.mean((vp.astype(np.float64) - a) / dvp)) / eps      # the trial's eps_scale
        return 1.0 + 1000.0 * (t - 0.03) ** 2

    got, ref, _, _ = _search_both(misfit, L2t0=1.0 + 1000.0 * 0.03 ** 2)
    assert got.step1 == 1 and got.step3 == 0
    assert abs(got.eps_scale - 0.03) < 1e-4
    assert got.epst[0] == 0.0 and got.epst[1] < 0.03 < got.epst[2]


def test_step_length_failure_keeps_old_third_slot():
    """Every trial raises the misfit: the step is halved STEPMAX+1 times, the
    search fails (step3 = 1) and slot 3 keeps the previous iteration's values."""
    def misfit(a, b, c, vp, dvp):
        return 2.0 + float(np.abs(vp.astype(np.float64) - a).sum())

    got, ref, _, _ = _search_both(misfit, L2t0=2.0, stepmax=3, carry=([0.0, 0.02, 0.04], [1.5, 1.2, 1.3]))
    assert got.step1 == 0 and got.step3 == 1 and got.nevals == 4
    assert abs(got.eps_scale - 0.01 / 2 ** 4) < 1e-9
    assert got.epst[2] == float(F32(0.04)) and got.L2t[2] == 1.3


def test_step_length_no_upper_bracket():
    """The misfit keeps falling: after STEPMAX growths the search stops with
    the last trial in slot 3 and fits the parabola anyway."""
    def misfit(a, b, c, vp, dvp):
        return 1.0 - float(np.abs(vp.astype(np.float64) - a).sum()) * 1e-6

    got, ref, _, _ = _search_both(misfit, L2t0=1.0, stepmax=2)
    assert got.step1 == 1 and got.step3 == 0
    assert got.eps_scale == float(got.epst[2])


def test_rtm_adjoint_source():
    """MODE 2 adjoint source: equal to the restatement bit for bit; pinned by
    the integral of a constant trace (j*DT*c, time-reversed), the sample-1
    rule (the synthetic's first sample replaces the data's) and GRAD_FORM 2
    (the time-reversed data)."""
    rng = np.random.default_rng(31)
    data = rng.standard_normal((5, 300)).astype(F32)
    syn = rng.standard_normal((5, 300)).astype(F32)
    for gf in (1, 2):
        for test in (0, 1):
            got, l2 = dnfwi.rtm_adjoint_source(data, syn, 1e-3, gf, test)
            ref, l2r = fo.rtm_adjoint_source(data, syn, 1e-3, gf, test)
            assert np.array_equal(got, ref) and l2 == l2r
            assert (l2 > 0) == bool(test)
    c = np.full((2, 200), 3.0, F32)
    got, _ = dnfwi.rtm_adjoint_source(c, c, 1e-3, 1, 1)
    ramp = 3.0 * 1e-3 * np.arange(1, 201)
    assert np.allclose(got[:, ::-1], ramp, rtol=2e-5)
    syn1 = c.copy()
    syn1[:, 0] = 7.0
    got1, _ = dnfwi.rtm_adjoint_source(c, syn1, 1e-3, 1, 1)
    assert np.allclose(got1[:, ::-1] - got[:, ::-1], 4.0 * 1e-3, rtol=1e-3)
    got2, _ = dnfwi.rtm_adjoint_source(data, syn, 1e-3, 2, 1)
    assert np.array_equal(got2[:, ::-1][:, 1:], data[:, 1:]) and np.array_equal(got2[:, -1], syn[:, 0])


@pytest.mark.parametrize("eprecond", [1, 3])
def test_energy_precond(eprecond):
    """dnfwi_energy_precond equals the restatement bit for bit; pins: for
    EPRECOND 3 the asinh factor is the integral over the receiver line
    x' in [DH, NX*DH] of 1/sqrt((x'-x)^2 + y^2) (quadrature); for EPRECOND 1
    on constant energies We = sqrt(Ws*Wr) * 1.005; the gradients are divided
    by We*C^2."""
    from scipy.integrate import quad
    rng = np.random.default_rng(eprecond)
    nx, ny, dh = 13, 9, 10.0
    Ws = rng.uniform(0.5, 2.0, (nx, ny)).astype(F32)
    Wr = rng.uniform(0.5, 2.0, (nx, ny)).astype(F32)
    g = [rnd(rng, shape=(nx, ny)) for _ in range(3)]
    got = dnfwi.energy_precond(eprecond, dh, Ws, Wr, 2500.0, 1400.0, 2100.0, *g)
    ref = fo.energy_precond(eprecond, dh, Ws, Wr if eprecond == 1 else None, 2500.0, 1400.0, 2100.0, *g)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    We = got[0]
    if eprecond == 3:
        raw = (We - F32(0.005) * (We.max() / 1.005)) / np.sqrt(Ws.astype(np.float64))
        for i, j in ((1, 1), (6, 4), (13, 9)):
            x, y = i * dh, j * dh
            val, _ = quad(lambda xp: 1.0 / np.sqrt((xp - x) ** 2 + y ** 2), dh, nx * dh)
            assert abs(raw[i - 1, j - 1] - val) < 1e-4 * val
    else:
        c = dnfwi.energy_precond(1, dh, np.full((nx, ny), 4.0, F32), np.full((nx, ny), 9.0, F32), 1.0, 1.0, 1.0,
                                 *g)[0]
        assert np.allclose(c, 6.0 * 1.005, rtol=1e-6)
    assert np.allclose(got[1], g[0] / (We.astype(np.float64) * 2500.0 ** 2), rtol=1e-5)
    _, _, gs0, _ = dnfwi.energy_precond(eprecond, dh, Ws, Wr, 2500.0, 0.0, 2100.0, *g)
    assert np.all(gs0 == 0.0)


def test_taper_hor():
    """dnfwi_taper_hor equals the restatement bit for bit; pins: 0 above
    GRADT1, 1 at GRADT2, the Gaussian ramp exp(-(3(k-L)/(L/2))^2/2) in between,
    (j*DH)^EXP below."""
    nx, ny, dh, g1, g2, ex = 7, 40, 10.0, 10, 20, 0.5
    got = dnfwi.taper_hor(nx, ny, dh, g1, g2, ex)
    assert np.array_equal(got, fo.taper_hor(nx, ny, dh, g1, g2, ex))
    L = g2 - g1
    assert np.all(got[:, :g1 - 1] == 0.0)
    assert np.all(got[:, g2 - 1] == 1.0)
    for j in range(g1, g2 + 1):
        k = j - g1 + 1
        want = np.exp(-0.5 * (3 * (k - L) / (L / 2)) ** 2) if k <= L else 1.0
        assert np.allclose(got[:, j - 1], want, rtol=1e-6, atol=1e-12)
    for j in range(g2 + 1, ny + 1):
        assert np.allclose(got[:, j - 1], (j * dh) ** ex, rtol=1e-6)


@pytest.mark.parametrize("pcg_beta", [1, 2, 3, 4])
def test_pcg_direction_and_conjugacy(pcg_beta):
    """dnfwi_pcg_direction equals the restatement bit for bit over several
    iterations (three classes, random gradients); pin: on a quadratic
    misfit of the vp class, 1/2 m^T A m - b^T m with A symmetric positive
    definite (n = 6 cells), exact line searches along the directions reach
    the minimum in n iterations for every beta formula (linear conjugate
    gradients), and iteration 1 is steepest descent."""
    rng = np.random.default_rng(pcg_beta)
    pc, po = dnfwi.PCG(NX, NY), fo.PCG(NX, NY)
    for it in range(5):
        g = [rnd(rng) for _ in range(3)]
        a = pc.direction(pcg_beta, *g)
        b = po.direction(pcg_beta, *g)
        for x, y in zip(a, b):
            assert np.array_equal(x, y), it
        if it == 0:
            assert np.array_equal(a[0], g[0]) and np.all(a[3] == 0)
    pc.close()
    # the pin
    nx, ny = 3, 2
    n = nx * ny
    M = rng.standard_normal((n, n))
    A = M @ M.T + n * np.eye(n)
    bvec = rng.standard_normal(n)
    m = np.zeros(n)
    cg = dnfwi.PCG(nx, ny)
    zero = np.zeros((nx, ny), F32)
    g0 = np.linalg.norm(A @ m - bvec)
    for it in range(n):
        g = (A @ m - bvec).astype(F32).reshape(nx, ny)
        d = cg.direction(pcg_beta, g, zero, zero)[0].astype(np.float64).ravel()
        p = -d                                   # the model moves along -direction
        eps = -float((A @ m - bvec) @ p) / float(p @ A @ p)
        m = m + eps * p
    cg.close()
    assert np.linalg.norm(A @ m - bvec) < 1e-4 * g0


def test_taper_vert():
    """dnfwi_taper_vert equals the restatement bit for bit; pins: 0 up to
    GRADT1 and from GRADT4, 1 between the ramps, both Gaussian ramps in
    closed form, and the gradient is 0 at every source and receiver cell."""
    nx, ny, dh, gt = 60, 9, 10.0, (5, 15, 40, 50)
    rng = np.random.default_rng(8)
    g = rnd(rng, shape=(nx, ny))
    src = np.zeros((8, 3), F32)
    src[0], src[1] = [100.0, 204.0, 305.0], [20.0, 30.0, 41.0]
    rec = np.array([[12, 20, 33], [2, 2, 2]], np.int32)
    coeff, got = dnfwi.taper_vert(nx, ny, dh, gt, src, rec, g)
    rc, rg = fo.taper_vert(nx, ny, dh, gt, src, rec, g)
    assert np.array_equal(coeff, rc) and np.array_equal(got, rg)
    assert np.all(coeff[:5] == 0.0) and np.all(coeff[49:] == 0.0)
    assert np.all(coeff[15:39] == 1.0)
    for i in range(6, 50):
        k = i - 5
        want = 1.0
        if k <= 10:
            want = np.exp(-0.5 * (3 * (k - 10) / 5.0) ** 2)
        if k >= 45 - 10:
            want = np.exp(-0.5 * (3 * (k - 35) / 5.0) ** 2)
        assert abs(coeff[i - 1, 0] - want) < 1e-6 * max(want, 1e-30) + 1e-30, i
    for (x, y) in ((10, 2), (20, 3), (31, 4), (12, 2), (20, 2), (33, 2)):
        assert got[x - 1, y - 1] == 0.0
    mask = np.ones_like(g, bool)
    for (x, y) in ((10, 2), (20, 3), (31, 4), (12, 2), (20, 2), (33, 2)):
        mask[x - 1, y - 1] = False
    assert np.array_equal(got[mask], (g * coeff)[mask])


@pytest.mark.parametrize("srtshape", [1, 2])
def test_taper_sources(srtshape):
    """dnfwi_taper_sources equals the restatement bit for bit; pins: the
    taper lies in [0, 1], is 1 away from the sources and 0 in the FILTSIZE
    box around each; for SRTSHAPE 1 it rises with the distance from a source
    like erf(r/rmax) (monotone along a row through the source)."""
    nx, ny, dh = 50, 30, 10.0
    rng = np.random.default_rng(srtshape)
    g = rnd(rng, shape=(nx, ny))
    src = np.zeros((8, 2), F32)
    src[0], src[1] = [120.0, 352.0], [100.0, 150.0]
    coeff, got = dnfwi.taper_sources(nx, ny, dh, 40.0, srtshape, 1, src, g)
    rc, rg = fo.taper_sources(nx, ny, dh, 40.0, srtshape, 1, src, g)
    assert np.array_equal(coeff, rc) and np.array_equal(got, rg)
    assert coeff.min() >= 0.0 and coeff.max() <= 1.0
    assert np.all(coeff[:, 25:] == 1.0) and np.all(coeff[45:, :] == 1.0)
    for (i, j) in ((12, 10), (35, 15)):
        assert np.all(coeff[i - 2:i + 1, j - 2:j + 1] == 0.0)
        assert np.all(got[i - 2:i + 1, j - 2:j + 1] == 0.0)
    if srtshape == 1:
        row = coeff[12:20, 9]          # x from the box edge outwards at the source depth
        assert np.all(np.diff(row) >= 0.0)


def _median_brute(a, fs):
    """The median filter by its definition: clamp every window index into
    the array, sort the fs*fs values, take the middle one."""
    nx, ny = a.shape
    h = fs // 2
    out = np.empty_like(a)
    for i in range(nx):
        for j in range(ny):
            v = sorted(float(a[min(max(i + di, 0), nx - 1), min(max(j + dj, 0), ny - 1)])
                       for di in range(-h, h + 1) for dj in range(-h, h + 1))
            out[i, j] = v[len(v) // 2]
    return out


@pytest.mark.parametrize("filt_size", [0, 1, 2, 3, -4, 5, 7])
def test_median_filter(filt_size):
    """dnfwi_median_filter equals the restatement bit for bit; pins: the
    definition (clamped window, sorted, middle value) on a small array,
    an even size acting as the next odd one, a negative size as its
    absolute value, 0 and 1 leaving the array unchanged."""
    rng = np.random.default_rng(11 + abs(filt_size))
    a = rng.uniform(1500.0, 4500.0, (37, 23)).astype(np.float32)
    a[5:9, 3:6] = 2000.0          # ties
    got = dnfwi.median_filter(a, filt_size)
    assert np.array_equal(got, fo.median_filter(a, filt_size))
    fs = abs(filt_size) + (1 - abs(filt_size) % 2) if filt_size else 1
    small = a[:9, :7]
    assert np.array_equal(fo.median_filter(small, filt_size), _median_brute(small, fs))
    if filt_size in (0, 1):
        assert np.array_equal(got, a)


def test_median_filter_properties():
    """Constant fields and straight step edges pass unchanged, an isolated
    spike is removed, and the input array of the binding is not modified."""
    a = np.full((20, 15), 3000.0, np.float32)
    assert np.array_equal(dnfwi.median_filter(a, 5), a)
    step = np.where(np.arange(20)[:, None] < 8, 2000.0, 3500.0).astype(np.float32) * np.ones((1, 15), np.float32)
    assert np.array_equal(dnfwi.median_filter(step, 5), step)
    spike = a.copy()
    spike[10, 7] = 9000.0
    assert np.array_equal(dnfwi.median_filter(spike, 3), a)
    assert spike[10, 7] == 9000.0


@pytest.mark.parametrize("wd_damp,c_vs,fc_end,dh", [(0.5, 2000.0, 10.0, 10.0), (0.3, 1800.0, 7.0, 5.0),
                                                    (0.25, 2400.0, 12.0, 10.0), (0.01, 2000.0, 10.0, 10.0),
                                                    (-0.45, 2000.0, 10.0, 10.0)])
def test_gauss_filter(wd_damp, c_vs, fc_end, dh):
    """dnfwi_gauss_filter equals the restatement bit for bit (widths 11, 15,
    5 (from 4), 0 and -9 (no normalisation, as in the reference))."""
    rng = np.random.default_rng(3)
    g = rng.standard_normal((41, 29)).astype(np.float32) * np.float32(1e-9)
    got, w = dnfwi.gauss_filter(g, dh, wd_damp, c_vs, fc_end)
    want, w2 = fo.gauss_filter(g, dh, wd_damp, c_vs, fc_end)
    assert w == w2
    assert np.array_equal(got, want)
    if w == 0:
        assert np.array_equal(got, g)


@pytest.mark.parametrize("wd_damp", [0.05, 0.15, -0.05, -0.15])
def test_gauss_filter_refuses_zero_sigma(wd_damp):
    """Widths 1 and 3 (here round(WD_DAMP 2000 / 10 / 10) = 1, 3 and their
    negatives) give hfs <= 1, sigma = hfs / 2 = 0 and a 0/0 kernel, NaN in
    every gradient cell in the reference: the filter refuses them."""
    g = np.ones((20, 15), np.float32)
    with pytest.raises(Exception, match="zero Gaussian"):
        dnfwi.gauss_filter(g, 10.0, wd_damp, 2000.0, 10.0)


def test_gauss_filter_pins():
    """Width round(WD_DAMP C_vs / FC_END / DH) = 10 -> 11, h = 5, sigma = 2
    (integer h/2), s = 8: a unit spike away from the edges spreads into the
    normalised product of the 1-D Gaussians exp(-k^2/8), k = -5..5; a
    constant field stays constant; the filter is linear; width 4 -> 5 has
    sigma 1 (s = 2).  A near-half width rounds away from zero (2.5 -> 3)."""
    g = np.zeros((31, 27), np.float32)
    g[15, 13] = 1.0
    out, w = fo.gauss_filter(g, 10.0, 0.5, 2000.0, 10.0)
    assert w == 11
    k1 = np.exp(-np.arange(-5, 6) ** 2 / 8.0)
    want = np.outer(k1, k1) / k1.sum() ** 2
    np.testing.assert_allclose(out[10:21, 8:19], want, rtol=2e-6, atol=1e-9)
    assert np.count_nonzero(out) == 121
    out5, w5 = fo.gauss_filter(g, 10.0, 0.2, 2000.0, 10.0)
    k1 = np.exp(-np.arange(-2, 3) ** 2 / 2.0)
    assert w5 == 5
    np.testing.assert_allclose(out5[13:18, 11:16], np.outer(k1, k1) / k1.sum() ** 2, rtol=2e-6)
    c = np.full((20, 17), 3.5, np.float32)
    np.testing.assert_allclose(fo.gauss_filter(c, 10.0, 0.5, 2000.0, 10.0)[0], c, rtol=1e-5)
    rng = np.random.default_rng(9)
    a, b = (rng.standard_normal((25, 19)).astype(np.float32) for _ in range(2))
    fa, fb, fab = (fo.gauss_filter(x, 10.0, 0.5, 2000.0, 10.0)[0] for x in (a, b, a + b))
    np.testing.assert_allclose(fab, fa + fb, atol=1e-5)
    assert fo.gauss_filter(a, 10.0, 0.25, 1000.0, 10.0)[1] == 3      # 2.5 -> 3 (odd already)
    with pytest.raises(Exception, match="filter width 3"):          # sigma 0: refused
        dnfwi.gauss_filter(a, 10.0, 0.25, 1000.0, 10.0)


@pytest.mark.parametrize("timewin", [0, 1, 2, 4, 5])
def test_time_window(timewin):
    """dnfwi_time_window equals the restatement bit for bit; pins: samples
    before the window start are untouched, the factor is 1 at the start
    and follows exp(-GAMMA (t - start)^2) after it (TIMEWIN 1, 4: start =
    pick + TWLENGTH_PLUS; 5: start = |offset / TWLENGTH_MINUS|), TIMEWIN 2
    multiplies from TWLENGTH_PLUS on by exp(-GAMMA (t + DT)^2), and column
    k of the time-reversed traces is time (ns - k) DT."""
    rng = np.random.default_rng(20 + timewin)
    ntr, ns, dt, gamma = 7, 300, 1e-3, 200.0
    d = (rng.standard_normal((ntr, ns)) + 2.0).astype(np.float32)
    picks = rng.uniform(0.02, 0.15, ntr).astype(np.float32)
    offset = rng.uniform(50.0, 900.0, ntr).astype(np.float32)
    twm, twp = 3000.0, 0.04
    got = dnfwi.time_window(timewin, d, dt, gamma, twm, twp, picks, offset)
    want = fo.time_window(timewin, d, dt, gamma, twm, twp, picks, offset)
    assert np.array_equal(got, want)
    t = np.arange(ns, 0, -1) * dt
    fac = got.astype(np.float64) / d
    for i in range(ntr):
        if timewin in (1, 4, 5):
            start = picks[i] + twp if timewin != 5 else offset[i] / twm
            ref = np.where(t >= start, np.exp(-gamma * (t - start) ** 2), 1.0)
        elif timewin == 2:
            ref = np.where(t >= twp, np.exp(-gamma * (t + dt) ** 2), 1.0)
        else:
            ref = np.ones(ns)
        np.testing.assert_allclose(fac[i], ref, rtol=3e-5, atol=1e-6)
        assert np.array_equal(got[i, t < start - 1e-5], d[i, t < start - 1e-5]) if timewin in (1, 4, 5) else True


@pytest.mark.parametrize("grad_form", [1, 2])
def test_residual_gc(grad_form):
    """dnfwi_residual_gc (LNORM 5) equals the restatement bit for bit, with
    skipped traces zero and the misfit running from a given start."""
    rng = np.random.default_rng(30 + grad_form)
    ntr, ns, dt = 6, 250, 1e-3
    syn = rng.standard_normal((ntr, ns)).astype(np.float32)
    data = (syn * np.float32(0.7) + rng.standard_normal((ntr, ns)).astype(np.float32) * np.float32(0.3)).astype(np.float32)
    skip = np.array([0, 1, 0, 0, 1, 0], np.int32)
    got, l2 = dnfwi.residual_gc(data, syn, dt, grad_form, 1, skip, l2=0.25)
    want, l2w = fo.residual_gc(data, syn, dt, grad_form, 1, skip, l2=0.25)
    assert np.array_equal(got, want)
    assert l2 == l2w
    assert not got[1].any() and not got[4].any()
    g0, l0 = dnfwi.residual_gc(data, syn, dt, grad_form, 0)
    assert l0 == 0.0 and np.array_equal(g0[0], got[0])


def test_residual_gc_pins():
    """What the mathematics fixes: the misfit -sum d.s/(|d||s|) is -1 per
    trace for data proportional to the synthetics (scale-free; traces that
    start at 0, as seismograms do), where the
    residual vanishes; the residual is the derivative of the misfit with
    respect to the synthetics (GRAD_FORM 2, central differences in float64,
    first samples equal so that the sample-1 replacement does not enter),
    time-reversed; GRAD_FORM 1 is GRAD_FORM 2 on the running integrals."""
    rng = np.random.default_rng(41)
    ntr, ns, dt = 3, 64, 1e-3
    syn = rng.standard_normal((ntr, ns)).astype(np.float32)
    syn[:, 0] = 0.0           # traces start at rest, so the sample-1 replacement changes nothing
    r, l2 = fo.residual_gc(syn * np.float32(3.0), syn, dt, 2, 1)
    assert abs(l2 + ntr) < 1e-5
    assert np.abs(r).max() < 1e-5 * np.abs(syn).max()
    data = rng.standard_normal((ntr, ns)).astype(np.float32)
    data[:, 0] = syn[:, 0]
    r, _ = fo.residual_gc(data, syn, dt, 2, 1)

    def misfit(s):
        d = data.astype(np.float64)
        return -sum(float(d[i] @ s[i]) / (np.linalg.norm(d[i]) * np.linalg.norm(s[i])) for i in range(ntr))

    s64 = syn.astype(np.float64)
    for i, j in [(0, 5), (1, 30), (2, 63), (0, 0)]:
        h = 1e-4
        sp, sm = s64.copy(), s64.copy()
        sp[i, j] += h
        sm[i, j] -= h
        fd = (misfit(sp) - misfit(sm)) / (2 * h)
        assert abs(r[i, ns - 1 - j] - fd) < 2e-4 * max(1.0, abs(fd)), (i, j, r[i, ns - 1 - j], fd)
    i1 = (np.cumsum(data, axis=1, dtype=np.float32) * np.float32(dt)).astype(np.float32)
    s1 = (np.cumsum(syn, axis=1, dtype=np.float32) * np.float32(dt)).astype(np.float32)
    r1, l1 = fo.residual_gc(data, syn, dt, 1, 1)
    r2, l2b = fo.residual_gc(i1, s1, dt, 2, 1)
    assert np.array_equal(r1[:, :-1], r2[:, :-1]) and abs(l1 - l2b) < 1e-4


def test_gauss_filter_var():
    """dnfwi_gauss_filter_var (SPATFILTER 4) equals the restatement bit for
    bit on a model with varying velocity and a WD_DAMP1 window wider than
    the WD_DAMP one."""
    rng = np.random.default_rng(8)
    nx, ny = 16, 12
    g = rng.standard_normal((nx, ny)).astype(np.float32)
    vel = np.linspace(1200.0, 2600.0, nx * ny, dtype=np.float32).reshape(nx, ny)
    got = dnfwi.gauss_filter_var(g, vel, 10.0, 0.4, 0.7, 12.0)
    want = fo.gauss_filter_var(g, vel, 10.0, 0.4, 0.7, 12.0)
    assert np.array_equal(got, want)


def test_gauss_filter_var_pins():
    """Uniform velocity 2000 m/s, FC_END 10 Hz, DH 10 m: WD_DAMP 0.5 gives
    x width 10 -> 11 (h 5, sigma 2), WD_DAMP1 0.2 gives y width 4 -> 5,
    raised to 6 (h 3, sigma 1); a unit spike spreads into the normalised
    product exp(-k^2/8) x exp(-l^2/2); constant fields stay constant;
    WD_DAMP1 0 (y width 0) leaves every cell unchanged."""
    nx, ny = 25, 17
    vel = np.full((nx, ny), 2000.0, np.float32)
    g = np.zeros((nx, ny), np.float32)
    g[12, 8] = 1.0
    out = fo.gauss_filter_var(g, vel, 10.0, 0.5, 0.2, 10.0)
    kx, kz = np.exp(-np.arange(-5, 6) ** 2 / 8.0), np.exp(-np.arange(-3, 4) ** 2 / 2.0)
    want = np.outer(kx, kz) / (kx.sum() * kz.sum())
    np.testing.assert_allclose(out[7:18, 5:12], want, rtol=3e-6, atol=1e-9)
    assert np.count_nonzero(out) == 11 * 7
    c = np.full((nx, ny), -2.5, np.float32)
    np.testing.assert_allclose(dnfwi.gauss_filter_var(c, vel, 10.0, 0.5, 0.2, 10.0), c, rtol=2e-6)
    r = np.random.default_rng(2).standard_normal((nx, ny)).astype(np.float32)
    assert np.array_equal(dnfwi.gauss_filter_var(r, vel, 10.0, 0.5, 0.0, 10.0), r)


# ---- the acoustic problem: vp and rho only (src/AC/FWI_AC.c) ----

def run_iterations_ac(nl, iters, seed):
    """One gradient/model sequence through the two-class product memory, the
    two-class oracle, and the three-class oracle with vs = 0 everywhere."""
    rng = np.random.default_rng(seed)
    lb, ob, o3 = dnfwi.LBFGS(NX, NY, nl, ac=True), fo.LBFGS(NX, NY, nl, nclass=2), fo.LBFGS(NX, NY, nl)
    vp, _, rho = models(rng)
    zero = np.zeros((NX, NY), F32)
    c = (fo.model_mean(vp), F32(0.0), fo.model_mean(rho))
    out = []
    for it in range(1, iters + 1):
        g = [rnd(rng, -1e-3, 1e-3), None, rnd(rng, -1e-3, 1e-3)]
        d_prod = lb.direction(it, *c, *g, vp, None, rho)
        d_orc = ob.direction(it, *c, *g, vp, None, rho)
        d_3 = o3.direction(it, *c, g[0], zero, g[2], vp, zero, rho)
        out.append((it, g, d_prod, d_orc, d_3, lb.pointer(), ob.pointer))
        vp = (vp - F32(0.5) * d_prod[0] / np.abs(d_prod[0]).max()).astype(F32)
        rho = (rho - F32(0.2) * d_prod[2] / np.abs(d_prod[2]).max()).astype(F32)
    lb.close()
    return out


def test_lbfgs_ac_matches_oracle_and_three_class_run():
    """store_LBFGS_AC / extract_LBFGS_AC ([rho | vp] vectors): the product
    equals the two-class oracle bit for bit, and the two-class oracle equals
    the three-class one (pinned against the BFGS matrix recursion above)
    with a zero vs class, whose products add nothing to the dot products."""
    for it, g, dp, do, d3, pp, po in run_iterations_ac(nl=3, iters=7, seed=21):
        assert pp == po
        assert dp[1] is None and do[1] is None
        for k in (0, 2):
            assert np.array_equal(dp[k], do[k]), (it, k)
            assert np.array_equal(do[k], d3[k]), (it, k)
        assert not d3[1].any()


def test_lbfgs_ac_files(tmp_path):
    """extract_LBFGS_AC.c:60-178: _p_vp, _p, _c, _p_mrho, _p_rho, _c_rho
    (model, normalised gradient, direction of vp and rho); no vs files."""
    rng = np.random.default_rng(23)
    lb, ob = dnfwi.LBFGS(NX, NY, 2, ac=True), fo.LBFGS(NX, NY, 2, nclass=2)
    vp, _, rho = models(rng)
    c = (fo.model_mean(vp), F32(0.0), fo.model_mean(rho))
    jac = str(tmp_path / "jac")
    nc = NX * NY
    for it in (1, 2, 3):
        g = [rnd(rng, -1e-3, 1e-3), None, rnd(rng, -1e-3, 1e-3)]
        d = lb.direction(it, *c, *g, vp, None, rho)
        do = ob.direction(it, *c, *g, vp, None, rho)
        lb.save(jac, d[0], None, d[2])
        for names, slot, dd in ((("p_vp", "p", "c"), 1, do[0]), (("p_mrho", "p_rho", "c_rho"), 0, do[2])):
            want = (ob.mold[slot * nc:(slot + 1) * nc], ob.gold[slot * nc:(slot + 1) * nc], dd.ravel())
            for n, w in zip(names, want):
                for sfx in (".old", ".old.0.0"):
                    assert np.array_equal(np.fromfile(f"{jac}_{n}{sfx}", np.float32), w), (it, n, sfx)
        assert np.array_equal(np.fromfile(f"{jac}_p_vp.old", np.float32), vp.ravel())
        assert np.array_equal(np.fromfile(f"{jac}_p.old", np.float32), (F32(c[0]) * g[0]).ravel())
        for n in ("p_vs", "p_u", "c_u"):
            assert not (tmp_path / f"jac_{n}.old").exists()
        vp = (vp - F32(0.5) * d[0] / np.abs(d[0]).max()).astype(F32)
        rho = (rho - F32(0.2) * d[2] / np.abs(d[2]).max()).astype(F32)
    lb.close()


def test_model_update_ac():
    """calc_mat_change_test_AC: the product equals the oracle, and the vp and
    rho of the three-class update (whose vs part is independent of them);
    eps_true is the vp step EPSILON (calc_mat_change_test_AC.c:103)."""
    rng = np.random.default_rng(25)
    for trial in range(5):
        vp, vs, rho = models(rng)
        d = [rnd(rng) for _ in range(3)]
        par = dict(PAR, inv_rho_iter=2 if trial == 2 else 0, invmat1=3 if trial == 3 else 1)
        eps = [0.01, 0.5, 0.05, 0.3, 0.0][trial]
        got = dnfwi.model_update(dnfwi.update_par(**par), 1, eps, vp, None, rho, d[0], None, d[2])
        ref = fo.mat_change_test(par, 1, eps, vp, None, rho, d[0], None, d[2])
        full = fo.mat_change_test(par, 1, eps, vp, vs, rho, *d)
        assert got[1] is None and ref[1] is None
        for k in (0, 2):
            assert np.array_equal(got[k], ref[k]), trial
            assert np.array_equal(ref[k], full[k]), trial
        assert got[3] == float(ref[3])
    one = np.ones((NX, NY), F32)
    nv, _, nr, eps_true = dnfwi.model_update(dnfwi.update_par(**PAR), 1, 0.01, vp, None, rho, one, None, one)
    assert np.array_equal(nv, (vp - F32(F32(0.01) * vp.max())).astype(F32))
    assert eps_true == float(F32(F32(0.01) * vp.max()))


def test_step_length_ac_on_parabolic_misfit():
    """step_length_est_ac: the misfit gets vs None; on a misfit parabolic in
    the vp step the search returns the parabola's minimiser, and the product
    equals the oracle's search."""
    rng = np.random.default_rng(27)
    vp, _, rho = models(rng)
    d_vp, d_rho = np.ones((NX, NY), F32), np.ones((NX, NY), F32)
    vmax = float(vp.max())
    best = 0.013

    def misfit(tv, ts, tr, itest):
        assert ts is None
        e = (float(vp[0, 0]) - float(tv[0, 0])) / vmax
        return 1.0 + (e - best) ** 2

    par = dnfwi.update_par(**dict(PAR, vplowerlim=0.0))
    st = dnfwi.step_length(par, 1, 0.01, 4.0, 6, 1.0 + best ** 2, vp, None, rho, d_vp, None, d_rho, misfit)
    ref = fo.step_length(dict(PAR, vplowerlim=0.0), 1, 0.01, 4.0, 6, 1.0 + best ** 2, vp, None, rho, d_vp, None,
                         d_rho, misfit)
    assert st.eps_scale == float(ref["eps_scale"]) and st.nevals == ref["nevals"]
    assert abs(st.eps_scale - best) < 1e-3 * best


def test_energy_precond_and_pcg_ac():
    """grad_obj_ac.c:225-233 and PCG with PCG_class 2: without a vs gradient
    the vp and rho results equal the three-class ones."""
    rng = np.random.default_rng(29)
    Ws = rnd(rng, 0.1, 2.0)
    g = [rnd(rng) for _ in range(3)]
    c = (F32(2500.0), F32(1400.0), F32(2100.0))
    We, gv, gs, gr = dnfwi.energy_precond(3, 10.0, Ws, None, *c, g[0], None, g[2])
    We3, gv3, _, gr3 = dnfwi.energy_precond(3, 10.0, Ws, None, *c, *g)
    oW, ov, os_, orr = fo.energy_precond(3, 10.0, Ws, None, *c, g[0], None, g[2])
    assert gs is None and os_ is None
    assert np.array_equal(gv, gv3) and np.array_equal(gr, gr3) and np.array_equal(We, We3)
    assert np.array_equal(gv, ov) and np.array_equal(gr, orr)
    for pcg_beta in (1, 2, 3, 4):
        p2, p3, po = dnfwi.PCG(NX, NY), dnfwi.PCG(NX, NY), fo.PCG(NX, NY)
        for it in range(4):
            gg = [rnd(rng) for _ in range(3)]
            a = p2.direction(pcg_beta, gg[0], None, gg[2])
            b = p3.direction(pcg_beta, gg[0], np.zeros((NX, NY), F32), gg[2])
            o = po.direction(pcg_beta, gg[0], None, gg[2])
            assert a[1] is None and o[1] is None
            for k in (0, 2):
                assert np.array_equal(a[k], b[k]) and np.array_equal(a[k], o[k]), (pcg_beta, it, k)
            assert a[3][0] == b[3][0] == o[3][0] and a[3][2] == b[3][2] == o[3][2] and a[3][1] == 0.0
        p2.close()
        p3.close()


# ---- the SH problem: vs and rho (src/SH/FWI_SH.c) ----

def test_lbfgs_sh_matches_oracle_and_normalised_three_class_run():
    """store_LBFGS_SH / extract_LBFGS_SH ([vs | rho] vectors of normalised
    models vs / C_vs, rho / C_rho and gradients C g; the direction left
    normalised): the product equals the SH oracle bit for bit; and the
    oracle's direction equals, up to float rounding of the normalisation,
    the three-class recursion run on the normalised models with C = 1 and
    a zero vp class (the same two-loop recursion, pinned against the BFGS
    matrix recursion above)."""
    rng = np.random.default_rng(31)
    lb, ob, o3 = dnfwi.LBFGS(NX, NY, 3, sh=True), fo.LBFGS(NX, NY, 3, sh=True), fo.LBFGS(NX, NY, 3)
    _, vs, rho = models(rng)
    zero = np.zeros((NX, NY), F32)
    c = (F32(0.0), fo.model_mean(vs), fo.model_mean(rho))
    for it in range(1, 7):
        g = [None, rnd(rng, -1e-3, 1e-3), rnd(rng, -1e-3, 1e-3)]
        dp = lb.direction(it, *c, *g, None, vs, rho)
        do = ob.direction(it, *c, *g, None, vs, rho)
        # three-class run on normalised quantities (C = 1): vs class <- vs / C_vs, gradient C_vs g
        gn = [zero, (c[1] * g[1]).astype(F32), (c[2] * g[2]).astype(F32)]
        mn = [zero, (vs / c[1]).astype(F32), (rho / c[2]).astype(F32)]
        d3 = o3.direction(it, 1.0, 1.0, 1.0, *gn, *mn)
        assert dp[0] is None and do[0] is None and lb.pointer() == ob.pointer
        for k in (1, 2):
            assert np.array_equal(dp[k], do[k]), (it, k)
            assert rel_l2(do[k], d3[k]) < 1e-5, (it, k)
        vs = (vs - F32(0.3) * dp[1] / np.abs(dp[1]).max()).astype(F32)
        rho = (rho - F32(0.2) * dp[2] / np.abs(dp[2]).max()).astype(F32)
    lb.close()


def test_lbfgs_sh_files(tmp_path):
    """extract_LBFGS_SH.c:50-178: _p_vs.old holds the normalised model
    vs / C_vs, _p_u.old the normalised gradient, _c_u.old the direction;
    likewise _p_mrho, _p_rho, _c_rho; no vp files."""
    rng = np.random.default_rng(33)
    lb = dnfwi.LBFGS(NX, NY, 2, sh=True)
    _, vs, rho = models(rng)
    c = (F32(0.0), fo.model_mean(vs), fo.model_mean(rho))
    jac = str(tmp_path / "jac")
    g = [None, rnd(rng, -1e-3, 1e-3), rnd(rng, -1e-3, 1e-3)]
    d = lb.direction(1, *c, *g, None, vs, rho)
    lb.save(jac, None, d[1], d[2])
    assert np.array_equal(np.fromfile(f"{jac}_p_vs.old", np.float32), ((vs - F32(0.0)) / c[1]).astype(F32).ravel())
    assert np.array_equal(np.fromfile(f"{jac}_p_mrho.old", np.float32), ((rho - F32(0.0)) / c[2]).astype(F32).ravel())
    assert np.array_equal(np.fromfile(f"{jac}_p_u.old", np.float32), (c[1] * g[1]).astype(F32).ravel())
    assert np.array_equal(np.fromfile(f"{jac}_c_rho.old", np.float32), d[2].ravel())
    for n in ("p_vp", "p", "c"):
        assert not (tmp_path / f"jac_{n}.old").exists()
    lb.close()


def test_model_update_sh():
    """calc_mat_change_test_SH: product = oracle; the vs and rho of the P/SV
    update with SCALERHO 1 (the SH rho step has no SCALERHO); eps_true is
    EPSILON_u."""
    rng = np.random.default_rng(35)
    for trial in range(5):
        vp, vs, rho = models(rng)
        vs[0, :2] = 0.0
        d = [rnd(rng) for _ in range(3)]
        par = dict(PAR, inv_vs_iter=2 if trial == 2 else 0, invmat1=3 if trial == 3 else 1)
        eps = [0.01, 0.5, 0.05, 0.3, 0.0][trial]
        got = dnfwi.model_update(dnfwi.update_par(**par), 1, eps, None, vs, rho, None, d[1], d[2])
        ref = fo.mat_change_test(par, 1, eps, None, vs, rho, None, d[1], d[2])
        full = fo.mat_change_test(dict(par, scalerho=1.0), 1, eps, vp, vs, rho, *d)
        assert got[0] is None and ref[0] is None
        for k in (1, 2):
            assert np.array_equal(got[k], ref[k]), trial
            assert np.array_equal(ref[k], full[k]), trial
        assert got[3] == float(ref[3]) == float(full[3])


def test_assemble_sh():
    """ass_gradSH on stacked sums: the product equals the oracle; INVMAT1 3
    gives mu' = -DT g_u and rho' = -DT g_rho; INVMAT1 1 GRAD_FORM 2 is the
    chain rule of those (2 rho vs mu', vs^2 mu' + rho'); INV_*_ITER zero
    the gradients; cells with vs = 0 keep the GRAD_FORM 1 sums."""
    rng = np.random.default_rng(37)
    _, vs, rho = models(rng)
    vs[1, :3] = 0.0
    gu, gr = rnd(rng, -1.0, 1.0), rnd(rng, -1.0, 1.0)
    dt = 1e-3
    for invmat1 in (1, 3):
        for gf in (1, 2):
            a = dnfwi.assemble_sh(dt, invmat1, gf, 1, 0, 0, vs, rho, gu, gr)
            b = fo.assemble_sh(dt, invmat1, gf, 1, 0, 0, vs, rho, gu, gr)
            for x, y in zip(a, b):
                assert np.array_equal(x, y), (invmat1, gf)
    mu3, rho3 = dnfwi.assemble_sh(dt, 3, 1, 1, 0, 0, vs, rho, gu, gr)
    assert np.array_equal(mu3, (F32(-dt) * gu).astype(F32)) and np.array_equal(rho3, (F32(-dt) * gr).astype(F32))
    v1, r1 = dnfwi.assemble_sh(dt, 1, 2, 1, 0, 0, vs, rho, gu, gr)
    v64, r64, m64 = vs.astype(np.float64), rho.astype(np.float64), mu3.astype(np.float64)
    assert rel_l2(v1, 2 * r64 * v64 * m64) < 1e-6
    assert rel_l2(r1, v64 * v64 * m64 + rho3) < 1e-6
    g1u, g1r = dnfwi.assemble_sh(dt, 1, 1, 1, 0, 0, vs, rho, gu, gr)
    assert np.array_equal(g1u[1, :3], gu[1, :3]) and np.array_equal(g1r[1, :3], gr[1, :3])
    z = dnfwi.assemble_sh(dt, 1, 1, 1, 2, 2, vs, rho, gu, gr)
    assert not z[0].any() and not z[1].any()


def test_energy_precond_sh():
    """grad_obj_sh.c:211-219: the vs and rho sums divided by We C^2, vp absent."""
    rng = np.random.default_rng(39)
    Ws = rnd(rng, 0.1, 2.0)
    g = [rnd(rng) for _ in range(3)]
    c = (F32(2500.0), F32(1400.0), F32(2100.0))
    We, gv, gs, gr = dnfwi.energy_precond(3, 10.0, Ws, None, *c, None, g[1], g[2])
    We3, _, gs3, gr3 = dnfwi.energy_precond(3, 10.0, Ws, None, *c, *g)
    assert gv is None and np.array_equal(gs, gs3) and np.array_equal(gr, gr3)
    oW, ov, os_, orr = fo.energy_precond(3, 10.0, Ws, None, *c, None, g[1], g[2])
    assert ov is None and np.array_equal(gs, os_) and np.array_equal(gr, orr)


def test_pcg_sh_without_vp():
    """store_PCG_SH / extract_PCG_SH (classes vs and rho): without g_vp the vs
    and rho directions and betas equal the three-class run with a zero vp
    class, in the product and in the oracle."""
    rng = np.random.default_rng(41)
    for pcg_beta in (1, 2, 3, 4):
        p2, p3, po = dnfwi.PCG(NX, NY), dnfwi.PCG(NX, NY), fo.PCG(NX, NY)
        for it in range(4):
            gg = [rnd(rng) for _ in range(3)]
            a = p2.direction(pcg_beta, None, gg[1], gg[2])
            b = p3.direction(pcg_beta, np.zeros((NX, NY), F32), gg[1], gg[2])
            o = po.direction(pcg_beta, None, gg[1], gg[2])
            assert a[0] is None and o[0] is None
            for k in (1, 2):
                assert np.array_equal(a[k], b[k]) and np.array_equal(a[k], o[k]), (pcg_beta, it, k)
            assert a[3][0] == 0.0 and a[3][1] == b[3][1] == o[3][1] and a[3][2] == b[3][2] == o[3][2]
        p2.close()
        p3.close()
