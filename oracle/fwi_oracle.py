"""CPU restatement of DENISE's FWI iteration logic (L-BFGS direction,
parabolic step length, bounded model update) in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the host FWI
functions of libdnio (include/denise_fwi.h).  The product never uses it.

Written from the reference's functions, one per function, keeping its
1-based slot bookkeeping and its single- and double-precision steps:
  - store_LBFGS / LBFGS / extract_LBFGS: src/PSV/store_LBFGS_PSV.c:37-205,
    src/LBFGS.c:26-95, src/PSV/extract_LBFGS_PSV.c:31-249, with the slot
    pointer of src/PSV/FWI_PSV.c:452-466 and dotp of src/dotp.c:12-40
    (a float accumulator, i.e. a sequential single-precision sum); with
    nclass 2 the acoustic src/AC/store_LBFGS_AC.c:37-152 and
    src/AC/extract_LBFGS_AC.c:28-57 ([rho | vp] vectors);
  - model_mean: src/PSV/FWI_PSV.c:540-580;
  - mat_change_test: src/PSV/calc_mat_change_test_PSV.c:10-251; without vs
    src/AC/calc_mat_change_test_AC.c:9-189;
  - opt_step: src/calc_opt_step.c:14-84 with src/solvelin.c:19-48;
  - step_length: src/PSV/step_length_est_psv.c:37-182;
  - rtm_adjoint_source: the MODE 2 branch of src/calc_res.c:416-446 and
    its L2 sum, src/calc_res.c:563-569;
  - energy_precond: src/eprecond1.c:9-60 with the gradient scaling of
    src/PSV/grad_obj_psv.c:219-238;
  - taper_hor: src/taper_grad.c:188-322 (SWS_TAPER_GRAD_HOR);
  - taper_vert: src/taper_grad.c:46-186 (SWS_TAPER_GRAD_VERT);
  - taper_sources: src/taper_grad.c:328-577 (SWS_TAPER_GRAD_SOURCES);
  - PCG: src/PCG.c:11-105 with descent, store_PCG_PSV and extract_PCG_PSV
    as src/PSV/FWI_PSV.c:637-662 calls them.
  - residual_gc: src/calc_res.c:130-157, 330-360, 414, 465-489 (LNORM 5);
  - time_window: src/time_window.c:7-146 (TIMEWIN 1, 2, 4) and
    src/calc_res.c:447-450 (TIMEWIN 5);
  - gauss_filter: src/gauss_filt.c:9-208 (SPATFILTER 3, through
    src/smooth_grad.c:17-33 at src/PSV/FWI_PSV.c:628-631);
  - gauss_filter_var: src/gauss_filt_var.c:12-240 (SPATFILTER 4);
  - median_filter: src/smooth_model.c:37-196 with src/median2d.c:9-44
    (MODEL_FILTER 1, called by src/PSV/FWI_PSV.c:752).
Vectors are numpy float32 arrays of shape (nx, ny) (the model-file order).
The pins of these functions against what the mathematics fixes (BFGS
matrix recursion, parabola vertex, limits) are in tests/test_fwi_host.py.
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32


def dotp(a, b):
    """src/dotp.c: tmp += vec1[i]*vec2[i] with float tmp, in index order."""
    prod = (a.ravel().astype(F32) * b.ravel().astype(F32)).astype(F32)
    if prod.size == 0:
        return F32(0.0)
    return F32(np.add.accumulate(prod, dtype=F32)[-1])


class LBFGS:
    """The reference's L-BFGS state: y, s of NLBFGS slots of 3*NX*NY floats
    ([rho | vs | vp]; nclass 2: 2*NX*NY floats [rho | vp], the acoustic
    problem), q, r, rho, alpha, the cyclic pointer, and the previous model /
    normalised gradients (the JACOBIAN_*.old files)."""

    def __init__(self, nx, ny, nlbfgs, nclass=3, sh=False):
        self.nclass = 2 if sh else nclass
        self.sh = sh
        nclass = self.nclass
        self.n = nclass * nx * ny
        self.nl = nlbfgs
        self.y = np.zeros((nlbfgs + 1, self.n), F32)   # row 0 unused: slots 1..NLBFGS
        self.s = np.zeros((nlbfgs + 1, self.n), F32)
        self.rho = np.zeros(nlbfgs + 1, F32)
        self.alpha = np.zeros(nlbfgs + 1, F32)
        self.q = np.zeros(self.n, F32)
        self.r = np.zeros(self.n, F32)
        self.pointer = 1
        self.gold = None
        self.mold = None

    def reset(self):
        """zero_LBFGS + LBFGS_pointer = 1 (FWI_PSV.c:807-811)."""
        for a in (self.y, self.s, self.rho, self.alpha, self.q, self.r):
            a[...] = 0
        self.pointer = 1

    def direction(self, it, c_vp, c_vs, c_rho, g_vp, g_vs, g_rho, vp, vs, rho):
        c_vp, c_vs, c_rho = F32(c_vp), F32(c_vs), F32(c_rho)
        # FWI_PSV.c:455-464
        if it > 2:
            self.pointer += 1
        if self.pointer > self.nl:
            self.pointer = 1
        # store_LBFGS_PSV.c:37-89: waveconv = C * waveconv; gradp = waveconv
        sh = self.sh
        ac = self.nclass == 2 and not sh
        w_vp = None if sh else (c_vp * g_vp.astype(F32)).astype(F32)
        w_vs = None if ac else (c_vs * g_vs.astype(F32)).astype(F32)
        w_rho = (c_rho * g_rho.astype(F32)).astype(F32)
        if sh:   # store_LBFGS_SH.c:36-134: [vs | rho], models (m - 0) / C
            gradp = np.concatenate([w_vs.ravel(), w_rho.ravel()])
            model = np.concatenate([((vs.astype(F32) - F32(0.0)) / c_vs).astype(F32).ravel(),
                                    ((rho.astype(F32) - F32(0.0)) / c_rho).astype(F32).ravel()]).astype(F32)
        elif ac:   # store_LBFGS_AC.c:63-150
            gradp = np.concatenate([w_rho.ravel(), w_vp.ravel()])
            model = np.concatenate([rho.ravel(), vp.ravel()]).astype(F32)
        else:
            gradp = np.concatenate([w_rho.ravel(), w_vs.ravel(), w_vp.ravel()])
            model = np.concatenate([rho.ravel(), vs.ravel(), vp.ravel()]).astype(F32)
        if it > 1:
            # store_LBFGS_PSV.c:91-205
            p = self.pointer
            self.y[p] = (gradp - self.gold).astype(F32)
            self.s[p] = (model - self.mold).astype(F32)
            self.q = gradp.copy()
            # LBFGS.c:26-95
            gamma = F32(dotp(self.y[p], self.s[p]) / dotp(self.y[p], self.y[p]))
            for k in range(1, self.nl + 1):
                sn = dotp(self.y[k], self.s[k])
                self.rho[k] = F32(1.0 / float(sn)) if abs(float(sn)) > 0.0 else F32(0.0)
            for k in range(self.nl, 0, -1):
                self.alpha[k] = F32(self.rho[k] * dotp(self.s[k], self.q))
                self.q = (self.q - (self.alpha[k] * self.y[k]).astype(F32)).astype(F32)
            self.r = (gamma * self.q).astype(F32)
            for k in range(1, self.nl + 1):
                beta = F32(self.rho[k] * dotp(self.y[k], self.r))
                self.r = (self.r + (self.s[k] * F32(self.alpha[k] - beta)).astype(F32)).astype(F32)
        # extract_LBFGS_PSV.c:80-249: old model and old normalised gradient
        self.mold = model.copy()
        self.gold = gradp.copy()
        if it > 1 and sh:
            # extract_LBFGS_SH.c:27-48: r itself, not scaled back
            nc = g_vs.size
            return None, self.r[:nc].reshape(g_vs.shape).copy(), self.r[nc:].reshape(g_rho.shape).copy()
        if it > 1:
            # extract_LBFGS_PSV.c:31-78
            nc = g_vp.size
            d_rho = self.r[:nc].reshape(g_rho.shape)
            d_vp = self.r[(self.nclass - 1) * nc:].reshape(g_vp.shape)
            d_vs = None if ac else (self.r[nc:2 * nc].reshape(g_vs.shape) * c_vs).astype(F32)
            return ((d_vp * c_vp).astype(F32), d_vs, (d_rho * c_rho).astype(F32))
        return w_vp, w_vs, w_rho


def model_mean(a):
    """Vp_avg += ppi[j][i] (float), then Vp_avg /= NXG*NYG."""
    s = F32(np.add.accumulate(a.ravel().astype(F32), dtype=F32)[-1])
    return F32(s / F32(a.size))


def _maxima(m, d):
    mmax = F32(max(0.0, float(m.max())))
    dmax = F32(np.abs(d).max())
    return mmax, dmax


def mat_change_test(par, it, eps_scale, vp, vs, rho, d_vp, d_vs, d_rho):
    """calc_mat_change_test_PSV for one test or final step.  par: dict with
    vpupperlim, vplowerlim, vsupperlim, vslowerlim, rhoupperlim,
    rholowerlim, scalerho, inv_vp_iter, inv_vs_iter, inv_rho_iter, invmat1.
    Returns (vp_new, vs_new, rho_new, eps_true).  vs None:
    calc_mat_change_test_AC (vp and rho; vs_new None, eps_true = EPSILON)."""
    if vs is None:
        return _mat_change_test_ac(par, it, eps_scale, vp, rho, d_vp, d_rho)
    if vp is None:
        return _mat_change_test_sh(par, it, eps_scale, vs, rho, d_vs, d_rho)
    eps_scale = F32(eps_scale)
    pimax, gmax = _maxima(vp, d_vp)
    umax, gmax_u = _maxima(vs, d_vs)
    rhomax, gmax_rho = _maxima(rho, d_rho)
    with np.errstate(divide="ignore", invalid="ignore"):
        eps = F32(eps_scale * F32(pimax / gmax))
        eps_u = F32(eps_scale * F32(umax / gmax_u))
        eps_rho = F32(F32(eps_scale * F32(rhomax / gmax_rho)) * F32(par["scalerho"]))
    if it < par["inv_vp_iter"]:
        eps = F32(0.0)
    if it < par["inv_vs_iter"]:
        eps_u = F32(0.0)
    if it < par["inv_rho_iter"]:
        eps_rho = F32(0.0)
    if par["invmat1"] not in (1, 3):
        return vp.copy(), vs.copy(), rho.copy(), eps_u
    p1 = (vp - (eps * d_vp).astype(F32)).astype(F32)
    u1 = (vs - (eps_u * d_vs).astype(F32)).astype(F32)
    r1 = (rho - (eps_rho * d_rho).astype(F32)).astype(F32)
    if par["invmat1"] == 1:
        p1 = np.where(p1 < F32(par["vplowerlim"]), vp, p1)
        p1 = np.where(p1 > F32(par["vpupperlim"]), vp, p1)
        u1 = np.where((u1 < F32(par["vslowerlim"])) & (u1.astype(np.float64) > 1e-6), vs, u1)
        u1 = np.where(u1 > F32(par["vsupperlim"]), vs, u1)
        r1 = np.where(r1 < F32(par["rholowerlim"]), rho, r1)
        r1 = np.where(r1 > F32(par["rhoupperlim"]), rho, r1)
    p1 = np.where(p1 < 0, vp, p1)
    u1 = np.where(u1 < 0, vs, u1)
    r1 = np.where(r1 < 0, rho, r1)
    return p1.astype(F32), u1.astype(F32), r1.astype(F32), eps_u


def _mat_change_test_ac(par, it, eps_scale, vp, rho, d_vp, d_rho):
    """calc_mat_change_test_AC.c:33-160: the steps of vp (EPSILON) and rho
    (EPSILON_rho, times SCALERHO), the limits for INVMAT1 1, no negative
    values; returns EPSILON as eps_true (:103)."""
    eps_scale = F32(eps_scale)
    pimax, gmax = _maxima(vp, d_vp)
    rhomax, gmax_rho = _maxima(rho, d_rho)
    with np.errstate(divide="ignore", invalid="ignore"):
        eps = F32(eps_scale * F32(pimax / gmax))
        eps_rho = F32(F32(eps_scale * F32(rhomax / gmax_rho)) * F32(par["scalerho"]))
    if it < par["inv_vp_iter"]:
        eps = F32(0.0)
    if it < par["inv_rho_iter"]:
        eps_rho = F32(0.0)
    if par["invmat1"] not in (1, 3):
        return vp.copy(), None, rho.copy(), eps
    p1 = (vp - (eps * d_vp).astype(F32)).astype(F32)
    r1 = (rho - (eps_rho * d_rho).astype(F32)).astype(F32)
    if par["invmat1"] == 1:
        p1 = np.where(p1 < F32(par["vplowerlim"]), vp, p1)
        p1 = np.where(p1 > F32(par["vpupperlim"]), vp, p1)
        r1 = np.where(r1 < F32(par["rholowerlim"]), rho, r1)
        r1 = np.where(r1 > F32(par["rhoupperlim"]), rho, r1)
    p1 = np.where(p1 < 0, vp, p1)
    r1 = np.where(r1 < 0, rho, r1)
    return p1.astype(F32), None, r1.astype(F32), eps


def _mat_change_test_sh(par, it, eps_scale, vs, rho, d_vs, d_rho):
    """calc_mat_change_test_SH.c:29-170: the steps of vs (EPSILON_u) and rho
    (EPSILON_rho, no SCALERHO), the limits for INVMAT1 1 (the vs lower limit
    above 1e-6 only), no negative values; eps_true = EPSILON_u."""
    eps_scale = F32(eps_scale)
    umax, gmax_u = _maxima(vs, d_vs)
    rhomax, gmax_rho = _maxima(rho, d_rho)
    with np.errstate(divide="ignore", invalid="ignore"):
        eps_u = F32(eps_scale * F32(umax / gmax_u))
        eps_rho = F32(eps_scale * F32(rhomax / gmax_rho))
    if it < par["inv_vs_iter"]:
        eps_u = F32(0.0)
    if it < par["inv_rho_iter"]:
        eps_rho = F32(0.0)
    if par["invmat1"] not in (1, 3):
        return None, vs.copy(), rho.copy(), eps_u
    u1 = (vs - (eps_u * d_vs).astype(F32)).astype(F32)
    r1 = (rho - (eps_rho * d_rho).astype(F32)).astype(F32)
    if par["invmat1"] == 1:
        u1 = np.where((u1 < F32(par["vslowerlim"])) & (u1.astype(np.float64) > 1e-6), vs, u1)
        u1 = np.where(u1 > F32(par["vsupperlim"]), vs, u1)
        r1 = np.where(r1 < F32(par["rholowerlim"]), rho, r1)
        r1 = np.where(r1 > F32(par["rhoupperlim"]), rho, r1)
    u1 = np.where(u1 < 0, vs, u1)
    r1 = np.where(r1 < 0, rho, r1)
    return None, u1.astype(F32), r1.astype(F32), eps_u


def assemble_sh(dt, invmat1, grad_form, it, inv_vs_iter, inv_rho_iter, vs, rho, g_u, g_rho):
    """ass_gradSH.c:19-96 on the stacked sums (nx, ny): (g_vs, g_rho)."""
    dt = F32(dt)
    vs = vs.astype(F32)
    rho = rho.astype(F32)
    wmu = ((-dt) * g_u.astype(F32)).astype(F32)
    gu = g_u.astype(F32).copy()
    gr = g_rho.astype(F32).copy()
    rho_s = ((-dt) * g_rho.astype(F32)).astype(F32)
    v64, r64, w64 = vs.astype(np.float64), rho.astype(np.float64), wmu.astype(np.float64)
    pos = vs > 0
    if invmat1 == 1 and grad_form == 1:
        vs3 = ((vs * vs).astype(F32) * vs).astype(F32).astype(np.float64)
        gu = np.where(pos, ((2.0 / r64 * vs3) * w64).astype(F32), gu)
        with np.errstate(divide="ignore"):
            den = ((((vs * vs).astype(F32) * rho).astype(F32)) * rho).astype(F32).astype(np.float64)
            vs2rho2 = (1.0 / den).astype(F32)
        gr = np.where(pos, ((vs2rho2 * wmu).astype(F32) + rho_s).astype(F32), gr)
    elif invmat1 == 1:
        gu = (((2.0 * r64) * v64) * w64).astype(F32)
        gr = ((((vs * vs).astype(F32) * wmu).astype(F32)) + rho_s).astype(F32)
    elif invmat1 == 3:
        gu = wmu
        gr = rho_s
    if it < inv_vs_iter:
        gu = np.zeros_like(gu)
    if it < inv_rho_iter:
        gr = np.zeros_like(gr)
    return gu.astype(F32), gr.astype(F32)


def opt_step(L2t, epst):
    """calc_opt_step with solvelin (method 1); L2t, epst: 3 values each
    (the reference's [1..3])."""
    A = np.zeros((4, 4), F32)
    b = np.zeros(4, F32)
    x = np.zeros(4, F32)
    for i in range(1, 4):
        e = F32(epst[i - 1])
        A[i, 3] = F32(e * e)
        A[i, 2] = e
        A[i, 1] = F32(1.0)
        b[i] = F32(L2t[i - 1])
    for k in range(1, 3):
        for n in range(k, 3):
            a = F32(A[n + 1, k] / A[k, k])
            for m in range(1, 4):
                A[n + 1, m] = F32(A[n + 1, m] - F32(a * A[k, m]))
            b[n + 1] = F32(b[n + 1] - F32(a * b[k]))
    for k in range(3, 0, -1):
        c = b[k]
        for m in range(3, k, -1):
            c = F32(c - F32(A[k, m] * x[m]))
        x[k] = F32(c / A[k, k])
    opteps = F32(-float(x[2]) / (2.0 * float(x[3])))
    if (2.0 * float(x[3])) < 0.0 and L2t[1] > L2t[2]:
        opteps = F32(epst[2])
    if opteps > epst[2]:
        opteps = F32(epst[2])
    if opteps < 0.0:
        opteps = F32(epst[0])
    return F32(opteps)


def step_length(par, it, eps_scale0, scalefac, stepmax, L2t0, vp, vs, rho, d_vp, d_vs, d_rho, misfit,
                carry=None):
    """step_length_est_psv (vs None: step_length_est_ac, the misfit gets vs
    None).  misfit(vp, vs, rho, itest) -> float.  carry:
    (epst[3], L2t[3]) of the previous iteration (the reference's arrays
    persist), zeros at the first.  Returns a dict like dnfwi_linesearch."""
    epst1 = [0.0, 0.0, 0.0, 0.0]
    L2t = [0.0, float(L2t0), 0.0, 0.0]
    if carry is not None:
        epst1[1:] = [F32(v) for v in carry[0]]
        L2t[2:] = [float(v) for v in carry[1][1:]]
    scalefac = F32(scalefac)
    eps_scale = F32(eps_scale0)
    step1 = step2 = step3 = countstep = 0
    itests = iteste = 2
    trials = []
    while step2 != 1 or step1 != 1:
        for itest in range(itests, iteste + 1):
            tv, ts, tr, _ = mat_change_test(par, it, eps_scale, vp, vs, rho, d_vp, d_vs, d_rho)
            L2t[itest] = float(misfit(tv, ts, tr, itest))
            trials.append(float(eps_scale))
            epst1[itest] = F32(eps_scale)
            epst1[1] = F32(0.0)
        if step1 == 0 and L2t[1] <= L2t[2]:
            eps_scale = F32(eps_scale / scalefac)
            countstep += 1
        if step1 == 1 and L2t[2] < L2t[3]:
            epst1[3] = F32(eps_scale)
            step2 = 1
        if step1 == 1 and L2t[2] >= L2t[3]:
            epst1[3] = F32(eps_scale)
            eps_scale = F32(eps_scale + F32(eps_scale / scalefac))
            countstep += 1
        if step1 == 0 and L2t[1] > L2t[2]:
            epst1[2] = F32(eps_scale)
            step1 = 1
            iteste = itests = 3
            countstep = 0
            eps_scale = F32(eps_scale + F32(eps_scale / scalefac))
        step3 = 0
        if step1 == 0 and countstep > stepmax:
            step3 = 1
            break
        if step1 == 1 and countstep > stepmax:
            step1 = 1
            step2 = 1
    if step1 == 1:
        eps_scale = opt_step(L2t[1:4], epst1[1:4])
    return {"eps_scale": F32(eps_scale), "epst": [F32(v) for v in epst1[1:4]], "L2t": L2t[1:4],
            "step1": step1, "step3": step3, "nevals": len(trials), "trials": trials}


def rtm_adjoint_source(data, syn, dt, grad_form=1, testshot=1):
    """calc_res for MODE 2, LNORM 2: sectiondata[i][1] = section[i][1];
    intseis += DT*sectiondata[i][j] (GRAD_FORM 1) or intseis =
    sectiondata[i][j] (GRAD_FORM 2); sectiondiff[i][invtime] = intseis with
    invtime = ns - j + 1.  Returns (diff, L2)."""
    ntr, ns = data.shape
    diff = np.zeros((ntr, ns), F32)
    L2 = 0.0
    dt = F32(dt)
    for i in range(ntr):
        sd = data[i].astype(F32).copy()
        sd[0] = syn[i, 0]
        intseis = F32(0.0)
        for j in range(1, ns + 1):
            if grad_form == 1:
                intseis = F32(intseis + F32(dt * sd[j - 1]))
            else:
                intseis = sd[j - 1]
            diff[i, ns - j] = intseis
        if testshot:
            for invtime in range(ns, 0, -1):
                L2 += 0.5 * float(diff[i, invtime - 1]) * float(diff[i, invtime - 1])
    return diff, L2


def energy_precond(eprecond, dh, Ws, Wr, c_vp, c_vs, c_rho, g_vp, g_vs, g_rho):
    """Returns (We, g_vp, g_vs, g_rho) as new float32 arrays; (nx, ny) inputs.
    g_vs None: the acoustic grad_obj_ac.c:225-233 (g_vs stays None)."""
    import math
    nx, ny = Ws.shape
    dh = F32(dh)
    xmin, xmax = dh, F32(F32(nx) * dh)
    We = np.zeros((nx, ny), F32)
    maxwe = F32(0.0)
    for i in range(1, nx + 1):
        for j in range(1, ny + 1):
            if eprecond == 1:
                w = F32(math.sqrt(float(F32(Ws[i - 1, j - 1] * Wr[i - 1, j - 1]))))
            else:
                x, y = F32(F32(i) * dh), F32(F32(j) * dh)
                w = F32(math.sqrt(float(Ws[i - 1, j - 1])) *
                        (math.asinh(float(F32((xmax - x) / y))) - math.asinh(float(F32((xmin - x) / y)))))
            We[i - 1, j - 1] = w
            if w > maxwe:
                maxwe = w
    We = (We + F32(F32(0.005) * maxwe)).astype(F32)
    c_vp, c_vs, c_rho = F32(c_vp), F32(c_vs), F32(c_rho)
    gv = None if g_vp is None else (g_vp / ((We * c_vp).astype(F32) * c_vp).astype(F32)).astype(F32)
    gs = None
    if g_vs is not None:
        gs = (g_vs / ((We * c_vs).astype(F32) * c_vs).astype(F32)).astype(F32)
        if c_vs == 0.0:
            gs[...] = 0.0
    gr = (g_rho / ((We * c_rho).astype(F32) * c_rho).astype(F32)).astype(F32)
    return We, gv, gs, gr


def _seqsum(x):
    """Sequential float32 sum (a float accumulator in index order)."""
    return np.cumsum(np.asarray(x, np.float32), dtype=np.float32)[-1] if len(x) else F32(0.0)


def residual_gc(data, syn, dt, grad_form, testshot, skip=None, l2=0.0):
    """LNORM 5 residuals (time-reversed (ntr, ns)) and the running misfit."""
    data = np.asarray(data, np.float32)
    syn = np.asarray(syn, np.float32)
    ntr, ns = syn.shape
    out = np.zeros((ntr, ns), np.float32)
    L2 = float(l2)
    for i in range(ntr):
        if skip is not None and skip[i] == 1:
            continue
        if grad_form == 1:
            d = (np.cumsum(data[i], dtype=np.float32) * F32(dt)).astype(np.float32)
            s = (np.cumsum(syn[i], dtype=np.float32) * F32(dt)).astype(np.float32)
        else:
            d, s = data[i], syn[i]
        abs_d = F32(math.sqrt(float(_seqsum((d * d).astype(np.float32)))))
        abs_s = F32(math.sqrt(float(_seqsum((s * s).astype(np.float32)))))
        dms = _seqsum((s * d).astype(np.float32))
        dj = d.copy()
        if grad_form == 2:
            dj[0] = syn[i, 0]                      # calc_res.c:414
        tmp = F32(F32(F32(abs_s * abs_s) * abs_s) * abs_d)
        tmp1 = F32(abs_s * abs_d)
        r = ((s * dms).astype(np.float32) / tmp).astype(np.float32) - (dj / tmp1).astype(np.float32)
        out[i] = r.astype(np.float32)[::-1]
        if testshot:
            terms = ((dj * s).astype(np.float32) / F32(abs_d * abs_s)).astype(np.float32)
            for v in terms.tolist():
                L2 -= v
    return out, L2


def time_window(timewin, diff, dt, gamma, twlength_minus, twlength_plus, picks=None, offset=None):
    """The windowed (ntr, ns) time-reversed residuals; column k holds time
    (ns - k) DT (calc_res.c:423, 441).  Exponents in float32, exp in
    double, as the C expressions do."""
    d = np.array(diff, np.float32, copy=True)
    ntr, ns = d.shape
    t = (np.arange(ns, 0, -1).astype(np.float32) * F32(dt)).astype(np.float32)
    g32 = F32(gamma)
    for i in range(ntr):
        if timewin == 5:
            t0 = F32(abs(F32(F32(offset[i]) / F32(twlength_minus))))
            x = (t - t0).astype(np.float32).astype(np.float64)
            fac = np.array([math.exp(-float(g32) * (v * v)) for v in x], np.float64)
            on = t >= t0
            d[i, on] = (d[i, on].astype(np.float64) * fac[on]).astype(np.float32)
        elif timewin in (1, 2, 4):
            pt = F32(-F32(dt)) if timewin == 2 else F32(F32(picks[i]) + F32(twlength_plus))
            pt1 = F32(twlength_plus) if timewin == 2 else F32(picks[i])
            dump = (t - pt).astype(np.float32)
            ex = ((-g32 * dump).astype(np.float32) * dump).astype(np.float32)
            taper = np.array([math.exp(float(v)) for v in ex], np.float64).astype(np.float32)
            on = t >= (pt1 if timewin == 2 else pt)
            d[i, on] = (d[i, on] * taper[on]).astype(np.float32)
    return d


def gauss_filter(g, dh, wd_damp, c_vs, fc_end):
    """The Gaussian gradient smoother: (filtered (nx, ny) float32, width).
    Kernel values are float32 of exp (C libm through math.exp) of a float32
    exponent; the convolution adds, per cell, the products of the
    edge-extended array and the kernel in the order x offset outer, y offset
    inner, in float32 (gauss_filt.c:117-147)."""
    import math
    g = np.asarray(g, np.float32)
    lam = F32(F32(c_vs) / F32(fc_end))
    x = float(F32(F32(F32(wd_damp) * lam) / F32(dh)))
    fs = int(math.copysign(math.floor(abs(x) + 0.5), x))      # C round(): halves away from 0
    if fs == 0:
        return g.copy(), 0
    if fs % 2 == 0:
        fs += 1 if fs > 0 else -1
    h = abs(fs) // 2
    sigma = F32(h // 2)
    s = F32(2.0 * float(sigma) * float(sigma))
    k = np.empty((2 * h + 1, 2 * h + 1), np.float32)
    tot = F32(0.0)
    for di in range(-h, h + 1):
        for dj in range(-h, h + 1):
            with np.errstate(invalid="ignore", divide="ignore"):
                e = F32(-(F32(di * di) / s) - (F32(dj * dj) / s))
            k[di + h, dj + h] = F32(math.exp(float(e))) if not np.isnan(e) else np.float32(np.nan)
            tot = F32(tot + k[di + h, dj + h])
    if fs > 0:
        k = (k / tot).astype(np.float32)
    nx, ny = g.shape
    pad = np.pad(g, h, mode="edge")
    conv = np.zeros_like(g)
    for di in range(-h, h + 1):
        for dj in range(-h, h + 1):
            conv = (conv + pad[h + di:h + di + nx, h + dj:h + dj + ny] * k[di + h, dj + h]).astype(np.float32)
    return conv, fs


def _c_round(x):
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


def gauss_filter_var(g, vel, dh, wd_damp, wd_damp1, fc_end):
    """The velocity-dependent Gaussian smoother, cell by cell in plain
    loops (small arrays): windows from the local wavelength, kernel sums in
    float32 in the order x offset outer, y offset inner."""
    g = np.asarray(g, np.float32)
    vel = np.asarray(vel, np.float32)
    nx, ny = g.shape
    out = g.copy()
    for i in range(nx):
        for j in range(ny):
            lam = F32(vel[i, j] / F32(fc_end))
            wx = _c_round(float(F32(F32(F32(wd_damp) * lam) / F32(dh))))
            wz = _c_round(float(F32(F32(F32(wd_damp1) * lam) / F32(dh))))
            if wz <= 0:
                continue
            wx += 1 - wx % 2
            wz += 1 - wz % 2
            wx, wz = max(wx, 6), max(wz, 6)
            hx, hz = abs(wx) // 2, abs(wz) // 2
            sx = F32(2.0 * float(hx // 2) ** 2)
            sz = F32(2.0 * float(hz // 2) ** 2)
            conv = F32(0.0)
            tot = F32(0.0)
            for di in range(-hx, hx + 1):
                ii = min(max(i + di, 0), nx - 1)
                for dj in range(-hz, hz + 1):
                    jj = min(max(j + dj, 0), ny - 1)
                    k = F32(math.exp(float(F32(-(F32(di * di) / sx) - (F32(dj * dj) / sz)))))
                    conv = F32(conv + F32(g[ii, jj] * k))
                    tot = F32(tot + k)
            out[i, j] = F32(conv / tot)
    return out


def median_filter(a, filt_size):
    """Median of the |filt_size|^2 window (an even size widened by one)
    around each cell of the (nx, ny) array, the array extended by repeating
    its edge cells (smooth_model.c:40-136); filt_size 0 returns a copy."""
    a = np.asarray(a, np.float32)
    if filt_size == 0:
        return a.copy()
    fs = abs(filt_size)
    fs += 1 - fs % 2
    h = fs // 2
    padded = np.pad(a, h, mode="edge")
    win = np.lib.stride_tricks.sliding_window_view(padded, (fs, fs))
    return np.median(win.reshape(a.shape[0], a.shape[1], fs * fs), axis=2).astype(np.float32)


def taper_hor(nx, ny, dh, gradt1, gradt2, exp_taper):
    """The taper coefficients, (nx, ny) float32."""
    import math
    L = gradt2 - gradt1
    ifw = L + 1
    a = F32(3.0)
    window = []
    for i in range(1, ifw + 1):
        w = F32(1.0)
        if i <= L:
            q = float(F32(a * F32(i - L))) / (L / 2.0)
            w = F32(math.exp(-(1.0 / 2.0) * q * q))
        window.append(w)
    coeff = np.zeros((nx, ny), F32)
    h = 1
    for j in range(1, ny + 1):
        t = F32(0.0)
        if gradt1 <= j <= gradt2:
            t = window[h - 1]
        if j > gradt2:
            t = F32(math.pow(float(F32(F32(j) * F32(dh))), float(F32(exp_taper))))
        coeff[:, j - 1] = t
        if j >= gradt1:
            h += 1
    return coeff


class PCG:
    """PCG_old / PCG_dir of [vp | vs | rho]; direction() returns the
    arrays the model update subtracts (-PCG_dir) and the three betas."""

    def __init__(self, nx, ny):
        self.n = nx * ny
        self.old = np.zeros(3 * self.n, F32)
        self.dir = np.zeros(3 * self.n, F32)

    def reset(self):
        self.old[...] = 0
        self.dir[...] = 0

    def direction(self, pcg_beta, g_vp, g_vs, g_rho):
        """g_vs None: the acoustic store_PCG_AC / extract_PCG_AC (vp and rho
        classes, the vs class left 0 and None on return)."""
        ac = g_vs is None
        if ac:
            return self._direction_ac(pcg_beta, g_vp, g_rho)
        if g_vp is None:   # SH (store_PCG_SH: vs and rho): an all-zero vp class keeps dir 0, beta 0
            d = self.direction(pcg_beta, np.zeros_like(g_vs), g_vs, g_rho)
            return (None, d[1], d[2], d[3])
        shape = g_vp.shape
        new = np.concatenate([-g_vp.ravel(), -g_vs.ravel(), -g_rho.ravel()]).astype(F32)
        betas = []
        for k in range(3):
            sl = slice(k * self.n, (k + 1) * self.n)
            nw, od, dr = new[sl], self.old[sl], self.dir[sl]
            if pcg_beta == 1:
                tz, tn = (nw * nw).astype(F32), (od * od).astype(F32)
            elif pcg_beta == 2:
                tz, tn = (nw * (nw - od).astype(F32)).astype(F32), (od * od).astype(F32)
            elif pcg_beta == 3:
                tz, tn = ((-nw) * (nw - od).astype(F32)).astype(F32), (dr * (nw - od).astype(F32)).astype(F32)
            else:
                tz, tn = ((-nw) * nw).astype(F32), (dr * (nw - od).astype(F32)).astype(F32)
            bz = np.add.accumulate(tz, dtype=F32)[-1]
            bn = np.add.accumulate(tn, dtype=F32)[-1]
            b = F32(0.0)
            if bn != 0:
                b = F32(bz / bn)
            if b < 0.0 and pcg_beta == 2:
                b = F32(0.0)
            betas.append(b)
            self.dir[sl] = (nw + (b * dr).astype(F32)).astype(F32)
        self.old = new.copy()
        d = -self.dir
        return (d[:self.n].reshape(shape), d[self.n:2 * self.n].reshape(shape), d[2 * self.n:].reshape(shape),
                np.array(betas, F32))

    def _direction_ac(self, pcg_beta, g_vp, g_rho):
        """PCG.c with PCG_class 2 over [vp | rho] (store_PCG_AC.c:11-44),
        kept in the vp and rho thirds of the state."""
        shape = g_vp.shape
        n = self.n
        sl = [slice(0, n), slice(2 * n, 3 * n)]
        new = [(-g_vp.ravel()).astype(F32), (-g_rho.ravel()).astype(F32)]
        betas = []
        for k in range(2):
            nw, od, dr = new[k], self.old[sl[k]], self.dir[sl[k]]
            if pcg_beta == 1:
                tz, tn = (nw * nw).astype(F32), (od * od).astype(F32)
            elif pcg_beta == 2:
                tz, tn = (nw * (nw - od).astype(F32)).astype(F32), (od * od).astype(F32)
            elif pcg_beta == 3:
                tz, tn = ((-nw) * (nw - od).astype(F32)).astype(F32), (dr * (nw - od).astype(F32)).astype(F32)
            else:
                tz, tn = ((-nw) * nw).astype(F32), (dr * (nw - od).astype(F32)).astype(F32)
            bz = np.add.accumulate(tz, dtype=F32)[-1]
            bn = np.add.accumulate(tn, dtype=F32)[-1]
            b = F32(0.0)
            if bn != 0:
                b = F32(bz / bn)
            if b < 0.0 and pcg_beta == 2:
                b = F32(0.0)
            betas.append(b)
            self.dir[sl[k]] = (nw + (b * dr).astype(F32)).astype(F32)
            self.old[sl[k]] = nw
        return ((-self.dir[sl[0]]).reshape(shape), None, (-self.dir[sl[1]]).reshape(shape),
                np.array([betas[0], F32(0.0), betas[1]], F32))


def taper_vert(nx, ny, dh, gradt, srcpos, recpos, g):
    """(coeff, tapered g); srcpos rows 0/1 = x/y in metres, recpos rows 0/1 =
    grid indices."""
    import math
    g1, g2, g3, g4 = [int(v) for v in gradt]
    tl, tl2, ifw = g2 - g1, g4 - g3, g4 - g1
    a = F32(3.0)
    window = []
    for i in range(1, ifw + 1):
        w = F32(1.0)
        if i <= tl:
            q = float(F32(a * F32(i - tl))) / (tl / 2.0)
            w = F32(math.exp(-(1.0 / 2.0) * q * q))
        if i >= ifw - tl2:
            q = float(F32(a * F32(i - (ifw - tl2)))) / (tl2 / 2.0)
            w = F32(math.exp(-(1.0 / 2.0) * q * q))
        window.append(w)
    row = np.zeros(nx, F32)
    h = 1
    for i in range(1, nx + 1):
        if g1 < i < g4:
            row[i - 1] = window[h - 1]
            h += 1
    coeff = np.repeat(row[:, None], ny, axis=1).astype(F32)
    out = (g * coeff).astype(F32)
    for n in range(srcpos.shape[1]):
        i = int(math.floor(float(F32(F32(srcpos[0, n]) / F32(dh))) + 0.5))
        j = int(math.floor(float(F32(F32(srcpos[1, n]) / F32(dh))) + 0.5))
        if 1 <= i <= nx and 1 <= j <= ny:
            out[i - 1, j - 1] = 0.0
    for n in range(recpos.shape[1]):
        i, j = int(recpos[0, n]), int(recpos[1, n])
        if 1 <= i <= nx and 1 <= j <= ny:
            out[i - 1, j - 1] = 0.0
    return coeff, out


def taper_sources(nx, ny, dh, srtradius, srtshape, filtsize, srcpos, g):
    """(coeff, tapered g) of the cylindrical source taper."""
    import math
    dh = F32(dh)
    gp = int(math.ceil(2.0 * float(F32(srtradius)) / float(dh)))
    gp = max(gp, 5)
    maxrad = F32(math.sqrt(2.0 * float(F32(srtradius)) * float(F32(srtradius))))
    m = np.zeros((gp + 1, gp + 1), F32)          # 1-based [j][i]
    for j in range(1, gp + 1):
        for i in range(1, gp + 1):
            x = F32((float(F32(i)) - float(F32(gp)) / 2.0 - 0.5) * float(dh))
            y = F32((float(F32(j)) - float(F32(gp)) / 2.0 - 0.5) * float(dh))
            rad = F32(math.sqrt(float(F32(F32(x * x) + F32(y * y)))))
            if srtshape == 1:
                m[j, i] = F32(math.erf(float(F32(F32(F32(1.0) * rad) / maxrad))))
            else:
                m[j, i] = F32(math.log(float(rad))) if rad > 0 else F32(0.0)
    core = m[1:, 1:]
    core -= F32(core.min())
    edge = [m[1, gp // 2], m[gp // 2, 1], m[gp // 2, gp], m[gp, gp // 2]]
    maxm = F32(max(edge))
    core /= maxm
    core[core > 1.0] = 1.0
    ijc = int(math.ceil(float(F32(F32(gp) / F32(2)))))
    msum = np.ones((nx, ny), F32)
    cells = []
    for n in range(srcpos.shape[1]):
        i = int(math.floor(float(F32(F32(srcpos[0, n]) / dh)) + 0.5))
        j = int(math.floor(float(F32(F32(srcpos[1, n]) / dh)) + 0.5))
        cells.append((i, j))
        for iy in range(1, gp + 1):
            for ix in range(1, gp + 1):
                xx, yy = i + ix - ijc, j + iy - ijc
                if 1 <= xx <= nx and 1 <= yy <= ny:
                    msum[xx - 1, yy - 1] = min(msum[xx - 1, yy - 1], m[iy, ix])
    msum = (msum - F32(msum.min())).astype(F32)
    coeff = (msum / F32(msum.max())).astype(F32)
    out = (g * coeff).astype(F32)
    for (i1, j1) in cells:
        for i in range(i1 - filtsize, i1 + filtsize + 1):
            for j in range(j1 - filtsize, j1 + filtsize + 1):
                if 1 <= i <= nx and 1 <= j <= ny:
                    out[i - 1, j - 1] = 0.0
                    coeff[i - 1, j - 1] = 0.0
    return coeff, out
