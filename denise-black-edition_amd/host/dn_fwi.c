/*
 * Host side of the FWI iteration (include/denise_fwi.h): L-BFGS search
 * direction, parabolic step-length search, bounded model update.
 *
 * Each function evaluates the reference's expressions in the reference's
 * order and precision (single-precision accumulators where the reference
 * has them, double where a double literal promotes), so the results equal
 * the reference's bit for bit; the build compiles this file with
 * -ffp-contract=off (no fused multiply-add).
 *
 * Index conventions: the reference's NR vectors are 1-based
 * (y_LBFGS[1..NLBFGS*NLBFGS_vec]); here every array is 0-based and slot k
 * (1..NLBFGS) starts at (k-1)*nvec.
 */
#include "denise_fwi.h"
#include "denise_io.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int dnio_fail_msg(const char *msg); /* dn_io.c */

struct dnfwi_lbfgs {
    int nx, ny, nlbfgs, pointer;
    int nclass;                  /* 3: [rho | vs | vp]; 2 (acoustic): [rho | vp]; 2 (SH): [vs | rho] */
    int sh;                      /* SH: models normalised by C, the direction left normalised */
    size_t ncell, nvec;          /* nvec = nclass * ncell */
    float *y, *s;                /* nlbfgs * nvec */
    float *q, *r;                /* nvec */
    float *rho_l, *alpha_l;      /* nlbfgs */
    float *gold, *mold;          /* previous normalised gradients / model, [rho | vs | vp] */
    int have_old;
};

static int lbfgs_create(int nx, int ny, int nlbfgs, int nclass, dnfwi_lbfgs **out)
{
    if (!out || nx < 1 || ny < 1 || nlbfgs < 1) return dnio_fail_msg("dnfwi_lbfgs_create: bad size");
    dnfwi_lbfgs *lb = calloc(1, sizeof *lb);
    if (!lb) return dnio_fail_msg("dnfwi_lbfgs_create: out of host memory");
    lb->nx = nx;
    lb->ny = ny;
    lb->nlbfgs = nlbfgs;
    lb->nclass = nclass;
    lb->ncell = (size_t)nx * ny;
    lb->nvec = (size_t)nclass * lb->ncell;
    /* FWI_PSV.c:289-298 / FWI_AC.c:269-283: NR vectors, zero-initialised */
    lb->y = calloc((size_t)nlbfgs * lb->nvec, sizeof(float));
    lb->s = calloc((size_t)nlbfgs * lb->nvec, sizeof(float));
    lb->q = calloc(lb->nvec, sizeof(float));
    lb->r = calloc(lb->nvec, sizeof(float));
    lb->gold = calloc(lb->nvec, sizeof(float));
    lb->mold = calloc(lb->nvec, sizeof(float));
    lb->rho_l = calloc((size_t)nlbfgs, sizeof(float));
    lb->alpha_l = calloc((size_t)nlbfgs, sizeof(float));
    if (!lb->y || !lb->s || !lb->q || !lb->r || !lb->gold || !lb->mold || !lb->rho_l || !lb->alpha_l) {
        dnfwi_lbfgs_destroy(lb);
        return dnio_fail_msg("dnfwi_lbfgs_create: out of host memory");
    }
    lb->pointer = 1; /* FWI_PSV.c:290 */
    *out = lb;
    return 0;
}

int dnfwi_lbfgs_create(int nx, int ny, int nlbfgs, dnfwi_lbfgs **out)
{
    return lbfgs_create(nx, ny, nlbfgs, 3, out);
}

int dnfwi_lbfgs_create_ac(int nx, int ny, int nlbfgs, dnfwi_lbfgs **out)
{
    return lbfgs_create(nx, ny, nlbfgs, 2, out); /* FWI_AC.c:271: NLBFGS_class = 2 */
}

int dnfwi_lbfgs_create_sh(int nx, int ny, int nlbfgs, dnfwi_lbfgs **out)
{
    int rc = lbfgs_create(nx, ny, nlbfgs, 2, out); /* FWI_SH.c: NLBFGS_class = 2 */
    if (!rc) (*out)->sh = 1;
    return rc;
}

void dnfwi_lbfgs_destroy(dnfwi_lbfgs *lb)
{
    if (!lb) return;
    free(lb->y);
    free(lb->s);
    free(lb->q);
    free(lb->r);
    free(lb->gold);
    free(lb->mold);
    free(lb->rho_l);
    free(lb->alpha_l);
    free(lb);
}

void dnfwi_lbfgs_reset(dnfwi_lbfgs *lb)
{
    if (!lb) return;
    /* zero_LBFGS.c:13-30; the old gradient/model files of the reference stay */
    memset(lb->y, 0, sizeof(float) * (size_t)lb->nlbfgs * lb->nvec);
    memset(lb->s, 0, sizeof(float) * (size_t)lb->nlbfgs * lb->nvec);
    memset(lb->q, 0, sizeof(float) * lb->nvec);
    memset(lb->r, 0, sizeof(float) * lb->nvec);
    memset(lb->rho_l, 0, sizeof(float) * (size_t)lb->nlbfgs);
    memset(lb->alpha_l, 0, sizeof(float) * (size_t)lb->nlbfgs);
    lb->pointer = 1;
}

int dnfwi_lbfgs_pointer(const dnfwi_lbfgs *lb) { return lb ? lb->pointer : 0; }

/* dotp.c:12-40: single-precision sequential sum of products over n values */
static float dotp(const float *a, const float *b, size_t n)
{
    float tmp = 0.0f;
    for (size_t i = 0; i < n; i++) tmp += a[i] * b[i];
    return tmp;
}

int dnfwi_lbfgs_direction(dnfwi_lbfgs *lb, int iter, float c_vp, float c_vs, float c_rho, float *g_vp,
                          float *g_vs, float *g_rho, const float *vp, const float *vs, const float *rho)
{
    const int sh = lb && lb->sh, ac = lb && lb->nclass == 2 && !sh;
    if (!lb || !g_rho || !rho || (!sh && (!g_vp || !vp)) || (!ac && (!g_vs || !vs)) || iter < 1)
        return dnio_fail_msg("dnfwi_lbfgs_direction: null argument or iter < 1");
    if (iter > 1 && !lb->have_old)
        return dnio_fail_msg("dnfwi_lbfgs_direction: iter > 1 needs the previous iteration's gradients");
    const size_t nc = lb->ncell, nv = lb->nvec;
    const int nl = lb->nlbfgs, ncl = lb->nclass;
    /* vector order of the L-BFGS vectors: [rho | vs | vp] (store_LBFGS_PSV.c)
     * or [rho | vp] (store_LBFGS_AC.c:63-140) */
    float *g[3] = {g_rho, ac ? g_vp : g_vs, g_vp};
    const float *m[3] = {rho, ac ? vp : vs, vp};
    float cs[3] = {c_rho, ac ? c_vp : c_vs, c_vp};
    if (sh) { /* store_LBFGS_SH.c:36-134: [vs | rho] */
        g[0] = g_vs;
        g[1] = g_rho;
        m[0] = vs;
        m[1] = rho;
        cs[0] = c_vs;
        cs[1] = c_rho;
    }

    /* FWI_PSV.c:455-464 */
    if (iter > 2) lb->pointer++;
    if (lb->pointer > nl) lb->pointer = 1;

    /* store_LBFGS_PSV.c:37-89: normalisation; gradp = the normalised gradient */
    for (int c = 0; c < ncl; c++)
        for (size_t k = 0; k < nc; k++) g[c][k] = cs[c] * g[c][k];

    if (iter > 1) {
        /* store_LBFGS_PSV.c:91-205: y, s into the current slot; q = gradient */
        float *ys = lb->y + (size_t)(lb->pointer - 1) * nv, *ss = lb->s + (size_t)(lb->pointer - 1) * nv;
        for (int c = 0; c < ncl; c++)
            for (size_t k = 0; k < nc; k++) {
                ys[c * nc + k] = g[c][k] - lb->gold[c * nc + k];
                /* store_LBFGS_SH.c:81, :104: the model normalised, (m - C_min) / C with C_min = 0 */
                const float mk = sh ? (m[c][k] - 0.0f) / cs[c] : m[c][k];
                ss[c * nc + k] = mk - lb->mold[c * nc + k];
            }
        for (int c = 0; c < ncl; c++) memcpy(lb->q + c * nc, g[c], sizeof(float) * nc);

        /* LBFGS.c:26-95 */
        const float *yp = lb->y + (size_t)(lb->pointer - 1) * nv, *sp = lb->s + (size_t)(lb->pointer - 1) * nv;
        float sum_nom = dotp(yp, sp, nv);
        float sum_denom = dotp(yp, yp, nv);
        const float gamma = sum_nom / sum_denom;
        for (int k = 1; k <= nl; k++) {
            sum_nom = dotp(lb->y + (size_t)(k - 1) * nv, lb->s + (size_t)(k - 1) * nv, nv);
            lb->rho_l[k - 1] = fabs(sum_nom) > 0.0 ? (float)(1.0 / sum_nom) : 0.0f;
        }
        for (int k = nl; k >= 1; k--) {
            const float *sk = lb->s + (size_t)(k - 1) * nv, *yk = lb->y + (size_t)(k - 1) * nv;
            sum_nom = dotp(sk, lb->q, nv);
            lb->alpha_l[k - 1] = lb->rho_l[k - 1] * sum_nom;
            const float a = lb->alpha_l[k - 1];
            for (size_t i = 0; i < nv; i++) lb->q[i] = lb->q[i] - a * yk[i];
        }
        for (size_t i = 0; i < nv; i++) lb->r[i] = gamma * lb->q[i];
        for (int k = 1; k <= nl; k++) {
            const float *sk = lb->s + (size_t)(k - 1) * nv, *yk = lb->y + (size_t)(k - 1) * nv;
            sum_nom = dotp(yk, lb->r, nv);
            const float beta = lb->rho_l[k - 1] * sum_nom;
            const float d = lb->alpha_l[k - 1] - beta;
            for (size_t i = 0; i < nv; i++) lb->r[i] = lb->r[i] + sk[i] * d;
        }
    }

    /* extract_LBFGS_PSV.c:80-249: keep the model and the normalised
     * gradients (before they are replaced by the direction) */
    for (int c = 0; c < ncl; c++) {
        if (sh) /* extract_LBFGS_SH.c:66-70: the normalised model */
            for (size_t k = 0; k < nc; k++) lb->mold[c * nc + k] = (m[c][k] - 0.0f) / cs[c];
        else
            memcpy(lb->mold + c * nc, m[c], sizeof(float) * nc);
        memcpy(lb->gold + c * nc, g[c], sizeof(float) * nc);
    }
    lb->have_old = 1;

    if (iter > 1 && sh) {
        /* extract_LBFGS_SH.c:27-48: the direction r, not scaled back */
        for (int c = 0; c < ncl; c++) memcpy(g[c], lb->r + c * nc, sizeof(float) * nc);
    } else if (iter > 1) {
        /* extract_LBFGS_PSV.c:31-78: direction = r, scaled back */
        for (int c = 0; c < ncl; c++) memcpy(g[c], lb->r + c * nc, sizeof(float) * nc);
        for (size_t k = 0; k < nc; k++) {
            g_vp[k] = g_vp[k] * c_vp;
            if (!ac) g_vs[k] = g_vs[k] * c_vs;
            g_rho[k] = g_rho[k] * c_rho;
        }
    }
    return 0;
}

int dnfwi_lbfgs_save(const dnfwi_lbfgs *lb, const char *jacobian, const float *d_vp, const float *d_vs,
                     const float *d_rho)
{
    const int sh = lb && lb->sh, ac = lb && lb->nclass == 2 && !sh;
    if (!lb || !jacobian || (!sh && !d_vp) || (!ac && !d_vs) || !d_rho)
        return dnio_fail_msg("dnfwi_lbfgs_save: null argument");
    if (!lb->have_old) return dnio_fail_msg("dnfwi_lbfgs_save: no direction taken yet");
    const size_t nc = lb->ncell;
    /* extract_LBFGS_PSV.c:80-249: per parameter the model, the normalised
     * gradient, then the direction; vp, vs, rho ([rho | vs | vp] slots 2, 1, 0).
     * extract_LBFGS_AC.c:60-178: vp, rho ([rho | vp] slots 1, 0) */
    static const char *const names[3][3] = {
        {"p_vp", "p", "c"}, {"p_vs", "p_u", "c_u"}, {"p_mrho", "p_rho", "c_rho"}};
    static const int slot[3] = {2, 1, 0};
    const float *dir[3] = {d_vp, d_vs, d_rho};
    char name[1024];
    for (int k = 0; k < 3; k++) {
        if ((ac && k == 1) || (sh && k == 0)) continue;
        /* extract_LBFGS_SH.c:50-178: vs, rho ([vs | rho] slots 0, 1) */
        const int sl = sh ? (k == 1 ? 0 : 1) : ac ? (k == 0 ? 1 : 0) : slot[k];
        const float *arr[3] = {lb->mold + (size_t)sl * nc, lb->gold + (size_t)sl * nc, dir[k]};
        for (int f = 0; f < 3; f++) {
            if ((size_t)snprintf(name, sizeof name, "%s_%s.old", jacobian, names[k][f]) >= sizeof name)
                return dnio_fail_msg("dnfwi_lbfgs_save: JACOBIAN path too long");
            if (dnio_write_model(name, arr[f], lb->nx, lb->ny)) return -1;
        }
    }
    return 0;
}

float dnfwi_model_mean(const float *a, int nx, int ny)
{
    /* FWI_PSV.c:549-569: Vp_avg += ppi[j][i] over i, j; Vp_avg /= NXG * NYG */
    float sum = 0.0f;
    const size_t n = (size_t)nx * ny;
    for (size_t k = 0; k < n; k++) sum += a[k];
    sum /= nx * ny;
    return sum;
}

/* calc_mat_change_test_PSV.c:41-104: max of the model (from 0) and of the
 * direction's magnitude (the first cell's, then any larger one) */
static void maxima(const float *m, const float *d, size_t n, float *mmax, float *dmax)
{
    float a = 0.0f, b = 0.0f;
    for (size_t k = 0; k < n; k++) {
        if (m[k] > a) a = m[k];
        if (k == 0 || b == 0.0f) b = fabsf(d[k]);
        else if (fabsf(d[k]) > b) b = fabsf(d[k]);
    }
    *mmax = a;
    *dmax = b;
}

/* calc_mat_change_test_SH.c:10-189: vs and rho, the rho step without SCALERHO */
static float model_update_sh(const dnfwi_update_par *u, int iter, float eps_scale, size_t n, const float *vs,
                             const float *rho, const float *d_vs, const float *d_rho, float *vs_new, float *rho_new)
{
    float umax, gradmax_u, rhomax, gradmax_rho;
    maxima(vs, d_vs, n, &umax, &gradmax_u);
    maxima(rho, d_rho, n, &rhomax, &gradmax_rho);
    float eps_u = eps_scale * (umax / gradmax_u);
    if (iter < u->inv_vs_iter) eps_u = 0.0f;
    float eps_rho = eps_scale * (rhomax / gradmax_rho);
    if (iter < u->inv_rho_iter) eps_rho = 0.0f;
    if (u->invmat1 != 1 && u->invmat1 != 3) {
        memcpy(vs_new, vs, sizeof(float) * n);
        memcpy(rho_new, rho, sizeof(float) * n);
        return eps_u;
    }
    for (size_t k = 0; k < n; k++) {
        float u1 = vs[k] - eps_u * d_vs[k];
        float r1 = rho[k] - eps_rho * d_rho[k];
        if (u->invmat1 == 1) {
            if ((u1 < u->vslowerlim) && (u1 > 1e-6)) u1 = vs[k];
            if (u1 > u->vsupperlim) u1 = vs[k];
            if (r1 < u->rholowerlim) r1 = rho[k];
            if (r1 > u->rhoupperlim) r1 = rho[k];
        }
        if (u1 < 0.0) u1 = vs[k];
        if (r1 < 0.0) r1 = rho[k];
        vs_new[k] = u1;
        rho_new[k] = r1;
    }
    return eps_u; /* eps_true = EPSILON_u (:104) */
}

float dnfwi_model_update(const dnfwi_update_par *u, int iter, float eps_scale, int nx, int ny, const float *vp,
                         const float *vs, const float *rho, const float *d_vp, const float *d_vs, const float *d_rho,
                         float *vp_new, float *vs_new, float *rho_new)
{
    const size_t n = (size_t)nx * ny;
    if (!vp) return model_update_sh(u, iter, eps_scale, n, vs, rho, d_vs, d_rho, vs_new, rho_new);
    /* no vs (acoustic): calc_mat_change_test_AC.c, vp and rho only */
    const int ac = !vs;
    float pimax, gradmax, umax = 0.0f, gradmax_u = 0.0f, rhomax, gradmax_rho;
    maxima(vp, d_vp, n, &pimax, &gradmax);
    if (!ac) maxima(vs, d_vs, n, &umax, &gradmax_u);
    maxima(rho, d_rho, n, &rhomax, &gradmax_rho);
    /* calc_mat_change_test_PSV.c:111-147, calc_mat_change_test_AC.c:70-98 */
    float eps = eps_scale * (pimax / gradmax);
    if (iter < u->inv_vp_iter) eps = 0.0f;
    float eps_u = 0.0f;
    if (!ac) {
        eps_u = eps_scale * (umax / gradmax_u);
        if (iter < u->inv_vs_iter) eps_u = 0.0f;
    }
    float eps_rho = eps_scale * (rhomax / gradmax_rho) * u->scalerho;
    if (iter < u->inv_rho_iter) eps_rho = 0.0f;
    /* calc_mat_change_test_PSV.c:157-215, calc_mat_change_test_AC.c:111-160 */
    if (u->invmat1 == 1 || u->invmat1 == 3) {
        for (size_t k = 0; k < n; k++) {
            float p1 = vp[k] - eps * d_vp[k];
            float u1 = ac ? 0.0f : vs[k] - eps_u * d_vs[k];
            float r1 = rho[k] - eps_rho * d_rho[k];
            if (u->invmat1 == 1) {
                if (p1 < u->vplowerlim) p1 = vp[k];
                if (p1 > u->vpupperlim) p1 = vp[k];
                if (!ac && (u1 < u->vslowerlim) && (u1 > 1e-6)) u1 = vs[k];
                if (!ac && u1 > u->vsupperlim) u1 = vs[k];
                if (r1 < u->rholowerlim) r1 = rho[k];
                if (r1 > u->rhoupperlim) r1 = rho[k];
            }
            if (p1 < 0.0) p1 = vp[k];
            if (!ac && u1 < 0.0) u1 = vs[k];
            if (r1 < 0.0) r1 = rho[k];
            vp_new[k] = p1;
            if (!ac) vs_new[k] = u1;
            rho_new[k] = r1;
        }
    } else {
        /* INVMAT1 = 2: the reference leaves the model unchanged */
        memcpy(vp_new, vp, sizeof(float) * n);
        if (!ac) memcpy(vs_new, vs, sizeof(float) * n);
        memcpy(rho_new, rho, sizeof(float) * n);
    }
    /* the reference's eps_true: EPSILON_u (P/SV), EPSILON (acoustic,
     * calc_mat_change_test_AC.c:103) */
    return ac ? eps : eps_u;
}

float dnfwi_opt_step(const double *L2t, const float *epst)
{
    /* calc_opt_step.c:31-46: A[i] = (1, eps, eps^2), b = L2 (stored as float) */
    float A[3][3], b[3], x[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < 3; i++) {
        A[i][2] = (epst[i] * epst[i]);
        A[i][1] = (epst[i]);
        A[i][0] = (1.0);
        b[i] = (float)(L2t[i]);
    }
    /* solvelin.c:30-46 (method 1): forward elimination, no pivoting */
    for (int k = 0; k < 2; k++)
        for (int nn = k; nn < 2; nn++) {
            const float a = A[nn + 1][k] / A[k][k];
            for (int m = 0; m < 3; m++) A[nn + 1][m] = A[nn + 1][m] - a * A[k][m];
            b[nn + 1] = b[nn + 1] - a * b[k];
        }
    for (int k = 2; k >= 0; k--) {
        float c = b[k];
        for (int m = 2; m >= k + 1; m--) c = c - A[k][m] * x[m];
        x[k] = c / A[k][k];
    }
    /* calc_opt_step.c:49-72 */
    float opteps = -x[1] / (2.0 * x[2]);
    if (((2.0 * x[2]) < 0.0) && (L2t[1] > L2t[2])) opteps = epst[2];
    if (opteps > epst[2]) opteps = epst[2];
    if (opteps < 0.0) opteps = epst[0];
    return opteps;
}

int dnfwi_step_length(const dnfwi_update_par *u, int iter, float eps_scale0, float scalefac, int stepmax,
                      double L2t0, int nx, int ny, const float *vp, const float *vs, const float *rho,
                      const float *d_vp, const float *d_vs, const float *d_rho, dnfwi_misfit_fn misfit,
                      void *user, dnfwi_linesearch *out)
{
    if (!u || !misfit || !out || !rho || !d_rho || (!vp && !vs) || (vp && !d_vp) || (vs && !d_vs) || nx < 1 ||
        ny < 1)
        return dnio_fail_msg("dnfwi_step_length: null argument");
    const size_t n = (size_t)nx * ny;
    /* no vs: the acoustic search (step_length_est_ac.c), test models of vp
     * and rho; no vp: the SH search (step_length_est_sh.c), vs and rho */
    float *tv = vp ? malloc(sizeof(float) * n) : NULL, *ts = vs ? malloc(sizeof(float) * n) : NULL;
    float *tr = malloc(sizeof(float) * n);
    if ((vp && !tv) || (vs && !ts) || !tr) {
        free(tv);
        free(ts);
        free(tr);
        return dnio_fail_msg("dnfwi_step_length: out of host memory");
    }
    /* step_length_est_psv.c:37-49 */
    int step1 = 0, step2 = 0, step3 = 0, countstep = 0, itests = 2, iteste = 2, nevals = 0;
    float eps_scale = eps_scale0;
    /* epst1 and L2t are the caller's arrays in the reference (FWI_PSV.c):
     * slots the search does not reach keep the previous iteration's values */
    float epst1[4] = {0.0f, out->epst[0], out->epst[1], out->epst[2]};
    double L2t[4] = {0.0, L2t0, out->L2t[1], out->L2t[2]};
    while ((step2 != 1) || (step1 != 1)) {
        for (int itest = itests; itest <= iteste; itest++) {
            /* step_length_est_psv.c:55-63: test model at eps_scale, its misfit */
            dnfwi_model_update(u, iter, eps_scale, nx, ny, vp, vs, rho, d_vp, d_vs, d_rho, tv, ts, tr);
            L2t[itest] = misfit(user, tv, ts, tr, itest);
            nevals++;
            epst1[itest] = eps_scale;
            epst1[1] = 0.0f;
        }
        /* step_length_est_psv.c:117-149 */
        if ((step1 == 0) && (L2t[1] <= L2t[2])) {
            eps_scale = eps_scale / scalefac;
            countstep++;
        }
        if ((step1 == 1) && (L2t[2] < L2t[3])) {
            epst1[3] = eps_scale;
            step2 = 1;
        }
        if ((step1 == 1) && (L2t[2] >= L2t[3])) {
            epst1[3] = eps_scale;
            eps_scale = eps_scale + (eps_scale / scalefac);
            countstep++;
        }
        if ((step1 == 0) && (L2t[1] > L2t[2])) {
            epst1[2] = eps_scale;
            step1 = 1;
            iteste = 3;
            itests = 3;
            countstep = 0;
            eps_scale = eps_scale + (eps_scale / scalefac);
        }
        /* step_length_est_psv.c:151-165 */
        step3 = 0;
        if ((step1 == 0) && (countstep > stepmax)) {
            step3 = 1;
            break;
        }
        if ((step1 == 1) && (countstep > stepmax)) {
            step1 = 1;
            step2 = 1;
        }
    }
    /* step_length_est_psv.c:170-179 */
    if (step1 == 1) eps_scale = dnfwi_opt_step(L2t + 1, epst1 + 1);
    free(tv);
    free(ts);
    free(tr);
    out->eps_scale = eps_scale;
    for (int k = 0; k < 3; k++) {
        out->epst[k] = epst1[k + 1];
        out->L2t[k] = L2t[k + 1];
    }
    out->step1 = step1;
    out->step3 = step3;
    out->nevals = nevals;
    return 0;
}

int dnfwi_rtm_adjoint_source(const float *data, const float *syn, int ntr, int ns, float dt, int grad_form,
                             int testshot, float *diff, double *l2)
{
    if (!data || !syn || !diff || ntr < 0 || ns < 1 || (grad_form != 1 && grad_form != 2))
        return dnio_fail_msg("dnfwi_rtm_adjoint_source: bad argument");
    double L2 = 0.0;
    for (int i = 0; i < ntr; i++) {
        const float *d = data + (size_t)i * ns, *s = syn + (size_t)i * ns;
        float *out = diff + (size_t)i * ns;
        float intseis = 0.0f;
        /* calc_res.c:416-446: invtime runs ns..1 while j runs 1..ns */
        for (int j = 0; j < ns; j++) {
            const float dj = j == 0 ? s[0] : d[j];
            if (grad_form == 1) intseis += dt * dj;
            else intseis = dj;
            out[ns - 1 - j] = intseis;
        }
        /* calc_res.c:563-569 */
        if (testshot)
            for (int j = 0; j < ns; j++) L2 += 0.5 * out[ns - 1 - j] * out[ns - 1 - j];
    }
    if (l2) *l2 = L2;
    return 0;
}

int dnfwi_energy_precond(int eprecond, int nx, int ny, float dh, const float *Ws, const float *Wr, float c_vp,
                         float c_vs, float c_rho, float *g_vp, float *g_vs, float *g_rho, float *We)
{
    if ((eprecond != 1 && eprecond != 3) || !Ws || (eprecond == 1 && !Wr) || (!g_vp && !g_vs) || !g_rho || nx < 1 ||
        ny < 1)
        return dnio_fail_msg("dnfwi_energy_precond: bad argument");
    const size_t n = (size_t)nx * ny;
    float *w = We ? We : malloc(sizeof(float) * n);
    if (!w) return dnio_fail_msg("dnfwi_energy_precond: out of host memory");
    /* eprecond1.c:13-46 */
    const float xmin = dh, xmax = nx * dh, eps_we = 0.005f;
    float maxwe = 0.0f;
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++) {
            const size_t k = (size_t)(i - 1) * ny + (j - 1);
            if (eprecond == 1) {
                w[k] = sqrt(Ws[k] * Wr[k]);
            } else {
                const float x = i * dh, y = j * dh;
                w[k] = sqrt(Ws[k]) * (asinh((xmax - x) / y) - asinh((xmin - x) / y));
            }
            if (w[k] > maxwe) maxwe = w[k];
        }
    /* eprecond1.c:50-58 */
    for (size_t k = 0; k < n; k++) w[k] = w[k] + (eps_we * maxwe);
    /* grad_obj_psv.c:224-237; grad_obj_ac.c:225-233 has no vs gradient */
    for (size_t k = 0; k < n; k++) {
        if (g_vp) g_vp[k] = g_vp[k] / (w[k] * c_vp * c_vp);
        if (g_vs) {
            g_vs[k] = g_vs[k] / (w[k] * c_vs * c_vs);
            if (c_vs == 0.0) g_vs[k] = 0.0f;
        }
        g_rho[k] = g_rho[k] / (w[k] * c_rho * c_rho);
    }
    if (!We) free(w);
    return 0;
}

int dnfwi_taper_hor(int nx, int ny, float dh, int gradt1, int gradt2, float exp_taper, float *coeff, float *g)
{
    if (!coeff || nx < 1 || ny < 1) return dnio_fail_msg("dnfwi_taper_hor: bad argument");
    const int taperlength = gradt2 - gradt1, ifw = gradt2 - gradt1 + 1;
    float *window = malloc(sizeof(float) * (size_t)(ifw > 0 ? ifw : 1));
    if (!window) return dnio_fail_msg("dnfwi_taper_hor: out of host memory");
    const float a = 3; /* damping coefficient */
    for (int i = 1; i <= ifw; i++) {
        window[i - 1] = 1.0f;
        if (i <= taperlength)
            window[i - 1] = exp(-(1.0 / 2.0) * (a * (i - taperlength) / (taperlength / 2.0)) *
                                (a * (i - taperlength) / (taperlength / 2.0)));
    }
    int h = 1;
    for (int j = 1; j <= ny; j++) {
        float grad_tap = 0.0f;
        if ((j >= gradt1) && (j <= gradt2)) grad_tap = window[h - 1];
        if (j > gradt2) grad_tap = pow((float)(j * dh), exp_taper);
        for (int i = 1; i <= nx; i++) coeff[(size_t)(i - 1) * ny + (j - 1)] = grad_tap;
        if (j >= gradt1) h++;
    }
    free(window);
    if (g)
        for (size_t k = 0; k < (size_t)nx * ny; k++) g[k] *= coeff[k];
    return 0;
}

int dnfwi_residual_gc(const float *data, const float *syn, int ntr, int ns, float dt, int grad_form, int testshot,
                      const int *skip, float *diff, double *l2)
{
    if (!data || !syn || !diff || !l2 || ntr < 0 || ns < 1 || (grad_form != 1 && grad_form != 2))
        return dnio_fail_msg("dnfwi_residual_gc: bad argument");
    const float eps = 0.0f; /* EPS_LNORM, calc_res.c:36 */
    float *id = NULL, *is = NULL;
    if (grad_form == 1) {
        id = malloc(sizeof(float) * (size_t)ns);
        is = malloc(sizeof(float) * (size_t)ns);
        if (!id || !is) {
            free(id);
            free(is);
            return dnio_fail_msg("dnfwi_residual_gc: out of host memory");
        }
    }
    memset(diff, 0, sizeof(float) * (size_t)ntr * ns); /* calc_res.c:41-43 */
    double L2 = *l2;
    for (int i = 0; i < ntr; i++) {
        if (skip && skip[i] == 1) continue;
        const float *dt_ = data + (size_t)i * ns, *st_ = syn + (size_t)i * ns;
        const float *d = dt_, *sy = st_;
        if (grad_form == 1) {
            float a = 0.0f, b = 0.0f;
            for (int j = 0; j < ns; j++) {
                a += st_[j];
                is[j] = dt * a;
                b += dt_[j];
                id[j] = dt * b;
            }
            d = id;
            sy = is;
        }
        float abs_d = 0.0f, abs_s = 0.0f, dms = 0.0f;
        for (int j = 0; j < ns; j++) {
            abs_d += d[j] * d[j];
            abs_s += sy[j] * sy[j];
            dms += sy[j] * d[j];
        }
        abs_d = sqrt(abs_d);
        abs_s = sqrt(abs_s);
        float *out = diff + (size_t)i * ns;
        for (int j = 0; j < ns; j++) {
            const float dj = (grad_form == 2 && j == 0) ? st_[0] : d[j];
            const float sj = sy[j];
            const float tmp = abs_s * abs_s * abs_s * abs_d, tmp1 = abs_s * abs_d;
            out[ns - 1 - j] = ((sj * dms) / (tmp + eps)) - (dj / (tmp1 + eps));
            if (testshot) {
                const float t2 = abs_d * abs_s;
                L2 -= (dj * sj) / (t2 + eps);
            }
        }
    }
    *l2 = L2;
    free(id);
    free(is);
    return 0;
}

int dnfwi_time_window(int timewin, int ntr, int ns, float dt, float gamma, float twlength_minus,
                      float twlength_plus, const float *picks, const float *offset, float *diff)
{
    if ((ntr > 0 && !diff) || ntr < 0 || ns < 0) return dnio_fail_msg("dnfwi_time_window: bad argument");
    if ((timewin == 1 || timewin == 4) && !picks) return dnio_fail_msg("dnfwi_time_window: TIMEWIN 1/4 need picks");
    if (timewin == 5 && !offset) return dnio_fail_msg("dnfwi_time_window: TIMEWIN 5 needs offsets");
    for (int i = 0; i < ntr; i++) {
        float *d = diff + (size_t)i * ns;
        if (timewin == 5) {
            const float t0 = fabs(offset[i] / twlength_minus);
            for (int k = 0; k < ns; k++) {
                const float t = (ns - k) * dt;
                if (t >= t0) d[k] *= exp(-gamma * pow(t - t0, 2));
            }
        } else if (timewin == 1 || timewin == 2 || timewin == 4) {
            const float pt = timewin == 2 ? -dt : picks[i] + twlength_plus;
            const float pt1 = timewin == 2 ? twlength_plus : picks[i];
            for (int k = 0; k < ns; k++) {
                const float t = (float)((ns - k) * dt);
                const float dump = t - pt;
                const float taper = exp(-gamma * dump * dump);
                if (timewin == 2 ? t >= pt1 : t >= pt) d[k] = d[k] * taper;
            }
        }
    }
    return 0;
}

int dnfwi_gauss_filter(int nx, int ny, float dh, float wd_damp, float c_vs, float fc_end, float *g, int *width)
{
    if (!g || nx < 1 || ny < 1 || !(dh > 0.0f) || !(fc_end > 0.0f))
        return dnio_fail_msg("dnfwi_gauss_filter: bad argument (FC_END and DH must be positive)");
    const float lam = c_vs / fc_end;
    int fs = (int)round((wd_damp * lam) / dh);
    if (width) *width = fs;
    if (fs == 0) return 0;
    if (fs % 2 == 0) fs += fs > 0 ? 1 : -1;
    if (width) *width = fs;
    const int hfs = abs(fs) / 2, n = 2 * hfs + 1;
    /* |width| 1 or 3 gives sigma = 0 and a 0/0 kernel (NaN everywhere, as
     * in the reference): refused */
    if (hfs < 2) {
        char msg[200];
        snprintf(msg, sizeof msg,
                 "dnfwi_gauss_filter: SPATFILTER 3 filter width %d (|width| below 5) gives a zero Gaussian "
                 "sigma and a 0/0 kernel; raise WD_DAMP",
                 fs);
        return dnio_fail_msg(msg);
    }
    const float sigma = hfs / 2; /* integer division, as the reference */
    const float s = 2.0 * sigma * sigma;
    const size_t ncell = (size_t)nx * ny;
    float *k = malloc(sizeof(float) * (size_t)n * n), *src = malloc(sizeof(float) * ncell);
    if (!k || !src) {
        free(k);
        free(src);
        return dnio_fail_msg("dnfwi_gauss_filter: out of host memory");
    }
    /* k[(di + hfs) * n + (dj + hfs)]: di along x, dj along y */
    float sum = 0.0f;
    for (int di = -hfs; di <= hfs; di++)
        for (int dj = -hfs; dj <= hfs; dj++) {
            const float e = -((di * di) / s) - ((dj * dj) / s);
            k[(di + hfs) * n + dj + hfs] = exp(e);
            sum += k[(di + hfs) * n + dj + hfs];
        }
    /* the reference normalises over 1..width, which is nothing for a
     * negative width */
    if (fs > 0)
        for (int q = 0; q < n * n; q++) k[q] /= sum;
    memcpy(src, g, sizeof(float) * ncell);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            float conv = 0.0f;
            for (int di = -hfs; di <= hfs; di++) {
                const int ii = i + di < 0 ? 0 : (i + di >= nx ? nx - 1 : i + di);
                for (int dj = -hfs; dj <= hfs; dj++) {
                    const int jj = j + dj < 0 ? 0 : (j + dj >= ny ? ny - 1 : j + dj);
                    conv += src[(size_t)ii * ny + jj] * k[(di + hfs) * n + dj + hfs];
                }
            }
            g[(size_t)i * ny + j] = conv;
        }
    free(k);
    free(src);
    return 0;
}

int dnfwi_gauss_filter_var(int nx, int ny, float dh, float wd_damp, float wd_damp1, float fc_end, const float *vel,
                           float *g)
{
    if (!g || !vel || nx < 1 || ny < 1 || !(dh > 0.0f) || !(fc_end > 0.0f))
        return dnio_fail_msg("dnfwi_gauss_filter_var: bad argument (FC_END and DH must be positive)");
    const size_t ncell = (size_t)nx * ny;
    float *src = malloc(sizeof(float) * ncell);
    if (!src) return dnio_fail_msg("dnfwi_gauss_filter_var: out of host memory");
    memcpy(src, g, sizeof(float) * ncell);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const float lam = vel[(size_t)i * ny + j] / fc_end;
            int wx = (int)round((wd_damp * lam) / dh), wz = (int)round((wd_damp1 * lam) / dh);
            if (wz <= 0) continue; /* g keeps the cell's value */
            if (wx % 2 == 0) wx += 1;
            if (wz % 2 == 0) wz += 1;
            if (wx < 6) wx = 6;
            if (wz < 6) wz = 6;
            const int hx = abs(wx) / 2, hz = abs(wz) / 2;
            const float sigmax = hx / 2, sigmaz = hz / 2; /* integer division, as the reference */
            const float sx = 2.0 * sigmax * sigmax, sz = 2.0 * sigmaz * sigmaz;
            float conv = 0.0f, sum = 0.0f;
            for (int di = -hx; di <= hx; di++) {
                const int ii = i + di < 0 ? 0 : (i + di >= nx ? nx - 1 : i + di);
                for (int dj = -hz; dj <= hz; dj++) {
                    const int jj = j + dj < 0 ? 0 : (j + dj >= ny ? ny - 1 : j + dj);
                    const float k = exp(-((di * di) / sx) - ((dj * dj) / sz));
                    conv += src[(size_t)ii * ny + jj] * k;
                    sum += k;
                }
            }
            g[(size_t)i * ny + j] = conv / sum;
        }
    free(src);
    return 0;
}

int dnfwi_median_filter(int nx, int ny, int filt_size, float *a)
{
    if (!a || nx < 1 || ny < 1) return dnio_fail_msg("dnfwi_median_filter: bad argument");
    if (filt_size == 0) return 0;
    int fs = abs(filt_size);
    if (fs % 2 == 0) fs++;
    const int hfs = fs / 2, n = fs * fs;
    const size_t ncell = (size_t)nx * ny;
    float *src = malloc(sizeof(float) * ncell), *win = malloc(sizeof(float) * (size_t)n);
    if (!src || !win) {
        free(src);
        free(win);
        return dnio_fail_msg("dnfwi_median_filter: out of host memory");
    }
    memcpy(src, a, sizeof(float) * ncell);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            /* the window row by row (y offset outer, x offset inner), each
             * value inserted into a descending run; with an odd count the
             * middle element is the median whatever the order of ties */
            int k = 0;
            for (int dj = -hfs; dj <= hfs; dj++) {
                const int jj = j + dj < 0 ? 0 : (j + dj >= ny ? ny - 1 : j + dj);
                for (int di = -hfs; di <= hfs; di++) {
                    const int ii = i + di < 0 ? 0 : (i + di >= nx ? nx - 1 : i + di);
                    const float v = src[(size_t)ii * ny + jj];
                    int m = k++;
                    while (m > 0 && win[m - 1] < v) {
                        win[m] = win[m - 1];
                        m--;
                    }
                    win[m] = v;
                }
            }
            a[(size_t)i * ny + j] = win[n / 2];
        }
    free(src);
    free(win);
    return 0;
}

struct dnfwi_pcg {
    size_t ncell;
    float *old, *dir, *nw; /* PCG_old, PCG_dir, PCG_new: [vp | vs | rho] */
};

int dnfwi_pcg_create(int nx, int ny, dnfwi_pcg **out)
{
    if (!out || nx < 1 || ny < 1) return dnio_fail_msg("dnfwi_pcg_create: bad size");
    dnfwi_pcg *pc = calloc(1, sizeof *pc);
    if (!pc) return dnio_fail_msg("dnfwi_pcg_create: out of host memory");
    pc->ncell = (size_t)nx * ny;
    pc->old = calloc(3 * pc->ncell, sizeof(float));
    pc->dir = calloc(3 * pc->ncell, sizeof(float));
    pc->nw = calloc(3 * pc->ncell, sizeof(float));
    if (!pc->old || !pc->dir || !pc->nw) {
        dnfwi_pcg_destroy(pc);
        return dnio_fail_msg("dnfwi_pcg_create: out of host memory");
    }
    *out = pc;
    return 0;
}

void dnfwi_pcg_destroy(dnfwi_pcg *pc)
{
    if (!pc) return;
    free(pc->old);
    free(pc->dir);
    free(pc->nw);
    free(pc);
}

void dnfwi_pcg_reset(dnfwi_pcg *pc)
{
    if (!pc) return;
    memset(pc->old, 0, sizeof(float) * 3 * pc->ncell);
    memset(pc->dir, 0, sizeof(float) * 3 * pc->ncell);
    memset(pc->nw, 0, sizeof(float) * 3 * pc->ncell);
}

int dnfwi_pcg_direction(dnfwi_pcg *pc, int pcg_beta, float *g_vp, float *g_vs, float *g_rho, float *beta_out)
{
    if (!pc || (!g_vp && !g_vs) || !g_rho || pcg_beta < 1 || pcg_beta > 4)
        return dnio_fail_msg("dnfwi_pcg_direction: bad argument (PCG_BETA 1..4)");
    const size_t nc = pc->ncell;
    float *g[3] = {g_vp, g_vs, g_rho};
    /* descent + store_PCG_PSV: PCG_new = -g.  Without g_vs (the acoustic
     * store_PCG_AC.c, classes vp and rho) the vs class stays 0: every class
     * has its own beta (PCG.c:21-78), so the others are those of a two-class
     * run. */
    for (int c = 0; c < 3; c++)
        for (size_t k = 0; k < nc; k++) pc->nw[c * nc + k] = g[c] ? -g[c][k] : 0.0f;
    /* PCG.c:21-78 */
    float beta[3];
    for (int c = 0; c < 3; c++) {
        const float *n = pc->nw + c * nc, *o = pc->old + c * nc, *d = pc->dir + c * nc;
        float betaz = 0.0f, betan = 0.0f;
        for (size_t k = 0; k < nc; k++) {
            if (pcg_beta == 1) {
                betaz += n[k] * n[k];
                betan += o[k] * o[k];
            } else if (pcg_beta == 2) {
                betaz += n[k] * (n[k] - o[k]);
                betan += o[k] * o[k];
            } else if (pcg_beta == 3) {
                betaz += -n[k] * (n[k] - o[k]);
                betan += d[k] * (n[k] - o[k]);
            } else {
                betaz += -n[k] * n[k];
                betan += d[k] * (n[k] - o[k]);
            }
        }
        beta[c] = 0.0f;
        if (betan != 0.0f) beta[c] = betaz / betan;
        if ((beta[c] < 0.0) && (pcg_beta == 2)) beta[c] = 0.0f;
    }
    /* PCG.c:82-93, extract_PCG_PSV, store_PCG_PSV(PCG_old), descent */
    for (int c = 0; c < 3; c++) {
        float *d = pc->dir + c * nc;
        const float *n = pc->nw + c * nc;
        for (size_t k = 0; k < nc; k++) {
            d[k] = n[k] + beta[c] * d[k];
            if (g[c]) g[c][k] = -d[k];
        }
        memcpy(pc->old + c * nc, n, sizeof(float) * nc);
        if (beta_out) beta_out[c] = beta[c];
    }
    return 0;
}

int dnfwi_taper_vert(int nx, int ny, float dh, const int *gradt, int nsrc, const float *srcpos, int ntr,
                     const int *recpos, float *coeff, float *g)
{
    if (!coeff || !gradt || nx < 1 || ny < 1 || (nsrc > 0 && !srcpos) || (ntr > 0 && !recpos))
        return dnio_fail_msg("dnfwi_taper_vert: bad argument");
    const int tl = gradt[1] - gradt[0], tl2 = gradt[3] - gradt[2], ifw = gradt[3] - gradt[0];
    float *window = malloc(sizeof(float) * (size_t)(ifw > 0 ? ifw : 1));
    if (!window) return dnio_fail_msg("dnfwi_taper_vert: out of host memory");
    const float a = 3; /* damping coefficient */
    for (int i = 1; i <= ifw; i++) {
        window[i - 1] = 1.0f;
        if (i <= tl)
            window[i - 1] = exp(-(1.0 / 2.0) * (a * (i - tl) / (tl / 2.0)) * (a * (i - tl) / (tl / 2.0)));
        if (i >= (ifw - tl2))
            window[i - 1] = exp(-(1.0 / 2.0) * (a * (i - (ifw - tl2)) / (tl2 / 2.0)) *
                                (a * (i - (ifw - tl2)) / (tl2 / 2.0)));
    }
    for (int j = 1; j <= ny; j++) {
        int h = 1;
        for (int i = 1; i <= nx; i++) {
            float grad_tap = 0.0f;
            if ((i > gradt[0]) && (i < gradt[3])) grad_tap = window[h++ - 1];
            coeff[(size_t)(i - 1) * ny + (j - 1)] = grad_tap;
        }
    }
    free(window);
    if (!g) return 0;
    for (size_t k = 0; k < (size_t)nx * ny; k++) g[k] *= coeff[k];
    /* taper_grad.c:120-170: the source and receiver cells */
    for (int n = 0; n < nsrc; n++) {
        const int i = (int)floor(srcpos[0 * nsrc + n] / dh + 0.5), j = (int)floor(srcpos[1 * nsrc + n] / dh + 0.5);
        if (i >= 1 && i <= nx && j >= 1 && j <= ny) g[(size_t)(i - 1) * ny + (j - 1)] = 0.0f;
    }
    for (int n = 0; n < ntr; n++) {
        const int i = recpos[n], j = recpos[ntr + n];
        if (i >= 1 && i <= nx && j >= 1 && j <= ny) g[(size_t)(i - 1) * ny + (j - 1)] = 0.0f;
    }
    return 0;
}

int dnfwi_taper_sources(int nx, int ny, float dh, float srtradius, int srtshape, int filtsize, int nsrc,
                        const float *srcpos, float *coeff, float *g)
{
    if (!coeff || nx < 1 || ny < 1 || (nsrc > 0 && !srcpos) || (srtshape != 1 && srtshape != 2))
        return dnio_fail_msg("dnfwi_taper_sources: bad argument (SRTSHAPE 1 or 2)");
    int gp = (int)(ceil(2.0 * srtradius / dh));
    if (gp < 5) gp = 5;
    const size_t n = (size_t)nx * ny;
    float *m = malloc(sizeof(float) * (size_t)gp * gp), *msum = malloc(sizeof(float) * n);
    if (!m || !msum) {
        free(m);
        free(msum);
        return dnio_fail_msg("dnfwi_taper_sources: out of host memory");
    }
#define M(j, i) m[(size_t)((j)-1) * gp + ((i)-1)]
    /* taper_grad.c:346-371: the stencil m[j][i] */
    const float a = 1.0;
    const float maxrad = sqrt(2.0 * srtradius * srtradius);
    for (int j = 1; j <= gp; j++)
        for (int i = 1; i <= gp; i++) {
            const float x = ((float)i - ((float)gp) / 2.0 - 0.5) * dh;
            const float y = ((float)j - ((float)gp) / 2.0 - 0.5) * dh;
            const float rad = sqrt(x * x + y * y);
            if (srtshape == 1)
                M(j, i) = erf(a * rad / maxrad);
            else
                M(j, i) = rad > 0 ? log(rad) : 0.0;
        }
    /* minimum_m / maximum_m (util.c:38-70): over all values */
    float minm = M(1, 1);
    for (int i = 1; i <= gp; i++)
        for (int j = 1; j <= gp; j++)
            if (M(j, i) < minm) minm = M(j, i);
    for (int j = 1; j <= gp; j++)
        for (int i = 1; i <= gp; i++) M(j, i) -= minm;
    const float edge[4] = {M(1, gp / 2), M(gp / 2, 1), M(gp / 2, gp), M(gp, gp / 2)};
    float maxm = edge[0];
    for (int k = 1; k < 4; k++)
        if (edge[k] > maxm) maxm = edge[k];
    for (int j = 1; j <= gp; j++)
        for (int i = 1; i <= gp; i++) {
            M(j, i) /= maxm;
            if (M(j, i) > 1.0) M(j, i) = 1.0;
        }
    /* taper_grad.c:387-419: the minimum over the sources' placements */
    const int ijc = (int)(ceil((float)gp / 2));
    for (size_t k = 0; k < n; k++) msum[k] = 1.0f;
    for (int s = 0; s < nsrc; s++) {
        const int i = (int)floor(srcpos[0 * nsrc + s] / dh + 0.5), j = (int)floor(srcpos[1 * nsrc + s] / dh + 0.5);
        for (int iy = 1; iy <= gp; iy++)
            for (int ix = 1; ix <= gp; ix++) {
                const int xx = i + ix - ijc, yy = j + iy - ijc;
                if ((xx < 1) || (xx > nx) || (yy < 1) || (yy > ny)) continue;
                float *v = msum + (size_t)(xx - 1) * ny + (yy - 1);
                if (*v > M(iy, ix)) *v = M(iy, ix);
            }
    }
#undef M
    /* taper_grad.c:456-479 */
    float mn = msum[0];
    for (size_t k = 0; k < n; k++)
        if (msum[k] < mn) mn = msum[k];
    for (size_t k = 0; k < n; k++) msum[k] -= mn;
    float mx = msum[0];
    for (size_t k = 0; k < n; k++)
        if (msum[k] > mx) mx = msum[k];
    for (size_t k = 0; k < n; k++) coeff[k] = msum[k] / mx;
    if (g)
        for (size_t k = 0; k < n; k++) g[k] *= coeff[k];
    /* taper_grad.c:488-512: the boxes around the sources */
    for (int s = 0; s < nsrc; s++) {
        const int i1 = (int)floor(srcpos[0 * nsrc + s] / dh + 0.5), j1 = (int)floor(srcpos[1 * nsrc + s] / dh + 0.5);
        for (int i = i1 - filtsize; i <= i1 + filtsize; i++)
            for (int j = j1 - filtsize; j <= j1 + filtsize; j++) {
                if (i < 1 || i > nx || j < 1 || j > ny) continue;
                const size_t k = (size_t)(i - 1) * ny + (j - 1);
                if (g) g[k] = 0.0f;
                coeff[k] = 0.0f;
            }
    }
    free(m);
    free(msum);
    return 0;
}

int dnfwi_assemble_sh(int nx, int ny, float dt, int invmat1, int grad_form, int iter, int inv_vs_iter,
                      int inv_rho_iter, const float *vs, const float *rho, float *g_u, float *g_rho)
{
    if (!vs || !rho || !g_u || !g_rho || nx < 1 || ny < 1 || (grad_form != 1 && grad_form != 2))
        return dnio_fail_msg("dnfwi_assemble_sh: bad argument");
    const size_t n = (size_t)nx * ny;
    float *wmu = malloc(sizeof(float) * n);
    if (!wmu) return dnio_fail_msg("dnfwi_assemble_sh: out of host memory");
    /* ass_gradSH.c:19-61: mu' = -DT w_u, then the vs (INVMAT1 1), Zs (2) or
     * mu (3) gradient; where the INVMAT1 1 / GRAD_FORM 1 branch writes
     * nothing (vs = 0) the sum stays */
    for (size_t k = 0; k < n; k++) {
        wmu[k] = -dt * g_u[k];
        if (invmat1 == 1) {
            if (grad_form == 1) {
                if (vs[k] > 0.0) {
                    const float vs3 = vs[k] * vs[k] * vs[k];
                    g_u[k] = (2.0 / rho[k] * vs3) * wmu[k];
                }
            } else {
                g_u[k] = 2.0 * rho[k] * vs[k] * wmu[k];
            }
        }
        if (invmat1 == 2) g_u[k] = 2.0 * vs[k] * wmu[k];
        if (invmat1 == 3) g_u[k] = wmu[k];
        if (iter < inv_vs_iter) g_u[k] = 0.0f;
    }
    /* ass_gradSH.c:63-96: rho' = -DT w_rho plus the mu term */
    for (size_t k = 0; k < n; k++) {
        const float rho_s = -dt * g_rho[k];
        if (invmat1 == 1) {
            if (grad_form == 1) {
                if (vs[k] > 0.0) {
                    const float vs2rho2 = 1.0 / (vs[k] * vs[k] * rho[k] * rho[k]);
                    g_rho[k] = (vs2rho2 * wmu[k]) + rho_s;
                }
            } else {
                g_rho[k] = (vs[k] * vs[k] * wmu[k]) + rho_s;
            }
        }
        if (invmat1 == 3) g_rho[k] = rho_s;
        if (iter < inv_rho_iter) g_rho[k] = 0.0f;
    }
    free(wmu);
    return 0;
}
