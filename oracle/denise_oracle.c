/*
 * CPU oracle for the DENISE Black Edition 2D P/SV time-stepping path.
 * TEST INFRASTRUCTURE ONLY -- the checker, never the product (see
 * denise_oracle.h).
 *
 * Compiled with -DORC_REAL=float  (reference FP32 path: float storage, the
 * reference's double-literal promotions kept) and -DORC_REAL=double (FP64
 * oracle).  -O2, no FMA contraction, no fast-math: the operation order of
 * every expression is the reference's.
 *
 * Each routine cites the reference source (paths relative to the reference
 * repository root) whose arithmetic it restates.
 */
#include "denise_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL real;

#define PI_D 3.141592653589793 /* include/fd.h:22 */

static char g_err[512];
static void seterr(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
const char *orc_error(void) { return g_err; }

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_real_size(void) { return (int)sizeof(real); }

/* iround, include/fd.h:17 */
static int iround_d(double x) { return (int)floor(x + 0.5); }

/* ------------------------------------------------------------------------- */
/* padded 2D arrays with reference (NR) indexing                              */
/* ------------------------------------------------------------------------- */
typedef struct {
    real *p;
    int jlo, jhi, ilo, ihi, ncol;
} arr;

#define AT(a, j, i) ((a).p[(size_t)((j) - (a).jlo) * (size_t)(a).ncol + (size_t)((i) - (a).ilo)])

static arr arr_new(int jlo, int jhi, int ilo, int ihi)
{
    arr a;
    a.jlo = jlo; a.jhi = jhi; a.ilo = ilo; a.ihi = ihi;
    a.ncol = ihi - ilo + 1;
    a.p = (real *)calloc((size_t)(jhi - jlo + 1) * (size_t)a.ncol, sizeof(real));
    return a;
}
static void arr_zero(arr a)
{
    if (a.p) memset(a.p, 0, (size_t)(a.jhi - a.jlo + 1) * (size_t)a.ncol * sizeof(real));
}
static size_t arr_len(arr a) { return (size_t)(a.jhi - a.jlo + 1) * (size_t)a.ncol; }

/* ------------------------------------------------------------------------- */
/* Holberg / Taylor staggered-grid coefficients, src/holbergcoeff.c:29-65     */
/* rows: MAXRELERROR 0..4; columns: FDORDER/2-1; entries: hc[0..6], hc[0] is  */
/* the minimum number of grid points per shortest wavelength.                */
/* ------------------------------------------------------------------------- */
static const float HOLBERG[5][6][7] = {
    {{23.0f, 1.0f, 0, 0, 0, 0, 0},
     {8.0f, (float)(9.0 / 8.0), (float)(-1.0 / 24.0), 0, 0, 0, 0},
     {6.0f, (float)(75.0 / 64.0), (float)(-25.0 / 384.0), (float)(3.0 / 640.0), 0, 0, 0},
     {5.0f, (float)(1225.0 / 1024.0), (float)(-245.0 / 3072.0), (float)(49.0 / 5120.0),
      (float)(-5.0 / 7168.0), 0, 0},
     {5.0f, (float)(19845.0 / 16384.0), (float)(-735.0 / 8192.0), (float)(567.0 / 40960.0),
      (float)(-405.0 / 229376.0), (float)(35.0 / 294912.0), 0},
     {4.0f, (float)(160083.0 / 131072.0), (float)(-12705.0 / 131072.0),
      (float)(22869.0 / 1310720.0), (float)(-5445.0 / 1835008.0), (float)(847.0 / 2359296.0),
      (float)(-63.0 / 2883584.0)}},
    {{49.7f, 1.0010f, 0, 0, 0, 0, 0},
     {8.32f, 1.1382f, -0.046414f, 0, 0, 0, 0},
     {4.77f, 1.1965f, -0.078804f, 0.0081781f, 0, 0, 0},
     {3.69f, 1.2257f, -0.099537f, 0.018063f, -0.0026274f, 0, 0},
     {3.19f, 1.2415f, -0.11231f, 0.026191f, -0.0064682f, 0.001191f, 0},
     {2.91f, 1.2508f, -0.12034f, 0.032131f, -0.010142f, 0.0029857f, -0.00066667f}},
    {{22.2f, 1.0050f, 0, 0, 0, 0, 0},
     {5.65f, 1.1534f, -0.052806f, 0, 0, 0, 0},
     {3.74f, 1.2111f, -0.088313f, 0.011768f, 0, 0, 0},
     {3.11f, 1.2367f, -0.10815f, 0.023113f, -0.0046905f, 0, 0},
     {2.80f, 1.2496f, -0.11921f, 0.031130f, -0.0093272f, 0.0025161f, 0},
     {2.62f, 1.2568f, -0.12573f, 0.036423f, -0.013132f, 0.0047484f, -0.0015979f}},
    {{15.8f, 1.0100f, 0, 0, 0, 0, 0},
     {4.80f, 1.1640f, -0.057991f, 0, 0, 0, 0},
     {3.39f, 1.2192f, -0.094070f, 0.014608f, 0, 0, 0},
     {2.90f, 1.2422f, -0.11269f, 0.026140f, -0.0064054f, 0, 0},
     {2.65f, 1.2534f, -0.12257f, 0.033755f, -0.011081f, 0.0036784f, 0},
     {2.51f, 1.2596f, -0.12825f, 0.038550f, -0.014763f, 0.0058619f, -0.0024538f}},
    {{9.16f, 1.0300f, 0, 0, 0, 0, 0},
     {3.47f, 1.1876f, -0.072518f, 0, 0, 0, 0},
     {2.91f, 1.2341f, -0.10569f, 0.022589f, 0, 0, 0},
     {2.61f, 1.2516f, -0.12085f, 0.032236f, -0.011459f, 0, 0},
     {2.45f, 1.2596f, -0.12829f, 0.038533f, -0.014681f, 0.0072580f, 0},
     {2.36f, 1.2640f, -0.13239f, 0.042217f, -0.017803f, 0.0081959f, -0.0051848f}},
};

/* src/holbergcoeff.c:21-90 */
int orc_holberg(int fdorder, int maxrelerror, float *hc7)
{
    if (fdorder < 2 || fdorder > 12 || (fdorder & 1)) {
        seterr("wrong FDORDER %d", fdorder);
        return -1;
    }
    if (maxrelerror < 0 || maxrelerror > 4) {
        seterr("wrong MAXRELERROR %d", maxrelerror);
        return -1;
    }
    for (int i = 0; i <= 6; i++) hc7[i] = HOLBERG[maxrelerror][fdorder / 2 - 1][i];
    return 0;
}

/* Ricker branch of src/wavelet.c:38-51 (QUELLART=1), times amplitude (:96).
 * out[0..nt-1] holds signals[k][1..NT]. */
void orc_ricker(int nt, float dt, float tshift, float fc, float amp, float *out)
{
    for (int n = 1; n <= nt; n++) {
        float t = (float)n * dt;
        float ts = (float)(1.0 / fc);
        float tau = (float)(PI_D * (t - 1.5 * ts - tshift) / (1.5 * ts));
        float a = (float)((1.0 - 4.0 * tau * tau) * exp(-2.0 * tau * tau));
        out[n - 1] = a * amp;
    }
}

/* ------------------------------------------------------------------------- */
/* context                                                                    */
/* ------------------------------------------------------------------------- */
enum { PML_DX = 0, PML_KX, PML_AX_ALPHA, PML_AX, PML_BX, PML_DXH, PML_KXH, PML_AXH_ALPHA, PML_AXH,
       PML_BXH, PML_DY, PML_KY, PML_AY_ALPHA, PML_AY, PML_BY, PML_DYH, PML_KYH, PML_AYH_ALPHA,
       PML_AYH, PML_BYH, PML_N };

struct orc_ctx {
    orc_params P;
    int nd, fdoh, ntdtinv, fw2;
    real hc[7];
    real *pml[PML_N]; /* index 1..2FW */
    /* material (padded) */
    arr rho, pi, u, taus, taup, uipjp, rip, rjp, tausipjp, fipjp, f, g;
    arr d[9], e[9], dip[9];
    real bip[9], bjm[9], cip[9], cjm[9], peta[9];
    /* wavefields */
    arr vx, vy, sxx, syy, sxy, vxp1, vyp1, ux, uy, uxy;
    arr r[9], pp[9], q[9];
    /* CPML memory: x-bands [1..NY][1..2FW], y-bands [1..2FW][1..NX] */
    real *psi_sxx_x, *psi_sxy_x, *psi_vxx, *psi_vyx, *psi_vxxs;
    real *psi_syy_y, *psi_sxy_y, *psi_vyy, *psi_vxy;
    /* acquisition */
    int nsrc;
    float *srcpos;  /* 8 x nsrc */
    real *signals;  /* nsrc x NT */
    int ntr;
    int *recpos;    /* 2 x ntr */
    real *sec[5];   /* vx, vy, p, div, curl : ntr x NT */
    real *adj[3];   /* adjoint sources vx, vy, p : ntr x NT */
    /* FWI */
    real *fwd[5];   /* rho_x, rho_y, x, y, u : NX*NY*NTDTINV each */
    int *dtinv_help;
    arr wconv, wconv_u, wconv_rho;     /* per-shot correlation accumulators */
    arr wlam, wmu, wrho_s;              /* ass_gradPSV intermediates */
    arr gvp, gvs, grho;                 /* assembled per-shot gradients */
    int qtype;                          /* global QUELLTYP */
    int nan_step;
    float resw[2];                      /* COMP_WEIGHT factors of vx / vy (0 = none set) */
    /* anisotropic P/SV (PHYSICS 3 VTI, 4 TTI; 1 = isotropic) */
    int physics;
    arr uip, ujp; /* SH shear moduli at the sxz / syz points (av_mu_SH) */
    arr c11, c13, c33, c44, c44h, theta;
    arr d11, d13, d15, d15h, d33, d35, d35h, d55, d55h;
};

#define PSIX(a, j, h) ((a)[(size_t)((j) - 1) * (size_t)c->fw2 + (size_t)((h) - 1)])
#define PSIY(a, h, i) ((a)[(size_t)((h) - 1) * (size_t)c->P.nx + (size_t)((i) - 1)])

/* abs() of a float argument in src/PML_pro.c:153-154 etc. is the int abs():
 * the value is truncated to int first, so the gate opens only for |d| >= 1. */
static int pml_gate(real d)
{
    int v = (int)d;
    return abs(v) > 1.0e-6;
}

/* src/PML_pro.c:65-407 (single domain: NXG=NX, NYG=NY); the SH loop's
 * src/SH/PML_pro_SH.c:65-407 adds 0.1 alpha_max to alpha inside the layer */
#define PML_ALPHA(an) \
    (c->physics == 5 ? (real)(alpha_max * (1.0 - (an)) + 0.1 * alpha_max) : (real)(alpha_max * (1.0 - (an))))
static void pml_profiles(orc_ctx *c)
{
    const orc_params *P = &c->P;
    const int FW = P->fw;
    real *dx = c->pml[PML_DX], *Kx = c->pml[PML_KX], *alx = c->pml[PML_AX_ALPHA], *ax = c->pml[PML_AX],
         *bx = c->pml[PML_BX];
    real *dxh = c->pml[PML_DXH], *Kxh = c->pml[PML_KXH], *alxh = c->pml[PML_AXH_ALPHA],
         *axh = c->pml[PML_AXH], *bxh = c->pml[PML_BXH];
    real *dy = c->pml[PML_DY], *Ky = c->pml[PML_KY], *aly = c->pml[PML_AY_ALPHA], *ay = c->pml[PML_AY],
         *by = c->pml[PML_BY];
    real *dyh = c->pml[PML_DYH], *Kyh = c->pml[PML_KYH], *alyh = c->pml[PML_AYH_ALPHA],
         *ayh = c->pml[PML_AYH], *byh = c->pml[PML_BYH];
    const real npower = (real)P->npower, kmax = (real)P->k_max_pml;
    const real DH = (real)P->dh, DT = (real)P->dt;
    const real alpha_max = (real)(2.0 * PI_D * (P->fpml / 2.0));
    const real thick_x = FW * DH, thick_y = FW * DH;
    const real Rcoef = (real)0.001;
    const real d0_x = (real)((double)((-(npower + 1)) * (real)P->damping) * log((double)Rcoef) /
                             (2.0 * (double)thick_x));
    const real d0_y = (real)((double)((-(npower + 1)) * (real)P->damping) * log((double)Rcoef) /
                             (2.0 * (double)thick_y));
    real absc, an;

    /* x: left band */
    const real xoL = thick_x;
    const real xoR = (P->nx - 1) * DH - thick_x;
    for (int i = 1; i <= FW; i++) {
        Kx[i] = 1.0;
        Kxh[i] = 1.0;
        real xval = DH * (i - 1);
        absc = xoL - xval;
        if (absc >= 0.0) {
            an = absc / thick_x;
            dx[i] = (real)(d0_x * pow(an, npower));
            Kx[i] = (real)(1.0 + (kmax - 1.0) * pow(an, npower));
            alx[i] = PML_ALPHA(an);
        }
        absc = (real)(xoL - (xval + DH / 2.0));
        if (absc >= 0.0) {
            an = absc / thick_x;
            dxh[i] = (real)(d0_x * pow(an, npower));
            Kxh[i] = (real)(1.0 + (kmax - 1.0) * pow(an, npower));
            alxh[i] = PML_ALPHA(an);
        }
        if (alx[i] < 0.0) alx[i] = 0.0;
        if (alxh[i] < 0.0) alxh[i] = 0.0;
        bx[i] = (real)exp((double)(-(dx[i] / Kx[i] + alx[i]) * DT));
        bxh[i] = (real)exp((double)(-(dxh[i] / Kxh[i] + alxh[i]) * DT));
        if (pml_gate(dx[i]))
            ax[i] = (real)((double)dx[i] * ((double)bx[i] - 1.0) / (double)(Kx[i] * (dx[i] + Kx[i] * alx[i])));
        if (pml_gate(dxh[i]))
            axh[i] = (real)((double)dxh[i] * ((double)bxh[i] - 1.0) /
                            (double)(Kxh[i] * (dxh[i] + Kxh[i] * alxh[i])));
    }
    /* x: right band */
    for (int i = P->nx - FW + 1; i <= P->nx; i++) {
        int h = i - P->nx + 2 * FW;
        Kx[h] = 1.0;
        Kxh[h] = 1.0;
        real xval = DH * (i - 1);
        absc = xval - xoR;
        if (absc >= 0.0) {
            an = absc / thick_x;
            dx[h] = (real)(d0_x * pow(an, npower));
            Kx[h] = (real)(1.0 + (kmax - 1.0) * pow(an, npower));
            alx[h] = PML_ALPHA(an);
        }
        absc = (real)(xval + DH / 2.0 - xoR);
        if (absc >= 0.0) {
            an = absc / thick_x;
            dxh[h] = (real)(d0_x * pow(an, npower));
            Kxh[h] = (real)(1.0 + (kmax - 1.0) * pow(an, npower));
            alxh[h] = PML_ALPHA(an);
        }
        if (alx[h] < 0.0) alx[h] = 0.0;
        if (alxh[h] < 0.0) alxh[h] = 0.0;
        bx[h] = (real)exp((double)(-(dx[h] / Kx[h] + alx[h]) * DT));
        bxh[h] = (real)exp((double)(-(dxh[h] / Kxh[h] + alxh[h]) * DT));
        if (pml_gate(dx[h]))
            ax[h] = (real)((double)dx[h] * ((double)bx[h] - 1.0) / (double)(Kx[h] * (dx[h] + Kx[h] * alx[h])));
        if (pml_gate(dxh[h]))
            axh[h] = (real)((double)dxh[h] * ((double)bxh[h] - 1.0) /
                            (double)(Kxh[h] * (dxh[h] + Kxh[h] * alxh[h])));
    }
    /* y: top band (no alpha clamp in the reference's y loops) */
    const real yoT = thick_y;
    const real yoB = (P->ny - 1) * DH - thick_y;
    for (int i = 1; i <= FW; i++) {
        Ky[i] = 1.0;
        Kyh[i] = 1.0;
        real yval = DH * (i - 1);
        absc = yoT - yval;
        if (absc >= 0.0) {
            an = absc / thick_y;
            dy[i] = (real)(d0_y * pow(an, npower));
            Ky[i] = (real)(1.0 + (kmax - 1.0) * pow(an, npower));
            aly[i] = PML_ALPHA(an);
        }
        absc = (real)(yoT - (yval + DH / 2.0));
        if (absc >= 0.0) {
            an = absc / thick_y;
            dyh[i] = (real)(d0_y * pow(an, npower));
            Kyh[i] = (real)(1.0 + (kmax - 1.0) * pow(an, npower));
            alyh[i] = PML_ALPHA(an);
        }
        by[i] = (real)exp((double)(-(dy[i] / Ky[i] + aly[i]) * DT));
        byh[i] = (real)exp((double)(-(dyh[i] / Kyh[i] + alyh[i]) * DT));
        if (pml_gate(dy[i]))
            ay[i] = (real)((double)dy[i] * ((double)by[i] - 1.0) / (double)(Ky[i] * (dy[i] + Ky[i] * aly[i])));
        if (pml_gate(dyh[i]))
            ayh[i] = (real)((double)dyh[i] * ((double)byh[i] - 1.0) /
                            (double)(Kyh[i] * (dyh[i] + Kyh[i] * alyh[i])));
    }
    /* y: bottom band; the half-point damping uses d0_x (src/PML_pro.c:281) */
    for (int i = P->ny - FW + 1; i <= P->ny; i++) {
        int h = i - P->ny + 2 * FW;
        Ky[h] = 1.0;
        Kyh[h] = 1.0;
        real yval = DH * (i - 1);
        absc = yval - yoB;
        if (absc >= 0.0) {
            an = absc / thick_y;
            dy[h] = (real)(d0_y * pow(an, npower));
            Ky[h] = (real)(1.0 + (kmax - 1.0) * pow(an, npower));
            aly[h] = PML_ALPHA(an);
        }
        absc = (real)(yval + DH / 2.0 - yoB);
        if (absc >= 0.0) {
            an = absc / thick_y;
            dyh[h] = (real)(d0_x * pow(an, npower));
            Kyh[h] = (real)(1.0 + (kmax - 1.0) * pow(an, npower));
            alyh[h] = PML_ALPHA(an);
        }
        by[h] = (real)exp((double)(-(dy[h] / Ky[h] + aly[h]) * DT));
        byh[h] = (real)exp((double)(-(dyh[h] / Kyh[h] + alyh[h]) * DT));
        if (pml_gate(dy[h]))
            ay[h] = (real)((double)dy[h] * ((double)by[h] - 1.0) / (double)(Ky[h] * (dy[h] + Ky[h] * aly[h])));
        if (pml_gate(dyh[h]))
            ayh[h] = (real)((double)dyh[h] * ((double)byh[h] - 1.0) /
                            (double)(Kyh[h] * (dyh[h] + Kyh[h] * alyh[h])));
    }
}

/* Ghost row/column fill of the material with one rank: the self send/receive
 * of src/PSV/matcopy_elastic_PSV.c:32-126 (and matcopy_PSV.c) makes row 0 a
 * copy of row NY, row NY+1 a copy of row 1, then columns 0 / NX+1 copies of
 * columns NX / 1 including the new ghost rows. */
static void matcopy_periodic(orc_ctx *c, arr a)
{
    const int NX = c->P.nx, NY = c->P.ny;
    for (int i = 0; i <= NX + 1; i++) {
        real top = AT(a, 1, i), bot = AT(a, NY, i);
        AT(a, NY + 1, i) = top;
        AT(a, 0, i) = bot;
    }
    for (int j = 0; j <= NY + 1; j++) {
        real lef = AT(a, j, 1), rig = AT(a, j, NX);
        AT(a, j, NX + 1) = lef;
        AT(a, j, 0) = rig;
    }
}

/* src/av_mue.c:5-51 */
static void av_mue(orc_ctx *c)
{
    const int NX = c->P.nx, NY = c->P.ny;
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            real u1, u2, u3, u4;
            if (c->P.invmat1 == 3) {
                u1 = AT(c->u, j, i); u2 = AT(c->u, j, i + 1);
                u3 = AT(c->u, j + 1, i); u4 = AT(c->u, j + 1, i + 1);
            } else {
                u1 = AT(c->rho, j, i) * AT(c->u, j, i) * AT(c->u, j, i);
                u2 = AT(c->rho, j, i + 1) * AT(c->u, j, i + 1) * AT(c->u, j, i + 1);
                u3 = AT(c->rho, j + 1, i) * AT(c->u, j + 1, i) * AT(c->u, j + 1, i);
                u4 = AT(c->rho, j + 1, i + 1) * AT(c->u, j + 1, i + 1) * AT(c->u, j + 1, i + 1);
            }
            real v = (real)(4.0 / ((1.0 / u1) + (1.0 / u2) + (1.0 / u3) + (1.0 / u4)));
            if (u1 == 0.0 || u2 == 0.0 || u3 == 0.0 || u4 == 0.0) v = 0.0;
            AT(c->uipjp, j, i) = v;
        }
}

/* src/av_rho.c:5-28 */
static void av_rho(orc_ctx *c)
{
    const int NX = c->P.nx, NY = c->P.ny;
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            AT(c->rjp, j, i) = (real)(1.0 / (0.5 * (AT(c->rho, j, i) + AT(c->rho, j + 1, i))));
            AT(c->rip, j, i) = (real)(1.0 / (0.5 * (AT(c->rho, j, i) + AT(c->rho, j, i + 1))));
            if (AT(c->rho, j, i) < 1e-4 && AT(c->rho, j + 1, i) < 1e-4) AT(c->rjp, j, i) = 0.0;
            if (AT(c->rho, j, i) < 1e-4 && AT(c->rho, j, i + 1) < 1e-4) AT(c->rip, j, i) = 0.0;
        }
}

/* src/av_tau.c:5-15 */
static void av_tau(orc_ctx *c)
{
    for (int j = 1; j <= c->P.ny; j++)
        for (int i = 1; i <= c->P.nx; i++)
            AT(c->tausipjp, j, i) = (real)(0.25 * (AT(c->taus, j, i) + AT(c->taus, j, i + 1) +
                                                   AT(c->taus, j + 1, i) + AT(c->taus, j + 1, i + 1)));
}

/* src/PSV/prepare_update_s_visc_PSV.c:8-94 (peta from readmod_visc_PSV.c:29-32) */
static void prepare_visc(orc_ctx *c)
{
    const int L = c->P.L;
    const real DT = (real)c->P.dt;
    real pts[9];
    for (int l = 1; l <= L; l++) {
        pts[l] = (real)(1.0 / (2.0 * PI_D * c->P.fl[l - 1]));
        c->peta[l] = DT / pts[l];
    }
    real ws = (real)(2.0 * PI_D * c->P.fl[0]);
    real sumu = 0.0, sumpi = 0.0;
    for (int l = 1; l <= L; l++) {
        sumu = (real)(sumu + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l])));
        sumpi = (real)(sumpi + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l])));
    }
    for (int l = 1; l <= L; l++) {
        real eta = c->peta[l];
        c->bip[l] = (real)(1.0 / (1.0 + (eta * 0.5)));
        c->bjm[l] = (real)(1.0 / (1.0 + (eta * 0.5)));
        c->cip[l] = (real)(1.0 - (eta * 0.5));
        c->cjm[l] = (real)(1.0 - (eta * 0.5));
    }
    for (int j = 1; j <= c->P.ny; j++)
        for (int i = 1; i <= c->P.nx; i++) {
            real mu, pi, muipjp;
            real ts = AT(c->taus, j, i), tp = AT(c->taup, j, i), tsij = AT(c->tausipjp, j, i);
            if (c->P.invmat1 == 3) {
                mu = (real)(AT(c->u, j, i) / (1.0 + sumu * ts));
                pi = (real)((AT(c->pi, j, i) + 2 * AT(c->u, j, i)) / (1.0 + sumpi * tp));
            } else {
                mu = (real)((AT(c->u, j, i) * AT(c->u, j, i) * AT(c->rho, j, i)) / (1.0 + sumu * ts));
                pi = (real)((AT(c->pi, j, i) * AT(c->pi, j, i) * AT(c->rho, j, i)) / (1.0 + sumpi * tp));
            }
            muipjp = (real)(AT(c->uipjp, j, i) / (1.0 + sumu * tsij));
            AT(c->fipjp, j, i) = (real)(muipjp * DT * (1.0 + L * tsij));
            AT(c->f, j, i) = (real)(mu * DT * (1.0 + L * ts));
            AT(c->g, j, i) = (real)(pi * DT * (1.0 + L * tp));
            for (int l = 1; l <= L; l++) {
                AT(c->dip[l], j, i) = muipjp * c->peta[l] * tsij;
                AT(c->d[l], j, i) = mu * c->peta[l] * ts;
                AT(c->e[l], j, i) = pi * c->peta[l] * tp;
            }
        }
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    for (int k = 0; k < PML_N; k++) free(c->pml[k]);
    arr *all[] = {&c->rho, &c->pi, &c->u, &c->taus, &c->taup, &c->uipjp, &c->rip, &c->rjp,
                  &c->tausipjp, &c->fipjp, &c->f, &c->g, &c->vx, &c->vy, &c->sxx, &c->syy,
                  &c->sxy, &c->vxp1, &c->vyp1, &c->ux, &c->uy, &c->uxy, &c->wconv, &c->wconv_u,
                  &c->wconv_rho, &c->wlam, &c->wmu, &c->wrho_s, &c->gvp, &c->gvs, &c->grho,
                  &c->c11, &c->c13, &c->c33, &c->c44, &c->c44h, &c->theta, &c->d11, &c->d13,
                  &c->d15, &c->d15h, &c->d33, &c->d35, &c->d35h, &c->d55, &c->d55h, &c->uip, &c->ujp};
    for (size_t k = 0; k < sizeof all / sizeof all[0]; k++) free(all[k]->p);
    for (int l = 0; l < 9; l++) {
        free(c->d[l].p); free(c->e[l].p); free(c->dip[l].p);
        free(c->r[l].p); free(c->pp[l].p); free(c->q[l].p);
    }
    free(c->psi_sxx_x); free(c->psi_sxy_x); free(c->psi_vxx); free(c->psi_vyx); free(c->psi_vxxs);
    free(c->psi_syy_y); free(c->psi_sxy_y); free(c->psi_vyy); free(c->psi_vxy);
    free(c->srcpos); free(c->signals); free(c->recpos);
    for (int k = 0; k < 5; k++) { free(c->sec[k]); free(c->fwd[k]); }
    for (int k = 0; k < 3; k++) free(c->adj[k]);
    free(c->dtinv_help);
    free(c);
}

/* Context with the grid, FD coefficients, CPML profiles, wavefields and
 * material arrays allocated (alloc_PSV, PML_pro, holbergcoeff of
 * src/PSV/FD_PSV.c:170-199); the model is loaded by the callers. */
static orc_ctx *ctx_new(const orc_params *p)
{
    if (p->nx < 1 || p->ny < 1 || p->nt < 2) { seterr("bad grid/NT"); return NULL; }
    if (p->L < 0 || p->L > 8) { seterr("L out of range"); return NULL; }
    if (p->fw < 0 || p->fw > p->nx || p->fw > p->ny) { seterr("FW out of range"); return NULL; }
    if (p->dtinv < 1) { seterr("DTINV must be >= 1"); return NULL; }
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof *c);
    c->P = *p;
    float hcf[7];
    if (orc_holberg(p->fdorder, p->maxrelerror, hcf)) { free(c); return NULL; }
    for (int k = 0; k < 7; k++) c->hc[k] = hcf[k];
    c->fdoh = p->fdorder / 2;
    c->nd = c->fdoh + 1;
    c->fw2 = 2 * p->fw;
    c->ntdtinv = (int)ceil((float)p->nt / (float)p->dtinv); /* src/PSV/FD_PSV.c:80 */
    const int NX = p->nx, NY = p->ny, nd = c->nd;
    for (int k = 0; k < PML_N; k++) c->pml[k] = (real *)calloc((size_t)c->fw2 + 2, sizeof(real));
    if (p->fw > 0) pml_profiles(c);

#define NEWA() arr_new(-nd + 1, NY + nd, -nd + 1, NX + nd)
    c->rho = NEWA(); c->pi = NEWA(); c->u = NEWA(); c->taus = NEWA(); c->taup = NEWA();
    c->uipjp = NEWA(); c->rip = NEWA(); c->rjp = NEWA(); c->tausipjp = NEWA();
    c->fipjp = NEWA(); c->f = NEWA(); c->g = NEWA();
    c->vx = NEWA(); c->vy = NEWA(); c->sxx = NEWA(); c->syy = NEWA(); c->sxy = NEWA();
    c->vxp1 = NEWA(); c->vyp1 = NEWA(); c->ux = NEWA(); c->uy = NEWA(); c->uxy = NEWA();
    c->wconv = NEWA(); c->wconv_u = NEWA(); c->wconv_rho = NEWA();
    c->wlam = NEWA(); c->wmu = NEWA(); c->wrho_s = NEWA();
    c->gvp = NEWA(); c->gvs = NEWA(); c->grho = NEWA();
    for (int l = 1; l <= p->L; l++) {
        c->d[l] = NEWA(); c->e[l] = NEWA(); c->dip[l] = NEWA();
        c->r[l] = NEWA(); c->pp[l] = NEWA(); c->q[l] = NEWA();
    }
#undef NEWA
    size_t nxb = (size_t)NY * c->fw2 + 1, nyb = (size_t)c->fw2 * NX + 1;
    c->psi_sxx_x = calloc(nxb, sizeof(real)); c->psi_sxy_x = calloc(nxb, sizeof(real));
    c->psi_vxx = calloc(nxb, sizeof(real)); c->psi_vyx = calloc(nxb, sizeof(real));
    c->psi_vxxs = calloc(nxb, sizeof(real));
    c->psi_syy_y = calloc(nyb, sizeof(real)); c->psi_sxy_y = calloc(nyb, sizeof(real));
    c->psi_vyy = calloc(nyb, sizeof(real)); c->psi_vxy = calloc(nyb, sizeof(real));
    c->dtinv_help = (int *)calloc((size_t)p->nt + 2, sizeof(int));
    c->nan_step = 0;
    c->physics = 1;
    return c;
}

orc_ctx *orc_create(const orc_params *p, const float *vp, const float *vs, const float *rho,
                    const float *taus, const float *taup)
{
    if (!p || !vp || !vs || !rho) { seterr("null argument"); return NULL; }
    if (p->L > 0 && (!taus || !taup)) { seterr("visco model needs taus/taup"); return NULL; }
    orc_ctx *c = ctx_new(p);
    if (!c) return NULL;
    const int NX = p->nx, NY = p->ny;
    /* model files: loop i=1..NX outer, j=1..NY inner (readmod_elastic_PSV.c:72-76) */
    for (int i = 1; i <= NX; i++)
        for (int j = 1; j <= NY; j++) {
            size_t k = (size_t)(i - 1) * NY + (j - 1);
            AT(c->pi, j, i) = vp[k];
            AT(c->u, j, i) = vs[k];
            AT(c->rho, j, i) = rho[k];
            if (p->L > 0) {
                AT(c->taus, j, i) = taus[k];
                AT(c->taup, j, i) = taup[k];
            }
        }
    matcopy_periodic(c, c->rho);
    matcopy_periodic(c, c->pi);
    matcopy_periodic(c, c->u);
    if (p->L > 0) {
        matcopy_periodic(c, c->taus);
        matcopy_periodic(c, c->taup);
    }
    av_mue(c);
    av_rho(c);
    if (p->L > 0) {
        av_tau(c);
        prepare_visc(c);
    }
    return c;
}

/* ------------------------------------------------------------------------- */
/* anisotropic P/SV: VTI (PHYSICS 3) and TTI (PHYSICS 4)                      */
/* ------------------------------------------------------------------------- */

/* src/av_harm.c:9-24: harmonic mean of the four cells around (i+1/2, j+1/2),
 * 0 where any of them is 0 */
static void av_harm(orc_ctx *c, arr m, arr mh)
{
    for (int j = 1; j <= c->P.ny; j++)
        for (int i = 1; i <= c->P.nx; i++) {
            real v = (real)(4.0 / ((1.0 / AT(m, j, i)) + (1.0 / AT(m, j, i + 1)) + (1.0 / AT(m, j + 1, i)) +
                                   (1.0 / AT(m, j + 1, i + 1))));
            if (AT(m, j, i) == 0.0 || AT(m, j, i + 1) == 0.0 || AT(m, j + 1, i) == 0.0 || AT(m, j + 1, i + 1) == 0.0)
                v = 0.0;
            AT(mh, j, i) = v;
        }
}

/* src/TTI/rot_el_tensor_TTI.c:18-40: the VTI tensor c11, c13, c33, c44
 * rotated by the tilt angle theta (radians); float locals as there */
static void rot_el_tensor(orc_ctx *c)
{
    for (int j = 1; j <= c->P.ny; j++)
        for (int i = 1; i <= c->P.nx; i++) {
            const real th = AT(c->theta, j, i);
            real cos2 = (real)(cos(th) * cos(th));
            real sin2 = (real)(sin(th) * sin(th));
            real cos4 = cos2 * cos2;
            real sin4 = sin2 * sin2;
            real sincos = (real)(sin(th) * cos(th));
            real cos2sin2 = cos2 * sin2;
            real cos2msin2 = cos2 - sin2;
            const real c11 = AT(c->c11, j, i), c13 = AT(c->c13, j, i), c33 = AT(c->c33, j, i), c44 = AT(c->c44, j, i);
            AT(c->d11, j, i) = (real)(cos2 * (c11 * cos2 + c13 * sin2) + sin2 * (c13 * cos2 + c33 * sin2) +
                                      4.0 * c44 * cos2sin2);
            AT(c->d13, j, i) = (real)(c11 * cos2sin2 + c13 * sin4 + c13 * cos4 + c33 * cos2sin2 - 4.0 * c44 * cos2sin2);
            AT(c->d15, j, i) = (real)((c13 - c11) * cos2 * sincos + (c33 - c13) * sin2 * sincos +
                                      2.0 * sincos * c44 * cos2msin2);
            AT(c->d33, j, i) = (real)(sin2 * (c11 * sin2 + c13 * cos2) + cos2 * (c33 * cos2 + c13 * sin2) +
                                      4.0 * c44 * cos2sin2);
            AT(c->d35, j, i) = (real)((c13 - c11) * sin2 * sincos + (c33 - c13) * cos2 * sincos -
                                      2.0 * sincos * c44 * cos2msin2);
            AT(c->d55, j, i) = (real)(c44 * (1.0 - 2.0 * sin2) * (1.0 - 2.0 * sin2) + (c33 - c13) * cos2sin2 -
                                      (c13 - c11) * cos2sin2);
        }
}

/* Model set-up of src/VTI/FD_VTI.c:180-197 (readmod_elastic_VTI.c:60-79,
 * matcopy_elastic_VTI with one rank, av_harm, av_rho) and of
 * src/TTI/FD_TTI.c:180-205 (readmod_elastic_TTI.c:64-85 with theta given in
 * degrees, rot_el_tensor_TTI, matcopy of d15, d35, d55, av_harm of each). */
orc_ctx *orc_create_aniso(const orc_params *p, int physics, const float *rho, const float *c11,
                          const float *c13, const float *c33, const float *c44, const float *theta)
{
    if (!p || !rho || !c11 || !c13 || !c33 || !c44) { seterr("null argument"); return NULL; }
    if (physics != 3 && physics != 4) { seterr("physics must be 3 (VTI) or 4 (TTI)"); return NULL; }
    if (physics == 4 && !theta) { seterr("TTI needs theta"); return NULL; }
    if (p->L != 0) { seterr("anisotropic runs are elastic (L = 0)"); return NULL; }
    /* TTI.c stores ux/uy/uxy for GRAD_FORM 2, which update_s_elastic_PML_TTI never writes */
    if (physics == 4 && p->mode > 0 && p->grad_form == 2) { seterr("TTI adjoint runs need GRAD_FORM 1"); return NULL; }
    orc_ctx *c = ctx_new(p);
    if (!c) return NULL;
    c->physics = physics;
    const int NX = p->nx, NY = p->ny, nd = c->nd;
#define NEWA() arr_new(-nd + 1, NY + nd, -nd + 1, NX + nd)
    c->c11 = NEWA(); c->c13 = NEWA(); c->c33 = NEWA(); c->c44 = NEWA(); c->c44h = NEWA();
    if (physics == 4) {
        c->theta = NEWA(); c->d11 = NEWA(); c->d13 = NEWA(); c->d15 = NEWA(); c->d15h = NEWA();
        c->d33 = NEWA(); c->d35 = NEWA(); c->d35h = NEWA(); c->d55 = NEWA(); c->d55h = NEWA();
    }
#undef NEWA
    for (int i = 1; i <= NX; i++)
        for (int j = 1; j <= NY; j++) {
            size_t k = (size_t)(i - 1) * NY + (j - 1);
            AT(c->c11, j, i) = c11[k];
            AT(c->c13, j, i) = c13[k];
            AT(c->c33, j, i) = c33[k];
            AT(c->c44, j, i) = c44[k];
            AT(c->rho, j, i) = rho[k];
            if (physics == 4) AT(c->theta, j, i) = (real)(theta[k] * PI_D / 180.0);
        }
    matcopy_periodic(c, c->rho);
    matcopy_periodic(c, c->c11);
    if (physics == 3) {
        matcopy_periodic(c, c->c44);
        av_harm(c, c->c44, c->c44h);
    } else {
        rot_el_tensor(c);
        matcopy_periodic(c, c->d15);
        matcopy_periodic(c, c->d35);
        matcopy_periodic(c, c->d55);
        av_harm(c, c->d15, c->d15h);
        av_harm(c, c->d35, c->d35h);
        av_harm(c, c->d55, c->d55h);
    }
    av_rho(c);
    return c;
}

int orc_set_sources(orc_ctx *c, int nsrc, const float *srcpos, const float *signals)
{
    free(c->srcpos); free(c->signals);
    c->nsrc = nsrc;
    c->srcpos = (float *)calloc((size_t)8 * nsrc + 1, sizeof(float));
    c->signals = (real *)calloc((size_t)nsrc * c->P.nt + 1, sizeof(real));
    if (nsrc > 0) memcpy(c->srcpos, srcpos, sizeof(float) * 8 * (size_t)nsrc);
    for (size_t k = 0; k < (size_t)nsrc * c->P.nt; k++) c->signals[k] = signals[k];
    /* FD_PSV.c:297: QUELLTYP of the shot */
    c->qtype = nsrc > 0 ? (int)c->srcpos[7 * nsrc + 0] : 0;
    for (int l = 0; l < nsrc; l++) {
        int i = (int)c->srcpos[0 * nsrc + l], j = (int)c->srcpos[1 * nsrc + l];
        if (i < 1 || i > c->P.nx || j < 1 || j > c->P.ny) {
            seterr("source %d outside the grid", l + 1);
            return -1;
        }
    }
    return 0;
}

int orc_set_receivers(orc_ctx *c, int ntr, const int *recpos)
{
    free(c->recpos);
    for (int k = 0; k < 5; k++) { free(c->sec[k]); c->sec[k] = NULL; }
    for (int k = 0; k < 3; k++) { free(c->adj[k]); c->adj[k] = NULL; }
    c->ntr = ntr;
    c->recpos = (int *)calloc((size_t)2 * ntr + 1, sizeof(int));
    if (ntr > 0) memcpy(c->recpos, recpos, sizeof(int) * 2 * (size_t)ntr);
    for (int k = 0; k < 5; k++) c->sec[k] = (real *)calloc((size_t)ntr * c->P.nt + 1, sizeof(real));
    for (int k = 0; k < 3; k++) c->adj[k] = (real *)calloc((size_t)ntr * c->P.nt + 1, sizeof(real));
    return 0;
}

/* src/PSV/zero_denise_elast_PSV.c:9-64, zero_denise_visc_PSV.c */
static void zero_wavefields(orc_ctx *c)
{
    arr_zero(c->vx); arr_zero(c->vy); arr_zero(c->sxx); arr_zero(c->syy); arr_zero(c->sxy);
    arr_zero(c->ux); arr_zero(c->uy); arr_zero(c->uxy); arr_zero(c->vxp1); arr_zero(c->vyp1);
    size_t nxb = (size_t)c->P.ny * c->fw2, nyb = (size_t)c->fw2 * c->P.nx;
    memset(c->psi_sxx_x, 0, nxb * sizeof(real)); memset(c->psi_sxy_x, 0, nxb * sizeof(real));
    memset(c->psi_vxx, 0, nxb * sizeof(real)); memset(c->psi_vyx, 0, nxb * sizeof(real));
    memset(c->psi_vxxs, 0, nxb * sizeof(real));
    memset(c->psi_syy_y, 0, nyb * sizeof(real)); memset(c->psi_sxy_y, 0, nyb * sizeof(real));
    memset(c->psi_vyy, 0, nyb * sizeof(real)); memset(c->psi_vxy, 0, nyb * sizeof(real));
    for (int l = 1; l <= c->P.L; l++) { arr_zero(c->r[l]); arr_zero(c->pp[l]); arr_zero(c->q[l]); }
}

/* src/PSV/update_v_PML_PSV.c:14-786 (FD blocks :54-715 are one unrolled case
 * per order with identical structure; sources :739-768) */
static void update_v(orc_ctx *c, int nt, int sw)
{
    const orc_params *P = &c->P;
    const int NX = P->nx, NY = P->ny, FW = P->fw, fdoh = c->fdoh;
    const real DT = (real)P->dt, DH = (real)P->dh;
    const real *hc = c->hc;
    real *const *pm = c->pml;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= NY; j++) {
        for (int i = 1; i <= NX; i++) {
            real sxx_x = hc[1] * (AT(c->sxx, j, i + 1) - AT(c->sxx, j, i));
            real sxy_x = hc[1] * (AT(c->sxy, j, i) - AT(c->sxy, j, i - 1));
            real sxy_y = hc[1] * (AT(c->sxy, j, i) - AT(c->sxy, j - 1, i));
            real syy_y = hc[1] * (AT(c->syy, j + 1, i) - AT(c->syy, j, i));
            for (int m = 2; m <= fdoh; m++) {
                sxx_x = sxx_x + hc[m] * (AT(c->sxx, j, i + m) - AT(c->sxx, j, i - m + 1));
                sxy_x = sxy_x + hc[m] * (AT(c->sxy, j, i + m - 1) - AT(c->sxy, j, i - m));
                sxy_y = sxy_y + hc[m] * (AT(c->sxy, j + m - 1, i) - AT(c->sxy, j - m, i));
                syy_y = syy_y + hc[m] * (AT(c->syy, j + m, i) - AT(c->syy, j - m + 1, i));
            }
            if (!P->boundary && i <= FW) {
                real *ps = &PSIX(c->psi_sxx_x, j, i), *pxy = &PSIX(c->psi_sxy_x, j, i);
                *ps = pm[PML_BXH][i] * *ps + pm[PML_AXH][i] * sxx_x;
                sxx_x = sxx_x / pm[PML_KXH][i] + *ps;
                *pxy = pm[PML_BX][i] * *pxy + pm[PML_AX][i] * sxy_x;
                sxy_x = sxy_x / pm[PML_KX][i] + *pxy;
            }
            if (!P->boundary && i >= NX - FW + 1) {
                int h1 = i - NX + 2 * FW;
                real *ps = &PSIX(c->psi_sxx_x, j, h1), *pxy = &PSIX(c->psi_sxy_x, j, h1);
                *ps = pm[PML_BXH][h1] * *ps + pm[PML_AXH][h1] * sxx_x;
                sxx_x = sxx_x / pm[PML_KXH][h1] + *ps;
                *pxy = pm[PML_BX][h1] * *pxy + pm[PML_AX][h1] * sxy_x;
                sxy_x = sxy_x / pm[PML_KX][h1] + *pxy;
            }
            if (!P->free_surf && j <= FW) {
                real *ps = &PSIY(c->psi_syy_y, j, i), *pxy = &PSIY(c->psi_sxy_y, j, i);
                *ps = pm[PML_BYH][j] * *ps + pm[PML_AYH][j] * syy_y;
                syy_y = syy_y / pm[PML_KYH][j] + *ps;
                *pxy = pm[PML_BY][j] * *pxy + pm[PML_AY][j] * sxy_y;
                sxy_y = sxy_y / pm[PML_KY][j] + *pxy;
            }
            if (j >= NY - FW + 1) {
                int h1 = j - NY + 2 * FW;
                real *ps = &PSIY(c->psi_syy_y, h1, i), *pxy = &PSIY(c->psi_sxy_y, h1, i);
                *ps = pm[PML_BYH][h1] * *ps + pm[PML_AYH][h1] * syy_y;
                syy_y = syy_y / pm[PML_KYH][h1] + *ps;
                *pxy = pm[PML_BY][h1] * *pxy + pm[PML_AY][h1] * sxy_y;
                sxy_y = sxy_y / pm[PML_KY][h1] + *pxy;
            }
            real rip = AT(c->rip, j, i), rjp = AT(c->rjp, j, i);
            if (P->grad_form == 1) {
                if (sw == 0) {
                    AT(c->vxp1, j, i) = rip * (sxx_x + sxy_y) / DH;
                    AT(c->vyp1, j, i) = rjp * (sxy_x + syy_y) / DH;
                } else {
                    AT(c->vxp1, j, i) += DT * AT(c->vx, j, i);
                    AT(c->vyp1, j, i) += DT * AT(c->vy, j, i);
                }
            } else if (P->grad_form == 2) {
                if (sw == 0) {
                    AT(c->vxp1, j, i) = rip * (sxx_x + sxy_y) / DH;
                    AT(c->vyp1, j, i) = rjp * (sxy_x + syy_y) / DH;
                } else {
                    AT(c->vxp1, j, i) = AT(c->vx, j, i);
                    AT(c->vyp1, j, i) = AT(c->vy, j, i);
                }
            }
            AT(c->vx, j, i) += DT * rip * (sxx_x + sxy_y) / DH;
            AT(c->vy, j, i) += DT * rjp * (sxy_x + syy_y) / DH;
        }
    }
    if (sw == 0) {
        for (int l = 0; l < c->nsrc; l++) {
            int i = (int)c->srcpos[0 * c->nsrc + l], j = (int)c->srcpos[1 * c->nsrc + l];
            real azi = (real)(c->srcpos[6 * c->nsrc + l] * PI_D / 180);
            int type = (int)c->srcpos[7 * c->nsrc + l];
            real s = c->signals[(size_t)l * P->nt + (nt - 1)];
            c->qtype = type;
            if (type == 2) AT(c->vx, j, i) += s;
            if (type == 3) AT(c->vy, j, i) += s;
            if (type == 4) {
                AT(c->vx, j, i) = (real)(AT(c->vx, j, i) + -sin(azi) * s);
                AT(c->vy, j, i) = (real)(AT(c->vy, j, i) + cos(azi) * s);
            }
        }
    } else {
        const int QB = P->quelltypb;
        for (int l = 0; l < c->ntr; l++) {
            int i = c->recpos[l], j = c->recpos[c->ntr + l];
            real sx = c->adj[0][(size_t)l * P->nt + (nt - 1)];
            real sy = c->adj[1][(size_t)l * P->nt + (nt - 1)];
            if (P->grad_form == 1 || P->grad_form == 2) {
                if (QB == 1) { AT(c->vx, j, i) += sx; AT(c->vy, j, i) += sy; }
                if (QB == 2 || QB == 6 || QB == 7) AT(c->vy, j, i) += sy;
                if (QB == 3 || QB == 5 || QB == 7) AT(c->vx, j, i) += sx;
            }
        }
    }
}

/* stress-sweep derivatives, src/PSV/update_s_elastic_PML_PSV.c:360-378 and
 * the CPML corrections :380-433 (identical in update_s_visc_PML_PSV.c) */
static void stress_derivs(orc_ctx *c, int j, int i, real dhi, real *pvxx, real *pvyy, real *pvxy,
                          real *pvyx)
{
    const orc_params *P = &c->P;
    const int NX = P->nx, NY = P->ny, FW = P->fw, fdoh = c->fdoh;
    const real *hc = c->hc;
    real *const *pm = c->pml;
    real vxx = hc[1] * (AT(c->vx, j, i) - AT(c->vx, j, i - 1));
    real vyx = hc[1] * (AT(c->vy, j, i + 1) - AT(c->vy, j, i));
    real vxy = hc[1] * (AT(c->vx, j + 1, i) - AT(c->vx, j, i));
    real vyy = hc[1] * (AT(c->vy, j, i) - AT(c->vy, j - 1, i));
    for (int m = 2; m <= fdoh; m++) {
        vxx = vxx + hc[m] * (AT(c->vx, j, i + m - 1) - AT(c->vx, j, i - m));
        vyx = vyx + hc[m] * (AT(c->vy, j, i + m) - AT(c->vy, j, i - m + 1));
        vxy = vxy + hc[m] * (AT(c->vx, j + m, i) - AT(c->vx, j - m + 1, i));
        vyy = vyy + hc[m] * (AT(c->vy, j + m - 1, i) - AT(c->vy, j - m, i));
    }
    vxx = vxx * dhi;
    vyx = vyx * dhi;
    vxy = vxy * dhi;
    vyy = vyy * dhi;
    if (!P->boundary && i <= FW) {
        real *a = &PSIX(c->psi_vxx, j, i), *b = &PSIX(c->psi_vyx, j, i);
        *a = pm[PML_BX][i] * *a + pm[PML_AX][i] * vxx;
        vxx = vxx / pm[PML_KX][i] + *a;
        *b = pm[PML_BXH][i] * *b + pm[PML_AXH][i] * vyx;
        vyx = vyx / pm[PML_KXH][i] + *b;
    }
    if (!P->boundary && i >= NX - FW + 1) {
        int h1 = i - NX + 2 * FW;
        real *a = &PSIX(c->psi_vxx, j, h1), *b = &PSIX(c->psi_vyx, j, h1);
        *a = pm[PML_BX][h1] * *a + pm[PML_AX][h1] * vxx;
        vxx = vxx / pm[PML_KX][h1] + *a;
        *b = pm[PML_BXH][h1] * *b + pm[PML_AXH][h1] * vyx;
        vyx = vyx / pm[PML_KXH][h1] + *b;
    }
    if (!P->free_surf && j <= FW) {
        real *a = &PSIY(c->psi_vyy, j, i), *b = &PSIY(c->psi_vxy, j, i);
        *a = pm[PML_BY][j] * *a + pm[PML_AY][j] * vyy;
        *b = pm[PML_BYH][j] * *b + pm[PML_AYH][j] * vxy;
        vyy = vyy / pm[PML_KY][j] + *a;
        vxy = vxy / pm[PML_KYH][j] + *b;
    }
    if (j >= NY - FW + 1) {
        int h1 = j - NY + 2 * FW;
        real *a = &PSIY(c->psi_vyy, h1, i), *b = &PSIY(c->psi_vxy, h1, i);
        *a = pm[PML_BY][h1] * *a + pm[PML_AY][h1] * vyy;
        vyy = vyy / pm[PML_KY][h1] + *a;
        *b = pm[PML_BYH][h1] * *b + pm[PML_AYH][h1] * vxy;
        vxy = vxy / pm[PML_KYH][h1] + *b;
    }
    *pvxx = vxx; *pvyy = vyy; *pvxy = vxy; *pvyx = vyx;
}

/* src/PSV/update_s_elastic_PML_PSV.c:13-742 */
static void update_s_elastic(orc_ctx *c, int mode)
{
    const orc_params *P = &c->P;
    const real DT = (real)P->dt, dhi = (real)P->dt / (real)P->dh;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= P->ny; j++) {
        for (int i = 1; i <= P->nx; i++) {
            real vxx, vyy, vxy, vyx, f, g;
            stress_derivs(c, j, i, dhi, &vxx, &vyy, &vxy, &vyx);
            real fipjp = AT(c->uipjp, j, i);
            if (P->invmat1 == 3) {
                f = AT(c->u, j, i);
                g = AT(c->pi, j, i);
            } else {
                f = AT(c->rho, j, i) * AT(c->u, j, i) * AT(c->u, j, i);
                g = AT(c->rho, j, i) * ((AT(c->pi, j, i) * AT(c->pi, j, i)) - 2 * AT(c->u, j, i) * AT(c->u, j, i));
            }
            if (mode == 0 && P->grad_form == 2) {
                AT(c->ux, j, i) = (real)((g * (vxx + vyy) + (2.0 * f * vxx)) / DT);
                AT(c->uy, j, i) = (real)((g * (vxx + vyy) + (2.0 * f * vyy)) / DT);
                AT(c->uxy, j, i) = fipjp * (vyx + vxy) / DT;
            }
            AT(c->sxy, j, i) += fipjp * (vyx + vxy);
            AT(c->sxx, j, i) = (real)(AT(c->sxx, j, i) + (g * (vxx + vyy) + (2.0 * f * vxx)));
            AT(c->syy, j, i) = (real)(AT(c->syy, j, i) + (g * (vxx + vyy) + (2.0 * f * vyy)));
        }
    }
}

/* src/PSV/update_s_visc_PML_PSV.c:13-825 (FD8 block :391-511) */
static void update_s_visc(orc_ctx *c, int mode)
{
    const orc_params *P = &c->P;
    const int L = P->L;
    const real dhi = (real)(1.0 / P->dh);
    const real dthalbe = (real)(P->dt / 2.0);
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= P->ny; j++) {
        for (int i = 1; i <= P->nx; i++) {
            real vxx, vyy, vxy, vyx;
            stress_derivs(c, j, i, dhi, &vxx, &vyy, &vxy, &vyx);
            real sumr = 0.0, sump = 0.0, sumq = 0.0;
            for (int l = 1; l <= L; l++) {
                sumr += AT(c->r[l], j, i);
                sump += AT(c->pp[l], j, i);
                sumq += AT(c->q[l], j, i);
            }
            real fipjp = AT(c->fipjp, j, i), f = AT(c->f, j, i), g = AT(c->g, j, i);
            AT(c->sxy, j, i) += (fipjp * (vxy + vyx)) + (dthalbe * sumr);
            AT(c->sxx, j, i) = (real)(AT(c->sxx, j, i) + ((g * (vxx + vyy)) - (2.0 * f * vyy) + (dthalbe * sump)));
            AT(c->syy, j, i) = (real)(AT(c->syy, j, i) + ((g * (vxx + vyy)) - (2.0 * f * vxx) + (dthalbe * sumq)));
            sumr = sump = sumq = 0.0;
            for (int l = 1; l <= L; l++) {
                real *r = &AT(c->r[l], j, i), *p = &AT(c->pp[l], j, i), *q = &AT(c->q[l], j, i);
                real dip = AT(c->dip[l], j, i), d = AT(c->d[l], j, i), e = AT(c->e[l], j, i);
                *r = c->bip[l] * (*r * c->cip[l] - (dip * (vxy + vyx)));
                *p = (real)(c->bjm[l] * (*p * c->cjm[l] - (e * (vxx + vyy)) + (2.0 * d * vyy)));
                *q = (real)(c->bjm[l] * (*q * c->cjm[l] - (e * (vxx + vyy)) + (2.0 * d * vxx)));
                sumr += *r;
                sump += *p;
                sumq += *q;
            }
            AT(c->sxy, j, i) += (dthalbe * sumr);
            AT(c->sxx, j, i) += (dthalbe * sump);
            AT(c->syy, j, i) += (dthalbe * sumq);
            if (mode == 0 && P->grad_form == 2) {
                AT(c->ux, j, i) = sump;
                AT(c->uy, j, i) = sumq;
                AT(c->uxy, j, i) = sumr;
            }
        }
    }
}

/* time derivative of the source signal by central differences, one-sided at
 * the ends (src/psource.c:28-31, src/msource.c:27-30) */
/* ------------------------------------------------------------------------- */
/* acoustic (PHYSICS 2): pressure p (kept in the sxx array) and velocities    */
/* ------------------------------------------------------------------------- */

static real source_amp(const orc_ctx *c, int l, int nt);

/* src/AC/update_v_PML_AC.c:11-676 (FD8 block :328-420; forces :640-656,
 * adjoint forces :660-672): sw 0 forward, 1 adjoint */
static void update_v_ac(orc_ctx *c, int nt, int sw)
{
    const orc_params *P = &c->P;
    const int NX = P->nx, NY = P->ny, FW = P->fw, fdoh = c->fdoh;
    const real DT = (real)P->dt, DH = (real)P->dh;
    const real *hc = c->hc;
    real *const *pm = c->pml;
    const arr p = c->sxx;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= NY; j++) {
        for (int i = 1; i <= NX; i++) {
            real p_x = hc[1] * (AT(p, j, i + 1) - AT(p, j, i));
            real p_y = hc[1] * (AT(p, j + 1, i) - AT(p, j, i));
            for (int m = 2; m <= fdoh; m++) {
                p_x = p_x + hc[m] * (AT(p, j, i + m) - AT(p, j, i - m + 1));
                p_y = p_y + hc[m] * (AT(p, j + m, i) - AT(p, j - m + 1, i));
            }
            if (!P->boundary && i <= FW) {
                real *ps = &PSIX(c->psi_sxx_x, j, i);
                *ps = pm[PML_BXH][i] * *ps + pm[PML_AXH][i] * p_x;
                p_x = p_x / pm[PML_KXH][i] + *ps;
            }
            if (!P->boundary && i >= NX - FW + 1) {
                int h1 = i - NX + 2 * FW;
                real *ps = &PSIX(c->psi_sxx_x, j, h1);
                *ps = pm[PML_BXH][h1] * *ps + pm[PML_AXH][h1] * p_x;
                p_x = p_x / pm[PML_KXH][h1] + *ps;
            }
            if (!P->free_surf && j <= FW) {
                real *ps = &PSIY(c->psi_syy_y, j, i);
                *ps = pm[PML_BYH][j] * *ps + pm[PML_AYH][j] * p_y;
                p_y = p_y / pm[PML_KYH][j] + *ps;
            }
            if (j >= NY - FW + 1) {
                int h1 = j - NY + 2 * FW;
                real *ps = &PSIY(c->psi_syy_y, h1, i);
                *ps = pm[PML_BYH][h1] * *ps + pm[PML_AYH][h1] * p_y;
                p_y = p_y / pm[PML_KYH][h1] + *ps;
            }
            real rip = AT(c->rip, j, i), rjp = AT(c->rjp, j, i);
            if (sw == 0) { /* :390-411, both GRAD_FORMs */
                AT(c->vxp1, j, i) = rip * p_x / DH;
                AT(c->vyp1, j, i) = rjp * p_y / DH;
            } else if (P->grad_form == 1) {
                AT(c->vxp1, j, i) += DT * AT(c->vx, j, i);
                AT(c->vyp1, j, i) += DT * AT(c->vy, j, i);
            } else {
                AT(c->vxp1, j, i) = AT(c->vx, j, i);
                AT(c->vyp1, j, i) = AT(c->vy, j, i);
            }
            AT(c->vx, j, i) += DT * rip * p_x / DH;
            AT(c->vy, j, i) += DT * rjp * p_y / DH;
        }
    }
    if (sw == 1) {
        const int QB = P->quelltypb;
        for (int l = 0; l < c->ntr; l++) {
            int i = c->recpos[l], j = c->recpos[c->ntr + l];
            real sx = c->adj[0][(size_t)l * P->nt + (nt - 1)];
            real sy = c->adj[1][(size_t)l * P->nt + (nt - 1)];
            if (QB == 1) { AT(c->vx, j, i) += sx; AT(c->vy, j, i) += sy; }
            if (QB == 2 || QB == 6 || QB == 7) AT(c->vy, j, i) += sy;
            if (QB == 3 || QB == 5 || QB == 7) AT(c->vx, j, i) += sx;
        }
        return;
    }
    for (int l = 0; l < c->nsrc; l++) {
        int i = (int)c->srcpos[0 * c->nsrc + l], j = (int)c->srcpos[1 * c->nsrc + l];
        real azi = (real)(c->srcpos[6 * c->nsrc + l] * PI_D / 180);
        int type = (int)c->srcpos[7 * c->nsrc + l];
        real s = c->signals[(size_t)l * P->nt + (nt - 1)];
        c->qtype = type; /* update_v_PML_AC.c:646 sets the global QUELLTYP */
        if (type == 2) AT(c->vx, j, i) += s;
        if (type == 3) AT(c->vy, j, i) += s;
        if (type == 4) {
            AT(c->vx, j, i) = (real)(AT(c->vx, j, i) + -sin(azi) * s);
            AT(c->vy, j, i) = (real)(AT(c->vy, j, i) + cos(azi) * s);
        }
    }
}

/* the pressure sweep's derivatives, src/AC/update_s_acoustic_PML_AC.c:253-298:
 * vxx and vyy only, CPML on psi_vxx and psi_vyy */
static void pressure_derivs(orc_ctx *c, int j, int i, real dhi, real *pvxx, real *pvyy)
{
    const orc_params *P = &c->P;
    const int NX = P->nx, NY = P->ny, FW = P->fw, fdoh = c->fdoh;
    const real *hc = c->hc;
    real *const *pm = c->pml;
    real vxx = hc[1] * (AT(c->vx, j, i) - AT(c->vx, j, i - 1));
    real vyy = hc[1] * (AT(c->vy, j, i) - AT(c->vy, j - 1, i));
    for (int m = 2; m <= fdoh; m++) {
        vxx = vxx + hc[m] * (AT(c->vx, j, i + m - 1) - AT(c->vx, j, i - m));
        vyy = vyy + hc[m] * (AT(c->vy, j + m - 1, i) - AT(c->vy, j - m, i));
    }
    vxx = vxx * dhi;
    vyy = vyy * dhi;
    if (!P->boundary && i <= FW) {
        real *a = &PSIX(c->psi_vxx, j, i);
        *a = pm[PML_BX][i] * *a + pm[PML_AX][i] * vxx;
        vxx = vxx / pm[PML_KX][i] + *a;
    }
    if (!P->boundary && i >= NX - FW + 1) {
        int h1 = i - NX + 2 * FW;
        real *a = &PSIX(c->psi_vxx, j, h1);
        *a = pm[PML_BX][h1] * *a + pm[PML_AX][h1] * vxx;
        vxx = vxx / pm[PML_KX][h1] + *a;
    }
    if (!P->free_surf && j <= FW) {
        real *a = &PSIY(c->psi_vyy, j, i);
        *a = pm[PML_BY][j] * *a + pm[PML_AY][j] * vyy;
        vyy = vyy / pm[PML_KY][j] + *a;
    }
    if (j >= NY - FW + 1) {
        int h1 = j - NY + 2 * FW;
        real *a = &PSIY(c->psi_vyy, h1, i);
        *a = pm[PML_BY][h1] * *a + pm[PML_AY][h1] * vyy;
        vyy = vyy / pm[PML_KY][h1] + *a;
    }
    *pvxx = vxx; *pvyy = vyy;
}

/* src/AC/update_s_acoustic_PML_AC.c:13-496 (FD8 block :249-316):
 * p += g (vxx + vyy), g = rho vp^2 (INVMAT1 1) or the modulus (3); a
 * forward run with GRAD_FORM 2 saves ux = g (vxx + vyy) / DT (:309-311) */
static void update_p_ac(orc_ctx *c, int mode)
{
    const orc_params *P = &c->P;
    const real dhi = (real)P->dt / (real)P->dh;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= P->ny; j++) {
        for (int i = 1; i <= P->nx; i++) {
            real vxx, vyy;
            pressure_derivs(c, j, i, dhi, &vxx, &vyy);
            real g = P->invmat1 == 3 ? AT(c->pi, j, i) : AT(c->rho, j, i) * AT(c->pi, j, i) * AT(c->pi, j, i);
            if (mode == 0 && P->grad_form == 2) AT(c->ux, j, i) = g * (vxx + vyy) / (real)P->dt;
            AT(c->sxx, j, i) += g * (vxx + vyy);
        }
    }
}

/* src/AC/psource_AC.c:9-51: sw 0 the explosive source on p, sw 1 the
 * adjoint pressure sources at the receivers (QUELLTYPB >= 4), once each */
static void psource_ac(orc_ctx *c, int nt, int sw)
{
    if (sw == 0) {
        for (int l = 0; l < c->nsrc; l++) {
            int i = (int)c->srcpos[0 * c->nsrc + l], j = (int)c->srcpos[1 * c->nsrc + l];
            AT(c->sxx, j, i) += source_amp(c, l, nt);
        }
    } else if (c->P.quelltypb >= 4) {
        for (int l = 0; l < c->ntr; l++) {
            int i = c->recpos[l], j = c->recpos[c->ntr + l];
            AT(c->sxx, j, i) += c->adj[2][(size_t)l * c->P.nt + (nt - 1)];
        }
    }
}

/* src/AC/surface_acoustic_PML_AC.c:10-27: p = 0 at row 1, ghost rows mirrored */
static void surface_ac(orc_ctx *c)
{
    for (int i = 1; i <= c->P.nx; i++) {
        AT(c->sxx, 1, i) = 0;
        for (int m = 1; m <= c->fdoh; m++) AT(c->sxx, 1 - m, i) = -AT(c->sxx, 1 + m, i);
    }
}

/* src/AC/seismo_AC.c:8-108 (called from src/AC/ac.c:199-201 with u = pi):
 * SEISMO 1 vx, vy; 2 p; 3 div, curl; 4 all five */
static void seismo_ac(orc_ctx *c, int nt)
{
    const orc_params *P = &c->P;
    const real dhi = (real)P->dt / (real)P->dh;
    const real *hc = c->hc;
    for (int itr = 0; itr < c->ntr; itr++) {
        int i = c->recpos[itr], j = c->recpos[c->ntr + itr];
        size_t o = (size_t)itr * P->nt + (nt - 1);
        if (P->seismo == 3 || P->seismo == 4) {
            real vxx = 0, vyy = 0, vyx = 0, vxy = 0;
            for (int m = 1; m <= c->fdoh; m++) {
                vxx += hc[m] * (AT(c->vx, j, i + m - 1) - AT(c->vx, j, i - m));
                vyy += hc[m] * (AT(c->vy, j + m - 1, i) - AT(c->vy, j - m, i));
                vyx += hc[m] * (AT(c->vy, j, i + m) - AT(c->vy, j, i - m + 1));
                vxy += hc[m] * (AT(c->vx, j + m, i) - AT(c->vx, j - m + 1, i));
            }
            vxx *= dhi; vyy *= dhi; vyx *= dhi; vxy *= dhi;
            if (P->invmat1 == 3) {
                c->sec[3][o] = (real)((vxx + vyy) * sqrt(AT(c->pi, j, i)));
                c->sec[4][o] = (real)((vxy - vyx) * sqrt(AT(c->pi, j, i)));
            } else {
                real muss = AT(c->pi, j, i) * AT(c->pi, j, i) * AT(c->rho, j, i);
                real lamss = AT(c->pi, j, i) * AT(c->pi, j, i) * AT(c->rho, j, i);
                c->sec[3][o] = (real)((vxx + vyy) * sqrt(lamss));
                c->sec[4][o] = (real)((vxy - vyx) * sqrt(muss));
            }
        }
        if (P->seismo == 1 || P->seismo == 4) {
            c->sec[0][o] = AT(c->vx, j, i);
            c->sec[1][o] = AT(c->vy, j, i);
        }
        if (P->seismo == 2 || P->seismo == 4) c->sec[2][o] = AT(c->sxx, j, i);
    }
}

/* Model set-up of src/AC/FD_AC.c:184-207: readmod_AC (MFILE.vp, .rho),
 * matcopy_acoustic_AC (one-rank periodic ghosts of rho and vp), av_rho */
orc_ctx *orc_create_ac(const orc_params *p, const float *vp, const float *rho)
{
    if (!p || !vp || !rho) { seterr("null argument"); return NULL; }
    if (p->L != 0) { seterr("acoustic runs have no relaxation mechanisms (L = 0)"); return NULL; }
    orc_ctx *c = ctx_new(p);
    if (!c) return NULL;
    c->physics = 2;
    const int NX = p->nx, NY = p->ny;
    for (int i = 1; i <= NX; i++)
        for (int j = 1; j <= NY; j++) {
            size_t k = (size_t)(i - 1) * NY + (j - 1);
            AT(c->pi, j, i) = vp[k];
            AT(c->rho, j, i) = rho[k];
        }
    matcopy_periodic(c, c->rho);
    matcopy_periodic(c, c->pi);
    av_rho(c);
    return c;
}

/* src/VTI/update_s_elastic_PML_VTI.c:13-637 (FD8 block :304-396): the
 * derivatives and CPML of the isotropic update, then
 * sxy += c44h (vyx + vxy), sxx += c11 vxx + c13 vyy, syy += c13 vxx + c33 vyy;
 * a forward run with GRAD_FORM 2 saves ux, uy, uxy (:384-389) */
static void update_s_vti(orc_ctx *c, int mode)
{
    const orc_params *P = &c->P;
    const real dhi = (real)P->dt / (real)P->dh;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= P->ny; j++) {
        for (int i = 1; i <= P->nx; i++) {
            real vxx, vyy, vxy, vyx;
            stress_derivs(c, j, i, dhi, &vxx, &vyy, &vxy, &vyx);
            if (mode == 0 && P->grad_form == 2) {
                AT(c->ux, j, i) = (AT(c->c11, j, i) * vxx + AT(c->c13, j, i) * vyy) / (real)P->dt;
                AT(c->uy, j, i) = (AT(c->c13, j, i) * vxx + AT(c->c33, j, i) * vyy) / (real)P->dt;
                AT(c->uxy, j, i) = AT(c->c44h, j, i) * (vyx + vxy) / (real)P->dt;
            }
            AT(c->sxy, j, i) += AT(c->c44h, j, i) * (vyx + vxy);
            AT(c->sxx, j, i) += AT(c->c11, j, i) * vxx + AT(c->c13, j, i) * vyy;
            AT(c->syy, j, i) += AT(c->c13, j, i) * vxx + AT(c->c33, j, i) * vyy;
        }
    }
}

/* src/TTI/update_s_elastic_PML_TTI.c:13-635 (FD8 block :304-395) */
static void update_s_tti(orc_ctx *c)
{
    const orc_params *P = &c->P;
    const real dhi = (real)P->dt / (real)P->dh;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= P->ny; j++) {
        for (int i = 1; i <= P->nx; i++) {
            real vxx, vyy, vxy, vyx;
            stress_derivs(c, j, i, dhi, &vxx, &vyy, &vxy, &vyx);
            AT(c->sxx, j, i) += AT(c->d11, j, i) * vxx + AT(c->d13, j, i) * vyy + AT(c->d15h, j, i) * (vyx + vxy);
            AT(c->syy, j, i) += AT(c->d13, j, i) * vxx + AT(c->d33, j, i) * vyy + AT(c->d35h, j, i) * (vyx + vxy);
            AT(c->sxy, j, i) += AT(c->d15, j, i) * vxx + AT(c->d35, j, i) * vyy + AT(c->d55h, j, i) * (vyx + vxy);
        }
    }
}

static real source_amp(const orc_ctx *c, int l, int nt)
{
    const int NT = c->P.nt;
    const real DT = (real)c->P.dt;
    const real *s = &c->signals[(size_t)l * NT]; /* s[n-1] = signals[l][n] */
    real amp = 0.0;
    if (nt == 1) amp = s[nt + 1 - 1] / DT;
    if (nt > 1 && nt < NT) amp = (s[nt + 1 - 1] - s[nt - 1 - 1]) / DT;
    if (nt == NT) amp = -s[nt - 1 - 1] / DT;
    return amp;
}

/* src/psource.c:9-55 */
static void psource(orc_ctx *c, int nt, int sw)
{
    const int NT = c->P.nt;
    if (sw == 0) {
        for (int l = 0; l < c->nsrc; l++) {
            int i = (int)c->srcpos[0 * c->nsrc + l], j = (int)c->srcpos[1 * c->nsrc + l];
            real amp = source_amp(c, l, nt);
            AT(c->sxx, j, i) += amp;
            AT(c->syy, j, i) += amp;
        }
    } else if (c->P.quelltypb >= 4) {
        for (int l = 0; l < c->ntr; l++) {
            int i = c->recpos[l], j = c->recpos[c->ntr + l];
            real s = c->adj[2][(size_t)l * NT + (nt - 1)];
            AT(c->sxx, j, i) += s;
            AT(c->syy, j, i) += s;
        }
    }
}

/* src/msource.c:9-42 (M11=0, M12=1, M22=0) */
static void msource(orc_ctx *c, int nt)
{
    const real M11 = 0.0, M12 = 1.0, M22 = 0.0;
    for (int l = 0; l < c->nsrc; l++) {
        int i = (int)c->srcpos[0 * c->nsrc + l], j = (int)c->srcpos[1 * c->nsrc + l];
        real amp = source_amp(c, l, nt);
        AT(c->sxx, j, i) += amp * M11;
        AT(c->syy, j, i) += amp * M22;
        AT(c->sxy, j, i) = (real)(AT(c->sxy, j, i) + 0.25 * amp * M12);
        AT(c->sxy, j, i - 1) = (real)(AT(c->sxy, j, i - 1) + 0.25 * amp * M12);
        AT(c->sxy, j - 1, i) = (real)(AT(c->sxy, j - 1, i) + 0.25 * amp * M12);
        AT(c->sxy, j - 1, i - 1) = (real)(AT(c->sxy, j - 1, i - 1) + 0.25 * amp * M12);
    }
}

/* x-derivatives at the free surface with the generic stencil and the
 * surface's own CPML memory psi_vxxs (surface_elastic_PML_PSV.c:37-87) */
static void surface_derivs(orc_ctx *c, int i, real dh24, real *pvxx, real *pvyy)
{
    const orc_params *P = &c->P;
    const int j = 1, NX = P->nx, FW = P->fw;
    const real *hc = c->hc;
    real *const *pm = c->pml;
    real vxx = 0.0, vyy = 0.0;
    for (int m = 1; m <= c->fdoh; m++) {
        AT(c->syy, j - m, i) = -AT(c->syy, j + m, i);
        AT(c->sxy, j - m, i) = -AT(c->sxy, j + m - 1, i);
        vxx += hc[m] * (AT(c->vx, j, i + m - 1) - AT(c->vx, j, i - m));
        vyy += hc[m] * (AT(c->vy, j + m - 1, i) - AT(c->vy, j - m, i));
    }
    vxx *= dh24;
    vyy *= dh24;
    if (!P->boundary && i <= FW) {
        real *a = &PSIX(c->psi_vxxs, j, i);
        *a = pm[PML_BX][i] * *a + pm[PML_AX][i] * vxx;
        vxx = vxx / pm[PML_KX][i] + *a;
    }
    if (!P->boundary && i >= NX - FW + 1) {
        int h1 = i - NX + 2 * FW;
        real *a = &PSIX(c->psi_vxxs, j, h1);
        *a = pm[PML_BX][h1] * *a + pm[PML_AX][h1] * vxx;
        vxx = vxx / pm[PML_KX][h1] + *a;
    }
    *pvxx = vxx;
    *pvyy = vyy;
}

/* src/PSV/surface_elastic_PML_PSV.c:10-105 */
static void surface_elastic(orc_ctx *c)
{
    const orc_params *P = &c->P;
    const int j = 1;
    const real DT = (real)P->dt, dh24 = (real)(1.0 / P->dh);
    for (int i = 1; i <= P->nx; i++) {
        AT(c->syy, j, i) = 0.0;
        real vxx, vyy, fjm, g;
        surface_derivs(c, i, dh24, &vxx, &vyy);
        if (P->invmat1 == 3) {
            fjm = (real)(AT(c->u, j, i) * 2.0);
            g = AT(c->pi, j, i);
        } else {
            fjm = (real)(AT(c->rho, j, i) * AT(c->u, j, i) * AT(c->u, j, i) * 2.0);
            g = AT(c->rho, j, i) * ((AT(c->pi, j, i) * AT(c->pi, j, i)) - 2 * AT(c->u, j, i) * AT(c->u, j, i));
        }
        AT(c->sxx, j, i) += -DT * ((g * g) / (g + fjm) * vxx + g * vyy);
    }
}

/* src/PSV/surface_visc_PML_PSV.c:10-173 */
static void surface_visc(orc_ctx *c)
{
    const orc_params *P = &c->P;
    const int j = 1, L = P->L;
    const real DT = (real)P->dt, dh24 = (real)(1.0 / P->dh), dthalbe = (real)(P->dt / 2.0);
    real pts[9];
    for (int l = 1; l <= L; l++) pts[l] = (real)(1.0 / (2.0 * PI_D * P->fl[l - 1]));
    real ws = (real)(2.0 * PI_D * (1.0 / P->ts));
    real sumu = 0.0, sumpi = 0.0;
    for (int l = 1; l <= L; l++) {
        sumu = (real)(sumu + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l])));
        sumpi = (real)(sumpi + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l])));
    }
    for (int i = 1; i <= P->nx; i++) {
        AT(c->syy, j, i) = 0.0;
        for (int l = 1; l <= L; l++) AT(c->q[l], j, i) = 0.0;
        real vxx, vyy, mu, pi;
        surface_derivs(c, i, dh24, &vxx, &vyy);
        real ts = AT(c->taus, j, i), tp = AT(c->taup, j, i);
        if (P->invmat1 == 3) {
            mu = (real)(AT(c->u, j, i) / (1.0 + sumu * ts));
            pi = (real)((AT(c->pi, j, i) + 2 * AT(c->u, j, i)) / (1.0 + sumpi * tp));
        } else {
            mu = (real)((AT(c->u, j, i) * AT(c->u, j, i) * AT(c->rho, j, i)) / (1.0 + sumu * ts));
            pi = (real)((AT(c->pi, j, i) * AT(c->pi, j, i) * AT(c->rho, j, i)) / (1.0 + sumpi * tp));
        }
        real sump = 0.0;
        for (int l = 1; l <= L; l++) sump += AT(c->pp[l], j, i);
        real fjm = (real)(mu * 2.0 * (1.0 + L * ts));
        real g = (real)(pi * (1.0 + L * tp));
        AT(c->sxx, j, i) += -(DT * (g - fjm) * (g - fjm) * vxx / g) - (DT * (g - fjm) * vyy) - (dthalbe * sump);
        sump = 0.0;
        for (int l = 1; l <= L; l++) {
            real etajm = c->peta[l];
            real bjm = (real)(etajm / (1.0 + (etajm * 0.5)));
            real djm = (real)(2.0 * mu * ts);
            real e = pi * tp;
            real *p = &AT(c->pp[l], j, i);
            *p = (real)(*p + bjm * (((djm - e) * ((fjm / g) - 1.0) * vxx) - ((djm - e) * vyy)));
            sump += *p;
        }
        AT(c->sxx, j, i) += (dthalbe * sump);
    }
}

/* src/seismo_ssg.c:8-108 */
static void seismo_ssg(orc_ctx *c, int nt)
{
    const orc_params *P = &c->P;
    const real dhi = (real)P->dt / (real)P->dh;
    const real *hc = c->hc;
    for (int itr = 0; itr < c->ntr; itr++) {
        int i = c->recpos[itr], j = c->recpos[c->ntr + itr];
        size_t o = (size_t)itr * P->nt + (nt - 1);
        if (P->seismo == 1 || P->seismo == 4) {
            c->sec[0][o] = AT(c->vx, j, i);
            c->sec[1][o] = AT(c->vy, j, i);
        }
        if (P->seismo == 2 || P->seismo == 4) c->sec[2][o] = -AT(c->sxx, j, i) - AT(c->syy, j, i);
        if (P->seismo == 3 || P->seismo == 4) {
            real vxx = 0, vyy = 0, vyx = 0, vxy = 0;
            for (int m = 1; m <= c->fdoh; m++) {
                vxx += hc[m] * (AT(c->vx, j, i + m - 1) - AT(c->vx, j, i - m));
                vyy += hc[m] * (AT(c->vy, j + m - 1, i) - AT(c->vy, j - m, i));
                vyx += hc[m] * (AT(c->vy, j, i + m) - AT(c->vy, j, i - m + 1));
                vxy += hc[m] * (AT(c->vx, j + m, i) - AT(c->vx, j - m + 1, i));
            }
            vxx *= dhi; vyy *= dhi; vyx *= dhi; vxy *= dhi;
            if (P->invmat1 == 3) {
                c->sec[3][o] = (real)((vxx + vyy) * sqrt(AT(c->pi, j, i)));
                c->sec[4][o] = (real)((vxy - vyx) * sqrt(AT(c->u, j, i)));
            } else {
                real muss = AT(c->u, j, i) * AT(c->u, j, i) * AT(c->rho, j, i);
                real lamss = AT(c->pi, j, i) * AT(c->pi, j, i) * AT(c->rho, j, i);
                c->sec[3][o] = (real)((vxx + vyy) * sqrt(lamss));
                c->sec[4][o] = (real)((vxy - vyx) * sqrt(muss));
            }
        }
    }
}

/* src/VTI/seismo_ssg_VTI.c:8-108 (also called by src/TTI/TTI.c:212): the
 * div/curl branches are commented out there, so SEISMO 3 samples nothing and
 * SEISMO 4 samples vx, vy and p */
static void seismo_ssg_aniso(orc_ctx *c, int nt)
{
    const orc_params *P = &c->P;
    for (int itr = 0; itr < c->ntr; itr++) {
        int i = c->recpos[itr], j = c->recpos[c->ntr + itr];
        size_t o = (size_t)itr * P->nt + (nt - 1);
        if (P->seismo == 1 || P->seismo == 4) {
            c->sec[0][o] = AT(c->vx, j, i);
            c->sec[1][o] = AT(c->vy, j, i);
        }
        if (P->seismo == 2 || P->seismo == 4) c->sec[2][o] = -AT(c->sxx, j, i) - AT(c->syy, j, i);
    }
}

/* forward snapshot store, src/PSV/psv.c:240-309 (IDXI=IDYI=1) */
static void store_snapshot(orc_ctx *c, int k)
{
    const int NX = c->P.nx, NY = c->P.ny;
    size_t base = (size_t)k * NX * NY;
    size_t need = (size_t)NX * NY * c->ntdtinv;
    for (int f = 0; f < 5; f++)
        if (!c->fwd[f]) c->fwd[f] = (real *)calloc(need, sizeof(real));
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= NX; i++)
        for (int j = 1; j <= NY; j++) {
            size_t o = base + (size_t)(i - 1) * NY + (j - 1);
            c->fwd[0][o] = AT(c->vxp1, j, i);
            c->fwd[1][o] = AT(c->vyp1, j, i);
            if (c->physics == 5) { /* sh.c:150-172: vzp1, and sxz / syz or uzx / uz */
                c->fwd[2][o] = AT(c->ux, j, i);
                c->fwd[3][o] = AT(c->uy, j, i);
                continue;
            }
            if (c->P.grad_form == 1) {
                c->fwd[2][o] = AT(c->sxx, j, i);
                c->fwd[3][o] = AT(c->syy, j, i);
                c->fwd[4][o] = AT(c->sxy, j, i);
            } else {
                c->fwd[2][o] = AT(c->ux, j, i);
                c->fwd[3][o] = AT(c->uy, j, i);
                c->fwd[4][o] = AT(c->uxy, j, i);
            }
        }
}

/* adjoint correlation, src/PSV/psv.c:313-387 (INVMAT1 1/3, GRAD_FORM 1/2) */
static void correlate(orc_ctx *c, int k)
{
    const orc_params *P = &c->P;
    const int NX = P->nx, NY = P->ny;
    size_t base = (size_t)k * NX * NY;
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= NX; i++)
        for (int j = 1; j <= NY; j++) {
            size_t o = base + (size_t)(i - 1) * NY + (j - 1);
            real frx = c->fwd[0][o], fry = c->fwd[1][o], fx = c->fwd[2][o], fy = c->fwd[3][o], fu = c->fwd[4][o];
            real axx = AT(c->sxx, j, i), ayy = AT(c->syy, j, i), axy = AT(c->sxy, j, i);
            if (c->physics == 5) { /* sh.c:186-223: density and shear-modulus terms */
                AT(c->wconv_rho, j, i) += AT(c->vxp1, j, i) * frx;
                real muss = P->invmat1 == 3 ? AT(c->u, j, i) : AT(c->rho, j, i) * AT(c->u, j, i) * AT(c->u, j, i);
                if (muss > 0.0) AT(c->wconv_u, j, i) += (fy * ayy) + (fx * axx);
                continue;
            }
            AT(c->wconv_rho, j, i) += (AT(c->vxp1, j, i) * frx) + (AT(c->vyp1, j, i) * fry);
            if (c->physics == 2) { /* ac.c:260-276: forward p (or ux) times adjoint p */
                AT(c->wconv, j, i) += fx * axx;
                continue;
            }
            AT(c->wconv, j, i) += (fx + fy) * (axx + ayy);
            /* VTI.c:306-331 / TTI.c:304-329: the shear term is commented out */
            if (c->physics == 3 || c->physics == 4) continue;
            real muss = 0, lamss = 0;
            if (P->invmat1 == 1) {
                muss = AT(c->rho, j, i) * AT(c->u, j, i) * AT(c->u, j, i);
                lamss = (real)(AT(c->rho, j, i) * AT(c->pi, j, i) * AT(c->pi, j, i) - 2.0 * muss);
            }
            if (P->invmat1 == 3) {
                muss = AT(c->u, j, i);
                lamss = AT(c->pi, j, i);
            }
            if (muss > 0.0) {
                if (P->grad_form == 1) {
                    AT(c->wconv_u, j, i) = (real)(AT(c->wconv_u, j, i) +
                        (((1.0 / (muss * muss)) * (fu * axy)) +
                         ((1.0 / 4.0) * ((fx + fy) * (axx + ayy)) / ((lamss + muss) * (lamss + muss))) +
                         ((1.0 / 4.0) * ((fx - fy) * (axx - ayy)) / (muss * muss))));
                } else {
                    real tmp = (real)((1.0 / (4.0 * (lamss + muss) * (lamss + muss))) - (1.0 / (4.0 * muss * muss)));
                    real tmp1 = (real)((1.0 / (4.0 * (lamss + muss) * (lamss + muss))) + (1.0 / (4.0 * muss * muss)));
                    AT(c->wconv_u, j, i) = (real)(AT(c->wconv_u, j, i) +
                        (((1.0 / (muss * muss)) * (fu * axy)) + (tmp1 * (fx * axx + fy * ayy)) +
                         (tmp * (fx * ayy + fy * axx))));
                }
            }
        }
}

/* ------------------------------------------------------------------------
 * SH (PHYSICS 5, src/SH): vz is kept in the vx array, sxz in sxx, syz in
 * syy; CPML memory psi_sxz_x in psi_sxx_x, psi_syz_y in psi_syy_y, psi_vzx
 * in psi_vxx, psi_vzy in psi_vyy; 1/rho in rip, the shear moduli in uip, ujp.
 * ---------------------------------------------------------------------- */

/* src/SH/update_v_PML_SH.c:11-602 (FD8 block :294-378, forces :556-568):
 * vz += DT rhoi (sxz_x + syz_y) / DH; QUELLTYP 1 adds the signal to vz.
 * Before the update the gradient field vzp1 (kept in vxp1, :341-372):
 * forward (sw 0) rhoi (sxz_x + syz_y) / DH; adjoint (sw 1) vzp1 + vz DT
 * (GRAD_FORM 1) or vz (GRAD_FORM 2).  The adjoint run adds the vz residual
 * at the receivers for QUELLTYPB 1 (:571-583). */
static void update_v_sh(orc_ctx *c, int nt, int sw)
{
    const orc_params *P = &c->P;
    const int NX = P->nx, NY = P->ny, FW = P->fw, fdoh = c->fdoh;
    const real DT = (real)P->dt, DH = (real)P->dh;
    const real *hc = c->hc;
    real *const *pm = c->pml;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= NY; j++) {
        for (int i = 1; i <= NX; i++) {
            real sxz_x = hc[1] * (AT(c->sxx, j, i) - AT(c->sxx, j, i - 1));
            real syz_y = hc[1] * (AT(c->syy, j, i) - AT(c->syy, j - 1, i));
            for (int m = 2; m <= fdoh; m++) {
                sxz_x = sxz_x + hc[m] * (AT(c->sxx, j, i + m - 1) - AT(c->sxx, j, i - m));
                syz_y = syz_y + hc[m] * (AT(c->syy, j + m - 1, i) - AT(c->syy, j - m, i));
            }
            if (!P->boundary && i <= FW) {
                real *ps = &PSIX(c->psi_sxx_x, j, i);
                *ps = pm[PML_BX][i] * *ps + pm[PML_AX][i] * sxz_x;
                sxz_x = sxz_x / pm[PML_KX][i] + *ps;
            }
            if (!P->boundary && i >= NX - FW + 1) {
                int h1 = i - NX + 2 * FW;
                real *ps = &PSIX(c->psi_sxx_x, j, h1);
                *ps = pm[PML_BX][h1] * *ps + pm[PML_AX][h1] * sxz_x;
                sxz_x = sxz_x / pm[PML_KX][h1] + *ps;
            }
            if (!P->free_surf && j <= FW) {
                real *ps = &PSIY(c->psi_syy_y, j, i);
                *ps = pm[PML_BY][j] * *ps + pm[PML_AY][j] * syz_y;
                syz_y = syz_y / pm[PML_KY][j] + *ps;
            }
            if (j >= NY - FW + 1) {
                int h1 = j - NY + 2 * FW;
                real *ps = &PSIY(c->psi_syy_y, h1, i);
                *ps = pm[PML_BY][h1] * *ps + pm[PML_AY][h1] * syz_y;
                syz_y = syz_y / pm[PML_KY][h1] + *ps;
            }
            if (sw == 0)
                AT(c->vxp1, j, i) = AT(c->rip, j, i) * (sxz_x + syz_y) / DH;
            else if (P->grad_form == 1)
                AT(c->vxp1, j, i) += AT(c->vx, j, i) * DT;
            else
                AT(c->vxp1, j, i) = AT(c->vx, j, i);
            AT(c->vx, j, i) += DT * AT(c->rip, j, i) * (sxz_x + syz_y) / DH;
        }
    }
    if (sw == 0) {
        for (int l = 0; l < c->nsrc; l++) {
            int i = (int)c->srcpos[0 * c->nsrc + l], j = (int)c->srcpos[1 * c->nsrc + l];
            int type = (int)c->srcpos[7 * c->nsrc + l];
            c->qtype = type; /* update_v_PML_SH.c:564 sets the global QUELLTYP */
            if (type == 1) AT(c->vx, j, i) += c->signals[(size_t)l * P->nt + (nt - 1)];
        }
    } else if (P->quelltypb == 1) {
        for (int l = 0; l < c->ntr; l++) {
            int i = c->recpos[l], j = c->recpos[c->ntr + l];
            AT(c->vx, j, i) += c->adj[1][(size_t)l * P->nt + (nt - 1)];
        }
    }
}

/* src/SH/update_s_elastic_PML_SH.c:12-470 (FD8 block :237-306):
 * sxz += uip vzx, syz += ujp vzy.  ADJ_SIGN is 1 in every run: sh.c:92 sets
 * it to -1 for the adjoint run, but update_s_elastic_PML_SH.c:38 sets it back
 * to 1 before each use.  The gradient fields uzx, uz (kept in ux, uy) are the
 * new sxz, syz (GRAD_FORM 1) or vzx, vzy (GRAD_FORM 2), :89-101. */
static void update_s_sh(orc_ctx *c)
{
    const orc_params *P = &c->P;
    const int NX = P->nx, NY = P->ny, FW = P->fw, fdoh = c->fdoh, L = P->L;
    /* update_s_visc_PML_SH.c:34-35: the viscoelastic update takes the
     * derivatives over DH and multiplies by DT inside fipjp / f */
    const real dhi = L ? (real)(1.0 / P->dh) : (real)P->dt / (real)P->dh;
    const real dthalbe = (real)(P->dt / 2.0);
    const real *hc = c->hc;
    real *const *pm = c->pml;
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= NY; j++) {
        for (int i = 1; i <= NX; i++) {
            real vzx = hc[1] * (AT(c->vx, j, i + 1) - AT(c->vx, j, i));
            real vzy = hc[1] * (AT(c->vx, j + 1, i) - AT(c->vx, j, i));
            for (int m = 2; m <= fdoh; m++) {
                vzx = vzx + hc[m] * (AT(c->vx, j, i + m) - AT(c->vx, j, i - m + 1));
                vzy = vzy + hc[m] * (AT(c->vx, j + m, i) - AT(c->vx, j - m + 1, i));
            }
            vzx = vzx * dhi;
            vzy = vzy * dhi;
            if (!P->boundary && i <= FW) {
                real *ps = &PSIX(c->psi_vxx, j, i);
                *ps = pm[PML_BXH][i] * *ps + pm[PML_AXH][i] * vzx;
                vzx = vzx / pm[PML_KXH][i] + *ps;
            }
            if (!P->boundary && i >= NX - FW + 1) {
                int h1 = i - NX + 2 * FW;
                real *ps = &PSIX(c->psi_vxx, j, h1);
                *ps = pm[PML_BXH][h1] * *ps + pm[PML_AXH][h1] * vzx;
                vzx = vzx / pm[PML_KXH][h1] + *ps;
            }
            if (!P->free_surf && j <= FW) {
                real *ps = &PSIY(c->psi_vyy, j, i);
                *ps = pm[PML_BYH][j] * *ps + pm[PML_AYH][j] * vzy;
                vzy = vzy / pm[PML_KYH][j] + *ps;
            }
            if (j >= NY - FW + 1) {
                int h1 = j - NY + 2 * FW;
                real *ps = &PSIY(c->psi_vyy, h1, i);
                *ps = pm[PML_BYH][h1] * *ps + pm[PML_AYH][h1] * vzy;
                vzy = vzy / pm[PML_KYH][h1] + *ps;
            }
            if (L) { /* update_s_visc_PML_SH.c:345-378 (FD8 block) */
                real sumr = 0.0, sumq = 0.0;
                for (int l = 1; l <= L; l++) {
                    sumr += AT(c->r[l], j, i);
                    sumq += AT(c->q[l], j, i);
                }
                AT(c->sxx, j, i) += (AT(c->fipjp, j, i) * vzx) + (dthalbe * sumr);
                AT(c->syy, j, i) += (AT(c->f, j, i) * vzy) + (dthalbe * sumq);
                sumr = sumq = 0.0;
                for (int l = 1; l <= L; l++) {
                    AT(c->r[l], j, i) = c->bip[l] * (AT(c->r[l], j, i) * c->cip[l] - (AT(c->dip[l], j, i) * vzx));
                    AT(c->q[l], j, i) = c->bjm[l] * (AT(c->q[l], j, i) * c->cjm[l] - (AT(c->d[l], j, i) * vzy));
                    sumr += AT(c->r[l], j, i);
                    sumq += AT(c->q[l], j, i);
                }
                AT(c->sxx, j, i) += (dthalbe * sumr);
                AT(c->syy, j, i) += (dthalbe * sumq);
                continue;
            }
            AT(c->sxx, j, i) += AT(c->uip, j, i) * vzx;
            AT(c->syy, j, i) += AT(c->ujp, j, i) * vzy;
            if (P->grad_form == 1) {
                AT(c->uy, j, i) = AT(c->syy, j, i);
                AT(c->ux, j, i) = AT(c->sxx, j, i);
            } else {
                AT(c->uy, j, i) = vzy;
                AT(c->ux, j, i) = vzx;
            }
        }
    }
}

/* src/seismo_ssg.c:8-102 as src/SH/sh.c:133-135 calls it, every section and
 * field argument vz: SEISMO 1 stores vz, which outseis_SHfor.c writes to
 * SEIS_FILE_VY (component 1 here); the other modes write no file */
static void seismo_sh(orc_ctx *c, int nt)
{
    if (c->P.seismo != 1) return;
    for (int itr = 0; itr < c->ntr; itr++) {
        int i = c->recpos[itr], j = c->recpos[c->ntr + itr];
        c->sec[1][(size_t)itr * c->P.nt + (nt - 1)] = AT(c->vx, j, i);
    }
}

/* Model set-up of src/SH/FD_SH.c:143-199: PML_pro_SH, readmod_elastic_SH
 * (MFILE.vs and .rho; INVMAT1 3: .mu), matcopy_elastic_SH (one-rank
 * periodic ghosts), av_mu_SH and inv_rho_SH */
orc_ctx *orc_create_sh(const orc_params *p, const float *vs, const float *rho)
{
    if (p && p->L != 0) { seterr("viscoelastic SH models need taus (orc_create_sh_visc)"); return NULL; }
    return orc_create_sh_visc(p, vs, rho, NULL);
}

/* ... and with L > 0 src/SH/FD_SH.c:171-206: readmod_visc_SH (taus = 2/Qs,
 * eta = DT / (1/(2 pi FL))), matcopy_SH (periodic ghosts of rho, vs, taus),
 * av_mu_SH, av_tau and prepare_update_s_visc_SH.c:8-86 */
orc_ctx *orc_create_sh_visc(const orc_params *p, const float *vs, const float *rho, const float *taus)
{
    if (!p || !vs || !rho || (p->L > 0 && !taus)) { seterr("null argument"); return NULL; }
    if (p->L > 0 && p->mode != 0) { seterr("viscoelastic SH runs are forward modelling (MODE 0)"); return NULL; }
    if (p->seismo > 1) { seterr("SH runs record vz (SEISMO 0 or 1)"); return NULL; }
    orc_ctx *c = ctx_new(p);
    if (!c) return NULL;
    c->physics = 5;
    for (int k = 0; k < PML_N; k++) memset(c->pml[k], 0, sizeof(real) * ((size_t)c->fw2 + 2));
    if (p->fw > 0) pml_profiles(c);
    const int NX = p->nx, NY = p->ny, nd = c->nd;
    c->uip = arr_new(-nd + 1, NY + nd, -nd + 1, NX + nd);
    c->ujp = arr_new(-nd + 1, NY + nd, -nd + 1, NX + nd);
    for (int i = 1; i <= NX; i++)
        for (int j = 1; j <= NY; j++) {
            size_t k = (size_t)(i - 1) * NY + (j - 1);
            AT(c->u, j, i) = vs[k];
            AT(c->rho, j, i) = rho[k];
        }
    matcopy_periodic(c, c->rho);
    matcopy_periodic(c, c->u);
    if (p->L > 0) {
        for (int i = 1; i <= NX; i++)
            for (int j = 1; j <= NY; j++) AT(c->taus, j, i) = taus[(size_t)(i - 1) * NY + (j - 1)];
        matcopy_periodic(c, c->taus);
    }
    for (int j = 1; j <= NY; j++)
        for (int i = 1; i <= NX; i++) {
            real u1, u2, u3;
            if (p->invmat1 == 3) {
                u1 = AT(c->u, j, i); u2 = AT(c->u, j, i + 1); u3 = AT(c->u, j + 1, i);
            } else {
                u1 = AT(c->rho, j, i) * AT(c->u, j, i) * AT(c->u, j, i);
                u2 = AT(c->rho, j, i + 1) * AT(c->u, j, i + 1) * AT(c->u, j, i + 1);
                u3 = AT(c->rho, j + 1, i) * AT(c->u, j + 1, i) * AT(c->u, j + 1, i);
            }
            AT(c->uip, j, i) = (real)(2.0 / ((1.0 / u1) + (1.0 / u2)));
            AT(c->ujp, j, i) = (real)(2.0 / ((1.0 / u1) + (1.0 / u3)));
            if (p->invmat1 == 1 && (u1 == 0.0 || u2 == 0.0)) AT(c->uip, j, i) = 0.0;
            if (p->invmat1 == 1 && (u1 == 0.0 || u3 == 0.0)) AT(c->ujp, j, i) = 0.0;
            AT(c->rip, j, i) = (real)(1.0 / AT(c->rho, j, i));
            if (AT(c->rho, j, i) < 1e-4) AT(c->rip, j, i) = 0.0;
        }
    if (p->L > 0) {
        const int L = p->L;
        const real DT = (real)p->dt;
        av_tau(c);
        real pts[9];
        for (int l = 1; l <= L; l++) {
            pts[l] = (real)(1.0 / (2.0 * PI_D * p->fl[l - 1]));
            c->peta[l] = DT / pts[l];
        }
        real ws = (real)(2.0 * PI_D * p->fl[0]), sumu = 0.0;
        for (int l = 1; l <= L; l++) sumu = (real)(sumu + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l])));
        for (int l = 1; l <= L; l++) {
            c->bip[l] = (real)(1.0 / (1.0 + (c->peta[l] * 0.5)));
            c->bjm[l] = (real)(1.0 / (1.0 + (c->peta[l] * 0.5)));
            c->cip[l] = (real)(1.0 - (c->peta[l] * 0.5));
            c->cjm[l] = (real)(1.0 - (c->peta[l] * 0.5));
        }
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                real ts = AT(c->taus, j, i), tsij = AT(c->tausipjp, j, i);
                real mujp = (real)(AT(c->ujp, j, i) / (1.0 + sumu * ts));
                real muip = (real)(AT(c->uip, j, i) / (1.0 + sumu * tsij));
                AT(c->fipjp, j, i) = (real)(muip * DT * (1.0 + L * tsij));
                AT(c->f, j, i) = (real)(mujp * DT * (1.0 + L * ts));
                for (int l = 1; l <= L; l++) {
                    AT(c->dip[l], j, i) = muip * c->peta[l] * tsij;
                    AT(c->d[l], j, i) = mujp * c->peta[l] * ts;
                }
            }
    }
    return c;
}

/* One time step after the NaN probe: src/PSV/psv.c:104-222 (velocity
 * update, stress update, explosive / adjoint pressure / moment-tensor
 * sources, free surface, receivers).  The VTI and TTI loops
 * (src/VTI/VTI.c:109-215, src/TTI/TTI.c:109-213) call their own stress
 * update and no free-surface routine: with FREE_SURF the top CPML is off
 * and the ghost rows above row 1 stay 0. */
static void step_body(orc_ctx *c, int nt, int mode)
{
    const orc_params *P = &c->P;
    if (c->physics == 5) { /* src/SH/sh.c:80-135 (no free-surface routine) */
        update_v_sh(c, nt, mode == 1 ? 1 : 0);
        update_s_sh(c);
        if (P->seismo && mode != 1) seismo_sh(c, nt);
        return;
    }
    if (c->physics == 2) { /* src/AC/ac.c:96-205 */
        update_v_ac(c, nt, mode == 1 ? 1 : 0);
        update_p_ac(c, mode);
        if (c->qtype == 1 && mode == 0) psource_ac(c, nt, 0);
        if (P->quelltypb >= 4 && mode == 1) psource_ac(c, nt, 1);
        if (P->free_surf) surface_ac(c);
        if (P->seismo && mode == 0) seismo_ac(c, nt);
        return;
    }
    update_v(c, nt, mode == 1 ? 1 : 0);
    if (c->physics == 3) update_s_vti(c, mode);
    else if (c->physics == 4) update_s_tti(c);
    else if (P->L) update_s_visc(c, mode);
    else update_s_elastic(c, mode);
    if (c->qtype == 1 && mode == 0) psource(c, nt, 0);
    if (P->quelltypb >= 4 && mode == 1) psource(c, nt, 1);
    if (c->qtype == 5 && mode == 0) msource(c, nt);
    if (P->free_surf && c->physics == 1) {
        if (P->L) surface_visc(c);
        else surface_elastic(c);
    }
    if (P->seismo && mode == 0) {
        if (c->physics == 1) seismo_ssg(c, nt);
        else seismo_ssg_aniso(c, nt);
    }
}

/* psv() mode 0 / 1, src/PSV/psv.c:17-390 */
static int run_psv(orc_ctx *c, int mode, int nsteps)
{
    const orc_params *P = &c->P;
    const int NT = P->nt;
    if (nsteps <= 0 || nsteps > NT) nsteps = NT;
    zero_wavefields(c);
    int hin = 1, hin1 = 1;
    if (mode == 1) {
        arr_zero(c->wconv); arr_zero(c->wconv_u); arr_zero(c->wconv_rho);
    }
    c->nan_step = 0;
    for (int nt = 1; nt <= nsteps; nt++) {
        /* psv.c:99-103 (vy); sh.c:74-77 (vz, kept in vx) */
        if (isnan((double)AT(c->physics == 5 ? c->vx : c->vy, P->ny / 2, P->nx / 2))) {
            c->nan_step = nt;
            seterr("Simulation is unstable at time step %d", nt);
            return -2;
        }
        step_body(c, nt, mode);
        if (nt == hin1 && mode == 0 && P->mode > 0) {
            store_snapshot(c, hin - 1);
            hin++;
            hin1 += P->dtinv;
            c->dtinv_help[nt] = 1;
        }
        if (mode == 1 && c->dtinv_help[NT - nt + 1] == 1) {
            correlate(c, c->ntdtinv - hin);
            hin++;
        }
    }
    return 0;
}

int orc_forward(orc_ctx *c, int nsteps) { return run_psv(c, 0, nsteps); }

/* Single time step nt of psv() on the current state (no zeroing, no store or
 * correlation): velocity sweep, stress sweep, sources, free surface, seismo.
 * Used by the tests that pin individual operators. */
int orc_step(orc_ctx *c, int nt, int mode)
{
    const orc_params *P = &c->P;
    if (nt < 1 || nt > P->nt) { seterr("step %d out of range", nt); return -1; }
    step_body(c, nt, mode);
    return 0;
}

int orc_zero(orc_ctx *c)
{
    zero_wavefields(c);
    return 0;
}
int orc_adjoint(orc_ctx *c, int nsteps) { return run_psv(c, 1, nsteps); }

/* L2 residual (LNORM=2, GRAD_FORM=1, MODE=1) of src/calc_res.c:296-452,
 * with the observed-trace reading of src/inseis.c:96-103 (sample 1 is zero
 * on input and replaced by the synthetic at calc_res.c:419) and the misfit
 * sum :553-556.  obs arrays hold the SU samples 1..NT per trace. */
static double calc_res_comp(orc_ctx *c, const real *syn, const float *obs, real *diff, double L2, int testshot,
                            float w)
{
    const int ns = c->P.nt;
    const real DT = (real)c->P.dt;
    real *data = (real *)malloc(sizeof(real) * (size_t)ns);
    real *sw = (real *)malloc(sizeof(real) * (size_t)ns);
    for (int i = 0; i < c->ntr; i++) {
        /* calc_res.c:100-123: with COMP_WEIGHT the synthetic and observed
         * samples are multiplied by tmp = C_vp*C_rho (float) first */
        const real tmp = (real)w;
        const real *s = sw;
        for (int j = 0; j < ns; j++) sw[j] = syn[(size_t)i * ns + j] * tmp;
        real *df = &diff[(size_t)i * ns];
        data[0] = 0.0;
        for (int j = 1; j < ns; j++) data[j] = obs[(size_t)i * ns + j] * tmp;
        for (int j = 0; j < ns; j++) df[j] = 0.0;
        real intseis = 0.0;
        int invtime = ns;
        for (int j = 1; j <= ns; j++) {
            if (j == 1) data[0] = s[0];
            intseis += DT * (s[j - 1] - data[j - 1]);
            df[invtime - 1] = intseis;
            invtime--;
        }
        invtime = ns;
        for (int j = 1; j <= ns; j++) {
            if (testshot) L2 += 0.5 * df[invtime - 1] * df[invtime - 1];
            invtime--;
        }
    }
    free(data);
    free(sw);
    return L2;
}

/* COMP_WEIGHT of src/calc_res.c:100-123 for the vx and vy residuals */
int orc_set_res_weights(orc_ctx *c, float w_vx, float w_vy)
{
    c->resw[0] = w_vx;
    c->resw[1] = w_vy;
    return 0;
}

/* src/PSV/calc_res_PSV.c:8-184 (vx, vy, p branches by QUELLTYPB) */
double orc_calc_res(orc_ctx *c, const float *obs_vx, const float *obs_vy, const float *obs_p, int testshot)
{
    const int QB = c->P.quelltypb;
    double L2 = 0.0;
    const float wx = c->resw[0] != 0.0f ? c->resw[0] : 1.0f, wy = c->resw[1] != 0.0f ? c->resw[1] : 1.0f;
    if (c->physics == 5) { /* src/SH/calc_res_SH.c:20-44: vz (component 1) for any QUELLTYPB, no weights */
        if (QB && obs_vy) L2 = calc_res_comp(c, c->sec[1], obs_vy, c->adj[1], L2, testshot, 1.0f);
        return L2;
    }
    if ((QB == 1 || QB == 3 || QB == 5 || QB == 7) && obs_vx)
        L2 = calc_res_comp(c, c->sec[0], obs_vx, c->adj[0], L2, testshot, wx);
    if ((QB == 1 || QB == 2 || QB == 6 || QB == 7) && obs_vy)
        L2 = calc_res_comp(c, c->sec[1], obs_vy, c->adj[1], L2, testshot, wy);
    if (QB >= 4 && obs_p) L2 = calc_res_comp(c, c->sec[2], obs_p, c->adj[2], L2, testshot, 1.0f);
    return L2;
}

int orc_set_adjoint(orc_ctx *c, const float *dvx, const float *dvy, const float *dp)
{
    size_t n = (size_t)c->ntr * c->P.nt;
    const float *src[3] = {dvx, dvy, dp};
    for (int k = 0; k < 3; k++)
        for (size_t t = 0; t < n; t++) c->adj[k][t] = src[k] ? src[k][t] : 0.0f;
    return 0;
}

/* src/PSV/ass_gradPSV.c:10-155 (INVMAT1 1/3, iter >= INV_*_ITER) */
int orc_assemble(orc_ctx *c)
{
    const orc_params *P = &c->P;
    const real DT = (real)P->dt;
    if (c->physics == 5) { /* src/SH/ass_gradSH.c:10-98 with INV_*_ITER 0: gvs, grho */
        for (int i = 1; i <= P->nx; i++)
            for (int j = 1; j <= P->ny; j++) {
                real pu = AT(c->u, j, i), prho = AT(c->rho, j, i);
                AT(c->wmu, j, i) = -DT * AT(c->wconv_u, j, i);
                real gu = AT(c->wconv_u, j, i); /* kept where ass_gradSH.c writes nothing */
                if (P->invmat1 == 1) {
                    if (P->grad_form == 1) {
                        if (pu > 0.0) {
                            real vs3 = pu * pu * pu;
                            gu = (real)((2.0 / prho * vs3) * AT(c->wmu, j, i));
                        }
                    } else {
                        gu = (real)(2.0 * prho * pu * AT(c->wmu, j, i));
                    }
                } else {
                    gu = AT(c->wmu, j, i);
                }
                AT(c->gvs, j, i) = gu;
                AT(c->gvp, j, i) = 0;
            }
        for (int i = 1; i <= P->nx; i++)
            for (int j = 1; j <= P->ny; j++) {
                real pu = AT(c->u, j, i), prho = AT(c->rho, j, i);
                AT(c->wrho_s, j, i) = -DT * AT(c->wconv_rho, j, i);
                real gr = AT(c->wconv_rho, j, i);
                if (P->invmat1 == 1) {
                    if (P->grad_form == 1) {
                        if (pu > 0.0) {
                            real vs2rho2 = (real)(1.0 / (pu * pu * prho * prho));
                            gr = (vs2rho2 * AT(c->wmu, j, i)) + AT(c->wrho_s, j, i);
                        }
                    } else {
                        gr = (pu * pu * AT(c->wmu, j, i)) + AT(c->wrho_s, j, i);
                    }
                } else {
                    gr = AT(c->wrho_s, j, i);
                }
                AT(c->grho, j, i) = gr;
            }
        return 0;
    }
    if (c->physics == 2) { /* src/AC/ass_gradAC.c:12-93 with INV_*_ITER 0 */
        for (int i = 1; i <= P->nx; i++)
            for (int j = 1; j <= P->ny; j++) {
                AT(c->wlam, j, i) = -DT * AT(c->wconv, j, i);
                real ppi = AT(c->pi, j, i), prho = AT(c->rho, j, i);
                real lamss = P->invmat1 == 3 ? ppi : prho * ppi * ppi;
                if (lamss > 0.0) {
                    if (P->grad_form == 1)
                        AT(c->wlam, j, i) = (real)((1.0 / (4.0 * lamss * lamss)) * AT(c->wlam, j, i));
                    else
                        AT(c->wlam, j, i) = (real)(-(1.0 / (lamss * lamss)) * AT(c->wlam, j, i));
                }
                AT(c->gvp, j, i) = P->invmat1 == 3 ? AT(c->wlam, j, i) : (real)(2.0 * ppi * prho * AT(c->wlam, j, i));
                AT(c->gvs, j, i) = 0;
                AT(c->wrho_s, j, i) = -DT * AT(c->wconv_rho, j, i);
                AT(c->grho, j, i) = P->invmat1 == 1 ? ((ppi * ppi) * AT(c->wlam, j, i)) + AT(c->wrho_s, j, i)
                                                    : AT(c->wrho_s, j, i);
            }
        return 0;
    }
    for (int i = 1; i <= P->nx; i++)
        for (int j = 1; j <= P->ny; j++) {
            AT(c->wlam, j, i) = -DT * AT(c->wconv, j, i);
            real ppi = AT(c->pi, j, i), pu = AT(c->u, j, i), prho = AT(c->rho, j, i);
            if (P->invmat1 == 3) {
                real muss = pu, lamss = ppi;
                if (P->grad_form == 1 && (muss > 0.0 || lamss > 0.0))
                    AT(c->wlam, j, i) = (real)((1.0 / (4.0 * (lamss + muss) * (lamss + muss))) * AT(c->wlam, j, i));
                AT(c->gvp, j, i) = AT(c->wlam, j, i);
            } else {
                real muss = prho * pu * pu;
                real lamss = (real)(prho * ppi * ppi - 2.0 * muss);
                if (muss > 0.0 || lamss > 0.0)
                    AT(c->wlam, j, i) = (real)((1.0 / (4.0 * (lamss + muss) * (lamss + muss))) * AT(c->wlam, j, i));
                AT(c->gvp, j, i) = (real)(2.0 * ppi * prho * AT(c->wlam, j, i));
            }
        }
    for (int i = 1; i <= P->nx; i++)
        for (int j = 1; j <= P->ny; j++) {
            AT(c->wmu, j, i) = -DT * AT(c->wconv_u, j, i);
            real pu = AT(c->u, j, i), prho = AT(c->rho, j, i);
            if (P->invmat1 == 1)
                AT(c->gvs, j, i) = (real)((-4.0 * prho * pu * AT(c->wlam, j, i)) + 2.0 * prho * pu * AT(c->wmu, j, i));
            else
                AT(c->gvs, j, i) = AT(c->wmu, j, i);
        }
    for (int i = 1; i <= P->nx; i++)
        for (int j = 1; j <= P->ny; j++) {
            AT(c->wrho_s, j, i) = -DT * AT(c->wconv_rho, j, i);
            real ppi = AT(c->pi, j, i), pu = AT(c->u, j, i);
            if (P->invmat1 == 1)
                AT(c->grho, j, i) = (real)(((((ppi * ppi) - (2.0 * pu * pu)) * AT(c->wlam, j, i)) +
                                             (pu * pu * AT(c->wmu, j, i)) + AT(c->wrho_s, j, i)));
            else
                AT(c->grho, j, i) = AT(c->wrho_s, j, i);
        }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* accessors                                                                  */
/* ------------------------------------------------------------------------- */
int orc_dims(const orc_ctx *c, int *nd, int *ntdtinv)
{
    if (nd) *nd = c->nd;
    if (ntdtinv) *ntdtinv = c->ntdtinv;
    return 0;
}

static const arr *field_ptr(const orc_ctx *c, int id)
{
    switch (id) {
    case ORC_VX: return &c->vx;
    case ORC_VY: return &c->vy;
    case ORC_SXX: return &c->sxx;
    case ORC_SYY: return &c->syy;
    case ORC_SXY: return &c->sxy;
    case ORC_VXP1: return &c->vxp1;
    case ORC_VYP1: return &c->vyp1;
    case ORC_UX: return &c->ux;
    case ORC_UY: return &c->uy;
    case ORC_UXY: return &c->uxy;
    default: return NULL;
    }
}

int orc_set_field(orc_ctx *c, int id, const void *in)
{
    const arr *a = field_ptr(c, id);
    if (!a) { seterr("bad field id %d", id); return -1; }
    memcpy(a->p, in, arr_len(*a) * sizeof(real));
    return 0;
}

int orc_get_field(const orc_ctx *c, int id, void *out)
{
    const arr *a = NULL;
    switch (id) {
    case ORC_VX: a = &c->vx; break;
    case ORC_VY: a = &c->vy; break;
    case ORC_SXX: a = &c->sxx; break;
    case ORC_SYY: a = &c->syy; break;
    case ORC_SXY: a = &c->sxy; break;
    case ORC_VXP1: a = &c->vxp1; break;
    case ORC_VYP1: a = &c->vyp1; break;
    case ORC_UX: a = &c->ux; break;
    case ORC_UY: a = &c->uy; break;
    case ORC_UXY: a = &c->uxy; break;
    case ORC_RIP: a = &c->rip; break;
    case ORC_RJP: a = &c->rjp; break;
    case ORC_UIPJP: a = &c->uipjp; break;
    case ORC_PRHO: a = &c->rho; break;
    case ORC_PPI: a = &c->pi; break;
    case ORC_PU: a = &c->u; break;
    case ORC_FIPJP: a = &c->fipjp; break;
    case ORC_F: a = &c->f; break;
    case ORC_G: a = &c->g; break;
    case ORC_C11: a = &c->c11; break;
    case ORC_C13: a = &c->c13; break;
    case ORC_C33: a = &c->c33; break;
    case ORC_C44H: a = &c->c44h; break;
    case ORC_D11: a = &c->d11; break;
    case ORC_D13: a = &c->d13; break;
    case ORC_D15: a = &c->d15; break;
    case ORC_D15H: a = &c->d15h; break;
    case ORC_D33: a = &c->d33; break;
    case ORC_D35: a = &c->d35; break;
    case ORC_D35H: a = &c->d35h; break;
    case ORC_D55H: a = &c->d55h; break;
    case ORC_UIP: a = &c->uip; break;
    case ORC_UJP: a = &c->ujp; break;
    default: seterr("bad field id %d", id); return -1;
    }
    memcpy(out, a->p, arr_len(*a) * sizeof(real));
    return 0;
}

/* CPML memory arrays: 0 psi_sxx_x 1 psi_sxy_x 2 psi_vxx 3 psi_vyx 4 psi_vxxs
 * ([NY][2FW]); 5 psi_syy_y 6 psi_sxy_y 7 psi_vyy 8 psi_vxy ([2FW][NX]) */
int orc_get_psi(const orc_ctx *c, int id, void *out)
{
    const real *src[9] = {c->psi_sxx_x, c->psi_sxy_x, c->psi_vxx, c->psi_vyx, c->psi_vxxs,
                          c->psi_syy_y, c->psi_sxy_y, c->psi_vyy, c->psi_vxy};
    if (id < 0 || id > 8) return -1;
    size_t n = id < 5 ? (size_t)c->P.ny * c->fw2 : (size_t)c->fw2 * c->P.nx;
    memcpy(out, src[id], n * sizeof(real));
    return 0;
}

/* memory variables: which 0 r, 1 p, 2 q; l = 1..L; padded layout */
int orc_get_mem(const orc_ctx *c, int which, int l, void *out)
{
    if (l < 1 || l > c->P.L || which < 0 || which > 2) return -1;
    const arr *a = which == 0 ? &c->r[l] : which == 1 ? &c->pp[l] : &c->q[l];
    memcpy(out, a->p, arr_len(*a) * sizeof(real));
    return 0;
}

int orc_get_seis(const orc_ctx *c, int comp, void *out)
{
    if (comp < 0 || comp > 4) return -1;
    memcpy(out, c->sec[comp], (size_t)c->ntr * c->P.nt * sizeof(real));
    return 0;
}

int orc_get_adjsrc(const orc_ctx *c, int comp, void *out)
{
    if (comp < 0 || comp > 2) return -1;
    memcpy(out, c->adj[comp], (size_t)c->ntr * c->P.nt * sizeof(real));
    return 0;
}

/* which: 0..2 raw per-shot accumulators (waveconv_shot, waveconv_u_shot,
 * waveconv_rho_shot), 3..5 assembled vp / vs / rho gradients.  Output is
 * x-major / y-fastest NX*NY (gradient-file order). */
int orc_get_grad(const orc_ctx *c, int which, void *out)
{
    const arr *src[6] = {&c->wconv, &c->wconv_u, &c->wconv_rho, &c->gvp, &c->gvs, &c->grho};
    if (which < 0 || which > 5) return -1;
    real *o = (real *)out;
    for (int i = 1; i <= c->P.nx; i++)
        for (int j = 1; j <= c->P.ny; j++) o[(size_t)(i - 1) * c->P.ny + (j - 1)] = AT(*src[which], j, i);
    return 0;
}

int orc_get_pml(const orc_ctx *c, int which, void *out)
{
    if (which < 0 || which >= PML_N) return -1;
    memcpy(out, c->pml[which] + 1, (size_t)c->fw2 * sizeof(real));
    return 0;
}

int orc_get_hc(const orc_ctx *c, void *out)
{
    memcpy(out, c->hc, 7 * sizeof(real));
    return 0;
}
