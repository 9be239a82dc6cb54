/*
 * denise_hip -- drop-in executable for DENISE's P/SV forward modelling and
 * FWI gradient on one MI355X:
 *
 *     denise_hip <model.inp> [<workflow.inp>]      (src/denise.c:60-88)
 *
 * MODE = 0: per shot, the forward run of src/PSV/FD_PSV.c:271-354 and the SU
 *           files of saveseis_glob (SEIS_FILE_*.shot<N>).
 * MODE = 1: per shot, the gradient of src/PSV/grad_obj_psv.c:130-262
 *           (forward with snapshot store, L2 residual against
 *           DATA_DIR_{x,y,p}.su.shot<N>, adjoint, ass_gradPSV), stacked over
 *           shots; writes the stacked vp/vs/rho gradients as JACOBIAN_grad_vp,
 *           JACOBIAN_grad_vs, JACOBIAN_grad_rho (float32, x-major, NX x NY)
 *           and, with RTM_SHOT = 1, the per-shot images JACOBIAN_P_image_shot_N
 *           and JACOBIAN_S_image_shot_N (src/PSV/RTM_PSV_out_shot.c).  The
 *           misfit is the L2 sum over the test shots.
 * Exit status 1 with a message on any error, as the reference's err().
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "denise_hip.h"
#include "denise_io.h"

static void die(const char *what, const char *msg)
{
    fprintf(stderr, "\n denise_hip: %s: %s\n", what, msg ? msg : "");
    exit(1);
}

static void io_ok(int rc, const char *what)
{
    if (rc) die(what, dnio_last_error());
}

static void dn_ok(int rc, const char *what)
{
    if (rc) die(what, dn_last_error());
}

static void *xmalloc(size_t n)
{
    void *p = calloc(1, n ? n : 1);
    if (!p) die("allocation", "out of host memory");
    return p;
}

/* what of the reference this build does not run; refuse instead of
 * silently computing something else */
static void check_supported(const dnio_par *p, const dnio_stage *st)
{
    if (p->physics != 1) die("PHYSICS", "only 2D P/SV (PHYSICS = 1) is built");
    if (p->mode != 0 && p->mode != 1) die("MODE", "MODE 0 (forward) and 1 (FWI gradient) are built; RTM is not");
    if (p->boundary != 0) die("BOUNDARY", "periodic boundaries are not supported");
    if (!p->srcrec) die("SRCREC", "sources must come from SOURCE_FILE (SRCREC = 1)");
    if (p->quellart == 6) die("QUELLART", "the spike source needs the Butterworth source filter, not built");
    if (p->snap) fprintf(stderr, " denise_hip: SNAP > 0 ignored (no snapshot output)\n");
    if (p->mode == 1) {
        if (!st) die("workflow", "MODE 1 needs a workflow file");
        if (st->lnorm != 2) die("LNORM", "only the L2 misfit (LNORM = 2) is built");
        if (st->time_filt) die("TIME_FILT", "time-domain filtering is not built");
        if (st->eprecond) die("EPRECOND", "energy preconditioning is not built");
        if (st->inv_stf || st->timewin || st->normalize || st->offset_mute || st->env || st->rowi)
            die("workflow", "STF inversion, time windows, normalisation, offset mutes, envelope and ROWI are not built");
        if (p->ndt != 1) die("NDT", "FWI needs NDT = 1");
        if (p->sws_taper_circular_per_shot) die("SWS_TAPER_CIRCULAR_PER_SHOT", "per-shot gradient tapers are not built");
        if (p->grad_form != 1 && p->grad_form != 2) die("GRAD_FORM", "1 or 2 expected");
    }
}

static void write_plane(const char *name, const float *a, int nx, int ny)
{
    io_ok(dnio_write_model(name, a, nx, ny), name);
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s <model.inp> [<workflow.inp>]\n", argv[0]);
        return 1;
    }
    dnio_par P;
    io_ok(dnio_read_par(argv[1], &P), "parameter file");
    dnio_stage stage, *st = NULL;
    if (P.mode == 1) {
        if (argc < 3) die("workflow", "MODE 1 needs a workflow file as second argument");
        io_ok(dnio_read_stage(argv[2], 1, &stage), "workflow file");
        st = &stage;
    }
    check_supported(&P, st);
    if (P.nprocx * P.nprocy > 1)
        printf(" NPROCX x NPROCY = %d x %d: the whole grid runs as one domain on one GPU\n", P.nprocx, P.nprocy);

    int nsrc = 0;
    float *srcpos = NULL, TS = P.ts;
    io_ok(dnio_read_sources(P.source_file, &nsrc, &srcpos, &TS), "SOURCE_FILE");
    const int nshots = P.run_multiple_shots ? nsrc : 1;
    printf(" %d source(s), %d shot(s); NX=%d NY=%d NT=%d DH=%g DT=%g FDORDER=%d L=%d\n", nsrc, nshots, P.nx, P.ny,
           P.nt, P.dh, P.dt, P.fdorder, P.L);

    const size_t ncell = (size_t)P.nx * P.ny;
    float *vp = xmalloc(sizeof(float) * ncell), *vs = xmalloc(sizeof(float) * ncell);
    float *rho = xmalloc(sizeof(float) * ncell);
    float *taus = P.L ? xmalloc(sizeof(float) * ncell) : NULL, *taup = P.L ? xmalloc(sizeof(float) * ncell) : NULL;
    io_ok(dnio_read_model(&P, vp, vs, rho, taus, taup), "model");

    dn_params dp;
    memset(&dp, 0, sizeof dp);
    dp.nx = P.nx;
    dp.ny = P.ny;
    dp.nt = P.nt;
    dp.fdorder = P.fdorder;
    dp.maxrelerror = P.maxrelerror;
    dp.fw = P.fw;
    dp.free_surf = P.free_surf;
    dp.boundary = P.boundary;
    dp.L = P.L;
    dp.seismo = P.mode == 1 && P.seismo == 0 ? 1 : P.seismo;
    dp.ndt = P.ndt;
    dp.invmat1 = P.invmat1;
    dp.grad_form = P.grad_form;
    dp.dtinv = P.dtinv > 0 ? P.dtinv : 1;
    dp.quelltypb = P.quelltypb;
    dp.mode = P.mode;
    dp.dh = P.dh;
    dp.dt = P.dt;
    dp.damping = P.damping;
    dp.fpml = P.fpml;
    dp.npower = P.npower;
    dp.k_max_pml = P.k_max_pml;
    dp.ts = TS;
    for (int l = 0; l < P.L && l < 8; l++) dp.fl[l] = P.fl[l];

    dn_ctx *ctx = NULL;
    dn_ok(dn_ctx_create(&dp, 0, &ctx), "device context");
    dn_ok(dn_set_model(ctx, vp, vs, rho, taus, taup), "model upload");

    int ntr = 0, *recpos = NULL;
    if (P.seismo || P.mode == 1) io_ok(dnio_read_receivers(&P, 1, &ntr, &recpos), "REC_FILE");

    float *g_sum[3] = {NULL, NULL, NULL}, *g_shot = NULL;
    double misfit = 0.0;
    if (P.mode == 1) {
        for (int k = 0; k < 3; k++) g_sum[k] = xmalloc(sizeof(float) * ncell);
        g_shot = xmalloc(sizeof(float) * ncell);
    }
    const int n_order = st ? st->n_order : 0;
    char name[4 * DNIO_STR];
    for (int ishot = 1; ishot <= nshots; ishot++) {
        if (P.readrec == 2 && ishot > 1) {
            dnio_free(recpos);
            io_ok(dnio_read_receivers(&P, ishot, &ntr, &recpos), "REC_FILE");
        }
        const int first = P.run_multiple_shots ? ishot - 1 : 0, count = P.run_multiple_shots ? 1 : nsrc;
        float *loc = xmalloc(sizeof(float) * 8 * (size_t)count);
        int nloc = 0;
        io_ok(dnio_split_sources(&P, srcpos, nsrc, first, count, loc, &nloc), "sources");
        /* pack the kept columns into an 8 x nloc table */
        float *sp = xmalloc(sizeof(float) * 8 * (size_t)(nloc ? nloc : 1));
        for (int r = 0; r < 8; r++)
            for (int l = 0; l < nloc; l++) sp[r * nloc + l] = loc[r * count + l];
        float *sig = xmalloc(sizeof(float) * (size_t)(nloc ? nloc : 1) * P.nt);
        io_ok(dnio_wavelet(&P, sp, nloc, ishot, TS, n_order, sig), "wavelet");
        dn_ok(dn_set_sources(ctx, nloc, sp, sig), "sources");
        int *rp = xmalloc(sizeof(int) * 2 * (size_t)(ntr ? ntr : 1));
        for (int k = 0; k < ntr; k++) {
            rp[k] = recpos[k];
            rp[ntr + k] = recpos[ntr + k];
        }
        dn_ok(dn_set_receivers(ctx, ntr, rp), "receivers");
        const float XS = srcpos[0 * nsrc + ishot - 1], YS = srcpos[1 * nsrc + ishot - 1];
        printf(" shot %d: %d source point(s), %d receivers\n", ishot, nloc, ntr);
        fflush(stdout);
        dn_ok(dn_run_forward(ctx, 0), "forward run");

        if (P.mode == 0 && ntr > 0 && P.seismo) {
            /* saveseis_glob: SEISMO 1 vx,vy; 2 p; 3 div,curl; 4 vx,vy,p */
            int comps[5], nc = 0;
            if (P.seismo == 1 || P.seismo == 4) comps[nc++] = 0, comps[nc++] = 1;
            if (P.seismo == 2 || P.seismo == 4) comps[nc++] = 2;
            if (P.seismo == 3) comps[nc++] = 3, comps[nc++] = 4;
            float *sec = xmalloc(sizeof(float) * (size_t)ntr * P.nt);
            for (int c = 0; c < nc; c++) {
                dn_ok(dn_get_seismogram(ctx, comps[c], sec), "seismograms");
                io_ok(dnio_seis_name(&P, comps[c], ishot, 1, name, sizeof name), "file name");
                io_ok(dnio_write_seis(name, &P, sec, ntr, P.nt, recpos, XS, YS), name);
                printf("   wrote %s\n", name);
            }
            free(sec);
        }
        if (P.mode == 1) {
            float *obs[3] = {NULL, NULL, NULL};
            const int qb = P.quelltypb;
            const int use[3] = {qb == 1 || qb == 3 || qb == 5 || qb == 7, qb == 1 || qb == 2 || qb == 6 || qb == 7,
                                qb >= 4};
            for (int c = 0; c < 3; c++)
                if (use[c]) {
                    obs[c] = xmalloc(sizeof(float) * (size_t)ntr * P.nt);
                    io_ok(dnio_data_name(&P, c, ishot, name, sizeof name), "file name");
                    io_ok(dnio_read_su(name, ntr, P.nt, obs[c]), name);
                }
            const int test = ishot >= P.testshot_start && ishot <= P.testshot_end && P.testshot_incr > 0 &&
                             (ishot - P.testshot_start) % P.testshot_incr == 0;
            double l2 = 0.0;
            dn_ok(dn_calc_residuals(ctx, obs[0], obs[1], obs[2], test, &l2), "residuals");
            misfit += l2;
            for (int c = 0; c < 3; c++) free(obs[c]);
            dn_ok(dn_run_adjoint(ctx, 0), "adjoint run");
            dn_ok(dn_assemble_gradient(ctx), "gradient");
            for (int k = 0; k < 3; k++) {
                dn_ok(dn_get_gradient(ctx, 3 + k, g_shot), "gradient");
                for (size_t q = 0; q < ncell; q++) g_sum[k][q] += g_shot[q];
                if (P.rtm_shot && k < 2) {
                    snprintf(name, sizeof name, "%s_%s_image_shot_%i", P.jacobian, k == 0 ? "P" : "S", ishot);
                    write_plane(name, g_shot, P.nx, P.ny);
                }
            }
            printf("   L2 of shot %d: %.9e%s\n", ishot, l2, test ? "" : " (not a test shot)");
        }
        free(loc);
        free(sp);
        free(sig);
        free(rp);
    }
    if (P.mode == 1) {
        const char *suffix[3] = {"vp", "vs", "rho"};
        for (int k = 0; k < 3; k++) {
            snprintf(name, sizeof name, "%s_grad_%s", P.jacobian, suffix[k]);
            write_plane(name, g_sum[k], P.nx, P.ny);
            printf(" wrote %s\n", name);
            free(g_sum[k]);
        }
        free(g_shot);
        printf(" misfit (L2, test shots): %.9e\n", misfit);
    }
    dn_ctx_destroy(ctx);
    dnio_free(recpos);
    dnio_free(srcpos);
    free(vp);
    free(vs);
    free(rho);
    free(taus);
    free(taup);
    return 0;
}
