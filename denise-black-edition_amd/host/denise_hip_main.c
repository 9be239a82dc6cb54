/*
 * denise_hip -- drop-in executable for DENISE's P/SV forward modelling and
 * full-waveform inversion on one MI355X:
 *
 *     denise_hip <model.inp> [<workflow.inp>]      (src/denise.c:60-88)
 *
 * MODE = 0: per shot, the forward run of src/PSV/FD_PSV.c:271-354 and the SU
 *           files of saveseis_glob (SEIS_FILE_*.shot<N>); with SNAP > 0 shot
 *           SNAP_SHOT only, with the wavefield snapshots of src/snap_ssg.c
 *           (SNAP_FILE.{bin,asc}.{x,y,p,div,rot}.0.0).
 * MODE = 1: the FWI workflow of src/PSV/FWI_PSV.c:381-836 with L-BFGS
 *           (GRAD_METHOD = 2) or PCG (1): for each stage of the workflow file and each
 *           iteration up to ITERMAX,
 *           - the gradient of src/PSV/grad_obj_psv.c:130-262 per shot
 *             (forward with snapshot store, L2 residual against
 *             DATA_DIR_{x,y,p}.su.shot<N>, adjoint, ass_gradPSV), stacked
 *             over shots, written after preconditioning as JACOBIAN_grad_vp,
 *             _vs, _rho (float32, x-major, NX x NY; files the reference does
 *             not write, kept for inspection) and, with RTM_SHOT = 1, the
 *             per-shot images JACOBIAN_P_image_shot_N and
 *             JACOBIAN_S_image_shot_N (src/PSV/RTM_PSV_out_shot.c); the
 *             misfit is the L2 sum over the test shots;
 *           - the L-BFGS or PCG direction; with L-BFGS the files of
 *             src/PSV/extract_LBFGS_PSV.c:80-249 (JACOBIAN_p_vp.old, _p.old,
 *             _c.old, _p_vs.old, _p_u.old, _c_u.old, _p_mrho.old, _p_rho.old,
 *             _c_rho.old, each also as .old.0.0);
 *           - the parabolic step-length search over
 *             the test shots (src/PSV/step_length_est_psv.c,
 *             src/PSV/obj_psv.c) and the model update
 *             (src/PSV/calc_mat_change_test_PSV.c), written as
 *             INV_MODELFILE_vp.bin, _vs.bin, _rho.bin;
 *           - a line of MISFIT_LOG_FILE (FWI_PSV.c:735-742), and the stage
 *             end test on the relative misfit change PRO (FWI_PSV.c:781-827)
 *             with the models of src/PSV/model_freq_out_PSV.c (INV_MOD_OUT = 0)
 *             or src/PSV/model_it_out_PSV.c (INV_MOD_OUT = 1).
 *           ITERMAX = 1 gives one gradient and one model update.
 * MODE = 2: reverse-time migration (src/PSV/RTM_PSV.c:265-385): the gradient
 *           pass with the time-integrated observed data as adjoint sources
 *           (calc_res.c:434-442), written as JACOBIAN_P_image and
 *           JACOBIAN_S_image (src/PSV/RTM_PSV_out.c) or per shot with
 *           RTM_SHOT = 1.
 * Exit status 1 with a message on any error, as the reference's err().
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "denise_fwi.h"
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

/* seconds spent in the time loops of the run (single-shot runs and engine
 * phases), printed at the end */
static double g_loop_s = 0.0;

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* dn_run_forward / dn_run_adjoint of the one-context loop, timed */
static void run_forward(dn_ctx *c, const char *what)
{
    const double t0 = now_s();
    dn_ok(dn_run_forward(c, 0), what);
    g_loop_s += now_s() - t0;
}

static void run_adjoint(dn_ctx *c)
{
    const double t0 = now_s();
    dn_ok(dn_run_adjoint(c, 0), "adjoint run");
    g_loop_s += now_s() - t0;
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
    if (p->physics < 1 || p->physics > 5)
        die("PHYSICS", "2D P/SV (1), acoustic (2), VTI (3), TTI (4) and SH (5) are built");
    /* physics_SH.c: forward modelling (MODE 0) and FWI (FWI_SH.c), no RTM
     * (its RTM call is commented out); physics_AC.c: MODE 0, FWI (FWI_AC.c)
     * and RTM (MODE 2); physics_VTI.c / physics_TTI.c offer MODE 0 and RTM,
     * their FWI branch is commented out */
    /* FD_SH.c:171-206: with L > 0 the SH forward run reads the viscoelastic
     * model (readmod_visc_SH) and updates the stresses with memory
     * variables; the viscoelastic SH FWI (FWI_SH_visc.c) is not built */
    if (p->physics == 5 && p->L > 0 && p->mode != 0)
        die("L", "viscoelastic SH runs forward modelling (MODE 0); FWI_SH_visc is not built");
    if (p->physics == 5 && p->L > 3) die("L", "viscoelastic SH runs take L <= 3");
    if (p->physics == 5 && p->mode == 2) die("MODE", "the SH loop has no RTM (physics_SH.c comments it out)");
    if ((p->physics == 3 || p->physics == 4) && p->mode == 1)
        die("MODE", "the VTI/TTI loops run forward modelling (MODE 0) and RTM (MODE 2); the reference has no "
                    "VTI/TTI FWI");
    /* TTI.c stores ux/uy/uxy for GRAD_FORM 2, which update_s_elastic_PML_TTI never writes */
    if (p->physics == 4 && p->mode == 2 && p->grad_form == 2)
        die("GRAD_FORM", "TTI RTM runs with GRAD_FORM 1 (the TTI update writes no ux/uy/uxy)");
    if (p->physics != 1 && p->snap)
        die("SNAP", "wavefield snapshots of the acoustic, VTI/TTI and SH loops (snap_AC, snap_ssg_VTI, snap) are not "
                    "built");
    if (p->mode < 0 || p->mode > 2) die("MODE", "MODE 0 (forward), 1 (FWI) and 2 (RTM) are built");
    if (p->boundary != 0) die("BOUNDARY", "periodic boundaries are not supported");
    if (!p->srcrec) die("SRCREC", "sources must come from SOURCE_FILE (SRCREC = 1)");
    if (p->snap && p->mode != 0) fprintf(stderr, " denise_hip: SNAP > 0 ignored in MODE %d (snapshots of MODE 0 only)\n", p->mode);
    if (p->snap && p->mode == 0 && p->snap_format != 2 && p->snap_format != 3)
        die("SNAP_FORMAT", "snapshots are written as ASCII (2) or binary (3); SU (1) is not implemented, as in the reference");
    if (p->mode == 1 || p->mode == 2) {
        if (!st) die("workflow", "MODE 1 and 2 need a workflow file");
        if (st->lnorm != 2 && st->lnorm != 5) die("LNORM", "the L2 (2) and global-correlation (5) misfits are built");
        if (st->lnorm == 5 && (p->mode != 1 || st->n_order != 0))
            die("LNORM", "the global-correlation misfit is built for FWI (MODE 1) without N_ORDER");
        /* smooth_grad.c:17-33: 1 wavenumber damping, 2 smooth2, 3 Gaussian, 4 velocity-dependent Gaussian */
        if (st->spatfilter == 1 || st->spatfilter == 2)
            die("SPATFILTER", "the Gaussian gradient filters (3, 4) are built; 1 and 2 are not");
        if ((st->spatfilter == 3 || st->spatfilter == 4) && !(st->fc_end > 0.0f))
            die("SPATFILTER", "SPATFILTER 3 and 4 need FC_END > 0");
        /* NORMALIZE acts in the STF inversion only (stf.c:176-186) and ENV with LNORM 6 only */
        if (st->rowi) die("ROWI", "ROWI is not built");
        /* stf.c:70-74: INV_STF 2 windows the data with the STA/LTA picker */
        if (st->inv_stf != 0 && st->inv_stf != 1) die("INV_STF", "0 or 1 (Wiener deconvolution) is built; 2 needs the STA/LTA picker");
        if (st->inv_stf && p->mode != 1) die("INV_STF", "the STF inversion runs in FWI (MODE 1)");
        if (st->inv_stf && (st->normalize < 0 || st->normalize > 2)) die("NORMALIZE", "0, 1 or 2 expected with INV_STF");
        if (st->timewin == 3) die("TIMEWIN", "time windows 1, 2, 4 and 5 are built; the STA/LTA picker (3) is not");
        if (st->offset_mute < 0 || st->offset_mute > 2) die("OFFSET_MUTE", "0, 1 or 2 expected");
        if (p->ndt != 1) die("NDT", "FWI needs NDT = 1");
        if (p->idx != 1 || p->idy != 1) die("IDX/IDY", "FWI runs on every grid point (IDX = IDY = 1)");
        if (p->grad_form != 1 && p->grad_form != 2) die("GRAD_FORM", "1 or 2 expected");
        if (p->trkill && p->readrec == 2) die("TRKILL", "trace killing needs one receiver file (READREC = 1)");
        /* calc_res_PSV.c:39-61: differences against a second data set */
        if (p->timelapse == 1) die("TIMELAPSE", "time-lapse residuals (TIMELAPSE = 1) are not built");
    }
    if (p->mode == 1) {
        if (p->grad_method != 1 && p->grad_method != 2) die("GRAD_METHOD", "PCG (1) or L-BFGS (2) expected");
        if (p->grad_method == 1 && (p->pcg_beta < 1 || p->pcg_beta > 4)) die("PCG_BETA", "1 to 4 expected");
        /* GRAD_FILTER is read by the SH inversion only (FWI_SH.c:635); P/SV ignores it */
        if (p->gravity) die("GRAVITY", "gravity modelling is not built");
        if (p->nlbfgs < 1) die("NLBFGS", "at least 1 expected");
        if (p->invmat1 != 1) die("INVMAT1", "the vp/vs/rho parameterisation (INVMAT1 = 1) is built");
    }
}

static void write_plane(const char *name, const float *a, int nx, int ny)
{
    io_ok(dnio_write_model(name, a, nx, ny), name);
}

/* everything a shot needs besides the model */
typedef struct run {
    dnio_par *P;
    dn_ctx *ctx;
    int nsrc, nshots, n_order, ntr;
    float *srcpos, TS;
    int *recpos;
    size_t ncell;
    const float *vp, *vs, *rho; /* the model as read (ppi, pu of the snapshot writer) */
    int eprecond;         /* EPRECOND of the current workflow stage */
    float c_vp, c_vs, c_rho; /* the initial model's averages (FWI_PSV.c:540-580) */
    int time_filt, order; /* TIME_FILT of the current workflow stage: Butterworth ORDER, */
    float fc, fc_start;   /* low-pass FC (= FC_END of the stage), high-pass FC_START */
    int *kill;            /* TRKILL: nsrc x ntr flags of TRKILL_FILE, or NULL */
    int offset_mute;      /* OFFSET_MUTE of the current stage, with OFFSETC */
    float offsetc;
    int *flag;            /* this shot's traces to leave out (ntr) */
    int flag_cap;         /* traces flag, offs and picks hold */
    int timewin;          /* TIMEWIN of the current stage, with GAMMA, TWLENGTH_MINUS/PLUS */
    float gamma, twlength_minus, twlength_plus;
    float *picks, *offs;  /* this shot's picked times (TIMEWIN 1, 4) and offsets (ntr) */
    int lnorm;            /* LNORM of the current stage */
    int inv_stf, normalize; /* INV_STF, NORMALIZE, OFFSETC_STF, EPS_STF of the current stage */
    int stf_stage_filter;   /* the SH STF run filters its spike with the stage filter too */
    float offsetc_stf, eps_stf;
    int stf_now;          /* invert the source wavelets in this gradient pass (iter 1) */
    float hc[7];
    char name[4 * DNIO_STR];
    /* multi-shot passes: the shots of a pass run together through the
     * engine, one context each, carved from one device arena */
    int batch;
    dn_params dp;                                  /* parameters of the shot contexts */
    const float *m_vp, *m_vs, *m_rho, *m_taus, *m_taup; /* the model of the next pass */
    dn_arena *arena;
    dn_engine *eng;
} run;

/* receivers, sources and wavelets of shot ishot on the context
 * (FD_PSV.c:271-300, grad_obj_psv.c:130-170, obj_psv.c:54-110) */
static void shot_setup(run *r, int ishot, float *first_signal)
{
    dnio_par *P = r->P;
    if (P->readrec == 2) {
        dnio_free(r->recpos);
        r->recpos = NULL;
        io_ok(dnio_read_receivers(P, ishot, &r->ntr, &r->recpos), "REC_FILE");
    }
    const int first = P->run_multiple_shots ? ishot - 1 : 0, count = P->run_multiple_shots ? 1 : r->nsrc;
    float *loc = xmalloc(sizeof(float) * 8 * (size_t)count);
    int nloc = 0;
    io_ok(dnio_split_sources(P, r->srcpos, r->nsrc, first, count, loc, &nloc), "sources");
    /* pack the kept columns into an 8 x nloc table */
    float *sp = xmalloc(sizeof(float) * 8 * (size_t)(nloc ? nloc : 1));
    for (int q = 0; q < 8; q++)
        for (int l = 0; l < nloc; l++) sp[q * nloc + l] = loc[q * count + l];
    float *sig = xmalloc(sizeof(float) * (size_t)(nloc ? nloc : 1) * P->nt);
    io_ok(dnio_wavelet(P, sp, nloc, ishot, r->TS, r->n_order, sig), "wavelet");
    /* grad_obj_psv.c:140-156 (and FD_PSV.c:327-333, obj_psv.c:106-120): the
     * band-limited spike, then the stage's time-domain filter */
    if (P->quellart == 6)
        io_ok(dnio_tdfilt(sig, nloc, P->nt, P->dt, P->order_spike, P->fc_spike_2, P->fc_spike_1), "spike filter");
    /* grad_obj_psv.c:151, obj_psv.c:116: no TIME_FILT of the source with
     * INV_STF (the inverted wavelet was fitted to the filtered data) */
    if (r->time_filt && (!r->inv_stf || r->stf_stage_filter))
        io_ok(dnio_tdfilt(sig, nloc, P->nt, P->dt, r->order, r->fc, r->fc_start), "TIME_FILT");
    dn_ok(dn_set_sources(r->ctx, nloc, sp, sig), "sources");
    if (first_signal) {
        if (nloc > 0)
            memcpy(first_signal, sig, sizeof(float) * P->nt);
        else
            memset(first_signal, 0, sizeof(float) * P->nt);
    }
    int *rp = xmalloc(sizeof(int) * 2 * (size_t)(r->ntr ? r->ntr : 1));
    for (int k = 0; k < r->ntr; k++) {
        rp[k] = r->recpos[k];
        rp[r->ntr + k] = r->recpos[r->ntr + k];
    }
    dn_ok(dn_set_receivers(r->ctx, r->ntr, rp), "receivers");
    printf(" shot %d: %d source point(s), %d receivers\n", ishot, nloc, r->ntr);
    fflush(stdout);
    free(loc);
    free(sp);
    free(sig);
    free(rp);
}

static int is_test_shot(const dnio_par *P, int ishot)
{
    return ishot >= P->testshot_start && ishot <= P->testshot_end && P->testshot_incr > 0 &&
           (ishot - P->testshot_start) % P->testshot_incr == 0;
}

/* TRKILL (calc_res.c:61-87, 308, 544): the traces flagged for this shot
 * get no adjoint source and no misfit; the misfit of the others is summed
 * in the reference's order (per component, trace, time reversed) */
/* the traces calc_res leaves out for this shot (calc_res.c:302-316,
 * 542-560): TRKILL flags, and OFFSET_MUTE 1 (offset >= OFFSETC) or 2
 * (offset <= OFFSETC); 1 when any trace can be left out */
static int shot_flags(run *r, int ishot)
{
    dnio_par *P = r->P;
    const int tw = r->timewin == 1 || r->timewin == 2 || r->timewin == 4 || r->timewin == 5;
    if (!r->kill && !r->offset_mute && !tw) return 0;
    /* READREC 2 reads a receiver file per shot: the buffers follow the
     * largest ntr so far */
    if (!r->flag || r->ntr > r->flag_cap) {
        const size_t n = (size_t)(r->ntr ? r->ntr : 1);
        free(r->flag);
        free(r->offs);
        free(r->picks);
        r->flag = xmalloc(sizeof(int) * n);
        r->offs = xmalloc(sizeof(float) * n);
        r->picks = xmalloc(sizeof(float) * n);
        r->flag_cap = (int)n;
    }
    if (r->timewin == 1 || r->timewin == 4) {
        /* time_window.c:29-44 (TIMEWIN 1: PICKS_FILE<shot>.dat) and 49-61
         * (TIMEWIN 4: PICKS_FILE shot<shot>.<POS1><POS2>, one domain: .00) */
        if (r->timewin == 1)
            snprintf(r->name, sizeof r->name, "%s%i.dat", P->picks_file, ishot);
        else
            snprintf(r->name, sizeof r->name, "%sshot%d.%d%d", P->picks_file, ishot, 0, 0);
        FILE *f = fopen(r->name, "r");
        if (!f) die(r->name, " picks_?.dat could not be opened !");
        for (int i = 0; i < r->ntr; i++)
            if (fscanf(f, r->timewin == 1 ? "%f" : "%e", &r->picks[i]) != 1) die(r->name, "fewer than NTR picked times");
        fclose(f);
    }
    const float xs = r->srcpos[0 * r->nsrc + ishot - 1], ys = r->srcpos[1 * r->nsrc + ishot - 1];
    for (int i = 0; i < r->ntr; i++) {
        int f = r->kill ? r->kill[(size_t)(ishot - 1) * r->ntr + i] == 1 : 0;
        if (r->offset_mute || r->timewin == 5) { /* calc_res.c:302-316 */
            const float xr = r->recpos[i] * P->dh, yr = r->recpos[r->ntr + i] * P->dh;
            const float offset = sqrt(((xs - xr) * (xs - xr)) + ((ys - yr) * (ys - yr)));
            r->offs[i] = offset;
            if ((r->offset_mute == 1) && (offset >= r->offsetc)) f = 1;
            if ((r->offset_mute == 2) && (offset <= r->offsetc)) f = 1;
        }
        r->flag[i] = f;
    }
    return 1;
}

static double kill_traces(run *r, int ishot, int test, float **diff)
{
    dnio_par *P = r->P;
    (void)ishot;
    const int *flag = r->flag;
    double l2 = 0.0;
    for (int c = 0; c < 3; c++) {
        if (!diff[c]) continue;
        for (int i = 0; i < r->ntr; i++) {
            float *d = diff[c] + (size_t)i * P->nt;
            if (flag[i] == 1) memset(d, 0, sizeof(float) * P->nt);
        }
        /* calc_res.c:447-450 and 526-537: TIMEWIN on the residuals before the misfit */
        io_ok(dnfwi_time_window(r->timewin, r->ntr, P->nt, P->dt, r->gamma, r->twlength_minus, r->twlength_plus,
                                r->picks, r->offs, diff[c]),
              "TIMEWIN");
        for (int i = 0; i < r->ntr; i++) {
            const float *d = diff[c] + (size_t)i * P->nt;
            if (flag[i] == 1) continue;
            if (test)
                for (int k = P->nt - 1; k >= 0; k--) l2 += 0.5 * d[k] * d[k];
        }
    }
    dn_ok(dn_set_adjoint_sources(r->ctx, diff[0], diff[1], diff[2]), "adjoint sources");
    return l2;
}

static int *read_kill(const char *path, int nsrc, int ntr)
{
    FILE *f = fopen(path, "r");
    if (!f) die("TRKILL_FILE", " Trace kill file could not be opened!");
    int *k = xmalloc(sizeof(int) * (size_t)nsrc * (ntr ? ntr : 1));
    for (size_t q = 0; q < (size_t)nsrc * ntr; q++)
        if (fscanf(f, "%d", &k[q]) != 1) die("TRKILL_FILE", "fewer than NSRC x NTR flags");
    fclose(f);
    return k;
}

/* COMP_WEIGHT (calc_res_PSV.c:65-67, 118-120; calc_res.c:100-123): with
 * QUELLTYPB 5 and 7 the vx traces, with 6 and 7 the vy traces, synthetic
 * and observed, are scaled by C_vp*C_rho before their residuals */
static float comp_weight(const run *r, int c)
{
    const int qb = r->P->quelltypb;
    if (r->P->physics == 5) return 1.0f; /* calc_res_SH.c:39: COMP_WEIGHT = 0 */
    const int weighted = (c == 0 && (qb == 5 || qb == 7)) || (c == 1 && (qb == 6 || qb == 7));
    if (!weighted) return 1.0f;
    const float tmp = r->c_vp * r->c_rho;
    return tmp;
}

static void scale_traces(float *a, size_t n, float w)
{
    if (w == 1.0f) return;
    for (size_t k = 0; k < n; k++) a[k] *= w;
}

/* residuals of the shot on the context against its observed data;
 * returns its L2 misfit (0 unless a test shot) */
static double shot_residuals(run *r, int ishot, int test)
{
    dnio_par *P = r->P;
    float *obs[3] = {NULL, NULL, NULL};
    const int qb = P->quelltypb;
    int use[3] = {qb == 1 || qb == 3 || qb == 5 || qb == 7, qb == 1 || qb == 2 || qb == 6 || qb == 7, qb >= 4};
    if (P->physics == 5) { /* calc_res_SH.c:20-44: the vz data (SEIS_FILE_VY) for any QUELLTYPB */
        use[0] = use[2] = 0;
        use[1] = qb != 0;
    }
    for (int c = 0; c < 3; c++)
        if (use[c]) {
            obs[c] = xmalloc(sizeof(float) * (size_t)r->ntr * P->nt);
            io_ok(dnio_data_name(P, c, ishot, r->name, sizeof r->name), "file name");
            io_ok(dnio_read_su(r->name, r->ntr, P->nt, obs[c]), r->name);
            /* calc_res_PSV.c:31-32 etc.: the observed data get the stage's filter */
            if (r->time_filt) io_ok(dnio_tdfilt(obs[c], r->ntr, P->nt, P->dt, r->order, r->fc, r->fc_start), "TIME_FILT");
        }
    double l2 = 0.0;
    if (r->lnorm == 5) {
        /* calc_res.c with LNORM 5: residuals and misfit on the host from the
         * device's seismograms, then the time windows, as RTM's sources */
        const int fl = shot_flags(r, ishot);
        float *diff[3] = {NULL, NULL, NULL}, *syn = xmalloc(sizeof(float) * (size_t)r->ntr * P->nt);
        for (int c = 0; c < 3; c++)
            if (use[c]) {
                diff[c] = xmalloc(sizeof(float) * (size_t)r->ntr * P->nt);
                dn_ok(dn_get_seismogram(r->ctx, c, syn), "seismograms");
                scale_traces(obs[c], (size_t)r->ntr * P->nt, comp_weight(r, c));
                scale_traces(syn, (size_t)r->ntr * P->nt, comp_weight(r, c));
                io_ok(dnfwi_residual_gc(obs[c], syn, r->ntr, P->nt, P->dt, P->grad_form, test, fl ? r->flag : NULL,
                                        diff[c], &l2),
                      "LNORM 5 residuals");
                /* calc_res.c:526-537: TIMEWIN 1, 2 and 4 act on every misfit;
                 * TIMEWIN 5 sits in the LNORM 2 branch (calc_res.c:447-450)
                 * and leaves LNORM 5 residuals alone */
                if (fl && r->timewin != 5)
                    io_ok(dnfwi_time_window(r->timewin, r->ntr, P->nt, P->dt, r->gamma, r->twlength_minus,
                                            r->twlength_plus, r->picks, r->offs, diff[c]),
                          "TIMEWIN");
            }
        dn_ok(dn_set_adjoint_sources(r->ctx, diff[0], diff[1], diff[2]), "adjoint sources");
        for (int c = 0; c < 3; c++) {
            free(diff[c]);
            free(obs[c]);
        }
        free(syn);
        return l2;
    }
    dn_ok(dn_set_residual_weights(r->ctx, comp_weight(r, 0), comp_weight(r, 1)), "residual weights");
    dn_ok(dn_calc_residuals(r->ctx, obs[0], obs[1], obs[2], test, &l2), "residuals");
    if (shot_flags(r, ishot)) {
        float *diff[3] = {NULL, NULL, NULL};
        for (int c = 0; c < 3; c++)
            if (use[c]) {
                diff[c] = xmalloc(sizeof(float) * (size_t)r->ntr * P->nt);
                dn_ok(dn_get_adjoint_sources(r->ctx, c, diff[c]), "adjoint sources");
            }
        l2 = kill_traces(r, ishot, test, diff);
        for (int c = 0; c < 3; c++) free(diff[c]);
    }
    for (int c = 0; c < 3; c++) free(obs[c]);
    return l2;
}

/* INV_STF 1 (grad_obj_psv.c:127-129 -> src/PSV/stf_psv.c:13-66): a forward
 * run of this shot with the band-limited spike (QUELLART 6, its
 * ORDER_SPIKE filter only; SH, src/SH/stf_sh.c:30-50: then the stage's
 * filter too), the observed vy (SH: vz) data with the stage's filter,
 * and the Wiener deconvolution of src/stf.c; the wavelet is written as
 * SIGNAL_FILE_shot_<ishot>.dat (%e per line), and from here on every
 * wavelet of the run is read from those files (QUELLART 3, as the
 * reference leaves it after stf_psv) */
static void stf_shot(run *r, int ishot)
{
    dnio_par *P = r->P;
    const size_t n = (size_t)r->ntr * P->nt;
    float *spike = xmalloc(sizeof(float) * P->nt), *stf = xmalloc(sizeof(float) * P->nt);
    P->quellart = 6;
    const int tf = r->time_filt;
    /* stf_sh.c:30-50 filters the spike with the stage's filter as well */
    if (P->physics == 5)
        r->stf_stage_filter = 1;
    else
        r->time_filt = 0;
    shot_setup(r, ishot, spike);
    r->time_filt = tf;
    r->stf_stage_filter = 0;
    run_forward(r->ctx, "STF forward run");
    float *syn = xmalloc(sizeof(float) * n), *obs = xmalloc(sizeof(float) * n);
    dn_ok(dn_get_seismogram(r->ctx, 1, syn), "seismograms");
    io_ok(dnio_data_name(P, 1, ishot, r->name, sizeof r->name), "file name");
    io_ok(dnio_read_su(r->name, r->ntr, P->nt, obs), r->name);
    if (r->time_filt) io_ok(dnio_tdfilt(obs, r->ntr, P->nt, P->dt, r->order, r->fc, r->fc_start), "TIME_FILT");
    /* stf.c:96-166: TRKILL, OFFSET_MUTE and OFFSETC_STF leave traces out */
    int *mute = xmalloc(sizeof(int) * (size_t)(r->ntr ? r->ntr : 1));
    const float xs = r->srcpos[0 * r->nsrc + ishot - 1], ys = r->srcpos[1 * r->nsrc + ishot - 1];
    for (int i = 0; i < r->ntr; i++) {
        int f = r->kill ? r->kill[(size_t)(ishot - 1) * r->ntr + i] == 1 : 0;
        const float xr = r->recpos[i] * P->dh, yr = r->recpos[r->ntr + i] * P->dh;
        const float offset = sqrt(((xs - xr) * (xs - xr)) + ((ys - yr) * (ys - yr)));
        if ((r->offset_mute == 1) && (offset >= r->offsetc)) f = 1;
        if ((r->offset_mute == 2) && (offset <= r->offsetc)) f = 1;
        if (r->offsetc_stf > 0.0f && fabs(offset) >= r->offsetc_stf) f = 1;
        mute[i] = f;
    }
    io_ok(dnfwi_stf(obs, syn, r->ntr, P->nt, mute, r->normalize, r->eps_stf, spike, stf), "STF inversion");
    snprintf(r->name, sizeof r->name, "%s_shot_%i.dat", P->signal_file, ishot);
    FILE *f = fopen(r->name, "w");
    if (!f) die("SIGNAL_FILE", r->name);
    for (int j = 0; j < P->nt; j++) fprintf(f, "%e\n", stf[j]);
    fclose(f);
    printf("   source wavelet of shot %d inverted (Wiener deconvolution), written to %s\n", ishot, r->name);
    P->quellart = 3;
    free(spike);
    free(stf);
    free(syn);
    free(obs);
    free(mute);
}

/* MODE 2: the observed data as adjoint sources (calc_res.c:416-446 with
 * MODE = 2); returns their L2 (0 unless a test shot) */
static double shot_rtm_sources(run *r, int ishot, int test)
{
    dnio_par *P = r->P;
    const int qb = P->quelltypb;
    const int use[3] = {qb == 1 || qb == 3 || qb == 5 || qb == 7, qb == 1 || qb == 2 || qb == 6 || qb == 7, qb >= 4};
    const size_t n = (size_t)r->ntr * P->nt;
    float *diff[3] = {NULL, NULL, NULL}, *obs = xmalloc(sizeof(float) * n), *syn = xmalloc(sizeof(float) * n);
    double l2 = 0.0;
    for (int c = 0; c < 3; c++)
        if (use[c]) {
            double l2c = 0.0;
            diff[c] = xmalloc(sizeof(float) * n);
            io_ok(dnio_data_name(P, c, ishot, r->name, sizeof r->name), "file name");
            io_ok(dnio_read_su(r->name, r->ntr, P->nt, obs), r->name);
            if (r->time_filt) io_ok(dnio_tdfilt(obs, r->ntr, P->nt, P->dt, r->order, r->fc, r->fc_start), "TIME_FILT");
            dn_ok(dn_get_seismogram(r->ctx, c, syn), "seismograms");
            scale_traces(obs, n, comp_weight(r, c));
            scale_traces(syn, n, comp_weight(r, c));
            io_ok(dnfwi_rtm_adjoint_source(obs, syn, r->ntr, P->nt, P->dt, P->grad_form, test, diff[c], &l2c),
                  "RTM adjoint source");
            l2 += l2c;
        }
    if (shot_flags(r, ishot))
        l2 = kill_traces(r, ishot, test, diff);
    else
        dn_ok(dn_set_adjoint_sources(r->ctx, diff[0], diff[1], diff[2]), "adjoint sources");
    for (int c = 0; c < 3; c++) free(diff[c]);
    free(obs);
    free(syn);
    return l2;
}

/* include/fd.h: iround(x) = (int)floor(x + 0.5) of a float quotient */
static int iround_f(float x) { return (int)floor(x + 0.5); }

static void snap_writer(void *user, int nt, int nsnap, const float *vx, const float *vy, const float *sxx,
                        const float *syy)
{
    run *r = user;
    printf("   snapshot %d at T = %f s\n", nsnap, nt * r->P->dt);
    io_ok(dnio_write_snap(r->P, nsnap, vx, vy, sxx, syy, r->vp, r->vs, r->hc), "snapshot");
}

/* MODE 0: forward runs and SU output; with SNAP > 0 only shot SNAP_SHOT
 * (FD_PSV.c:264-268), with its snapshots (psv.c:75-77, 225-231) */
/* ---- multi-shot passes ------------------------------------------------
 * FD_PSV.c:271-354, grad_obj_psv.c:56-262 and obj_psv.c:54-124 loop over the
 * shots one by one.  Here the shots of a pass run together through the
 * multi-shot engine (dn_engine_*), one device context per shot, in groups
 * that fit the device arena; every host step around the time loops
 * (sources and wavelets, residuals and misfits, gradient weights, tapers,
 * stacking, output) runs per shot in the reference's shot order, so the
 * results equal the one-context loop bit for bit. */

/* shots go through the engine unless the run needs what it does not take:
 * VTI/TTI, moment-tensor sources, snapshots, per-shot receiver files, a
 * single shot; DENISE_HIP_SEQUENTIAL=1 keeps the one-context loop */
static int batching_applies(const run *r)
{
    const dnio_par *P = r->P;
    const char *env = getenv("DENISE_HIP_SEQUENTIAL");
    if (env && env[0] && env[0] != '0') return 0;
    if (P->physics != 1 || P->snap || P->readrec == 2 || r->nshots < 2) return 0;
    for (int l = 0; l < r->nsrc; l++)
        if ((int)r->srcpos[7 * r->nsrc + l] == 5) return 0;
    return 1;
}

/* the arena and engine of the multi-shot passes: room for every shot of the
 * largest pass (mode: params.mode of its contexts) and the test shots, up
 * to the free device memory less a reserve */
static void batch_init(run *r)
{
    if (!r->batch) return;
    dn_params p = r->dp;
    size_t fp = 0, fp0 = 0, fr = 0, tot = 0;
    p.mode = r->dp.mode;
    dn_ok(dn_estimate_footprint(&p, r->nsrc, r->ntr, &fp), "footprint");
    p.mode = 0;
    dn_ok(dn_estimate_footprint(&p, r->nsrc, r->ntr, &fp0), "footprint");
    dn_ok(dn_device_memory(0, &fr, &tot), "device memory");
    int ntest = 0;
    for (int ishot = 1; ishot <= r->nshots; ishot++) ntest += is_test_shot(r->P, ishot);
    const size_t slack = 1 << 20; /* cell lists and alignment */
    size_t need = (size_t)r->nshots * (fp + slack) + (size_t)ntest * (fp0 + slack);
    const size_t reserve = (size_t)4 << 30;
    size_t cap = fr > reserve ? (fr - reserve) / 10 * 9 : 0;
    /* DENISE_HIP_ARENA_BYTES: a smaller arena (the passes then run in
     * several groups of shots) */
    const char *lim = getenv("DENISE_HIP_ARENA_BYTES");
    if (lim && atoll(lim) > 0 && (size_t)atoll(lim) < cap) cap = (size_t)atoll(lim);
    if (need > cap) need = cap;
    if (need < fp + slack) { /* not even one shot: the one-context loop */
        r->batch = 0;
        return;
    }
    dn_ok(dn_arena_create(0, need, &r->arena), "device arena");
    dn_ok(dn_engine_create(0, &r->eng), "multi-shot engine");
    printf(" multi-shot passes: %d shots through the engine, %.2f GB device arena\n", r->nshots, need / 1e9);
}

static void batch_free(run *r)
{
    if (r->eng) dn_engine_destroy(r->eng);
    if (r->arena) dn_arena_destroy(r->arena);
    r->eng = NULL;
    r->arena = NULL;
}

/* a context for one shot of a pass on the pass's model, or NULL when the
 * arena has no room left for it */
static dn_ctx *shot_ctx(run *r, int mode)
{
    dn_params p = r->dp;
    p.mode = mode;
    int fits = 0;
    dn_ok(dn_arena_fits(r->arena, &p, r->nsrc, r->ntr, &fits), "arena");
    if (!fits) return NULL;
    dn_ctx *c = NULL;
    dn_ok(dn_ctx_create_in(&p, r->arena, &c), "shot context");
    dn_ok(dn_set_model(c, r->m_vp, r->m_vs, r->m_rho, r->m_taus, r->m_taup), "model upload");
    if (mode > 0 && r->eprecond) dn_ok(dn_set_energy_weights(c, r->eprecond), "energy weights");
    return c;
}

/* one phase (0 forward, 1 adjoint) of n prepared shots through the engine */
static void engine_phase(run *r, dn_ctx **cs, int n, int phase)
{
    const char *what = phase ? "adjoint run" : "forward run";
    printf(" engine: %d shot(s), %s\n", n, what);
    fflush(stdout);
    const double t0 = now_s();
    for (int k = 0; k < n; k++) dn_ok(dn_engine_attach(r->eng, cs[k], phase), what);
    int running = n;
    while (running > 0) dn_ok(dn_engine_step(r->eng, 512, &running), what);
    g_loop_s += now_s() - t0;
    for (int k = 0; k < n; k++) {
        int st = 0;
        dn_ok(dn_engine_status(r->eng, cs[k], &st), what);
        if (st == DN_ERR_UNSTABLE) die(what, " Simulation is unstable !");
        dn_ok(dn_engine_forget(r->eng, cs[k]), what);
    }
}

/* saveseis_glob for one shot from the context's seismograms */
static void write_shot_seis(run *r, int ishot)
{
    dnio_par *P = r->P;
    if (r->ntr <= 0 || !P->seismo) return;
    const float XS = r->srcpos[0 * r->nsrc + ishot - 1], YS = r->srcpos[1 * r->nsrc + ishot - 1];
    /* saveseis_glob: SEISMO 1 vx,vy; 2 p; 3 div,curl; 4 vx,vy,p;
     * saveseis_glob_SH: SEISMO 1 vz to SEIS_FILE_VY, nothing else */
    int comps[5], nc = 0;
    if (P->physics == 5) {
        if (P->seismo == 1) comps[nc++] = 1;
    } else {
        if (P->seismo == 1 || P->seismo == 4) comps[nc++] = 0, comps[nc++] = 1;
        if (P->seismo == 2 || P->seismo == 4) comps[nc++] = 2;
        if (P->seismo == 3) comps[nc++] = 3, comps[nc++] = 4;
    }
    float *sec = xmalloc(sizeof(float) * (size_t)r->ntr * P->nt);
    for (int c = 0; c < nc; c++) {
        dn_ok(dn_get_seismogram(r->ctx, comps[c], sec), "seismograms");
        io_ok(dnio_seis_name(P, comps[c], ishot, 1, r->name, sizeof r->name), "file name");
        io_ok(dnio_write_seis(r->name, P, sec, r->ntr, P->nt, r->recpos, XS, YS), r->name);
        printf("   wrote %s\n", r->name);
    }
    free(sec);
}

static void forward_pass(run *r)
{
    dnio_par *P = r->P;
    int first = 1, last = r->nshots;
    if (P->snap) {
        if (P->snap_shot < 1 || P->snap_shot > r->nshots) die("SNAP_SHOT", "no such shot");
        first = last = P->snap_shot;
        dn_ok(dn_get_hc(r->ctx, r->hc), "FD coefficients");
        dn_ok(dn_set_snapshots(r->ctx, iround_f(P->tsnap1 / P->dt), iround_f(P->tsnapinc / P->dt),
                               iround_f(P->tsnap2 / P->dt), snap_writer, r),
              "snapshots");
    }
    if (!r->batch) {
        for (int ishot = first; ishot <= last; ishot++) {
            shot_setup(r, ishot, NULL);
            run_forward(r->ctx, "forward run");
            write_shot_seis(r, ishot);
        }
        return;
    }
    dn_ctx *own = r->ctx, **cs = xmalloc(sizeof(dn_ctx *) * (size_t)r->nshots);
    for (int k0 = first; k0 <= last;) {
        int n = 0;
        for (int ishot = k0; ishot <= last; ishot++) {
            dn_ctx *c = shot_ctx(r, 0);
            if (!c) break;
            cs[n++] = r->ctx = c;
            shot_setup(r, ishot, NULL);
        }
        if (n == 0) die("device memory", "one shot does not fit the device arena");
        engine_phase(r, cs, n, 0);
        for (int q = 0; q < n; q++) {
            r->ctx = cs[q];
            write_shot_seis(r, k0 + q);
            dn_ctx_destroy(cs[q]);
        }
        k0 += n;
    }
    r->ctx = own;
    free(cs);
}

/* grad_obj_psv: stacked vp/vs/rho gradients over all shots; returns the
 * L2 misfit summed over the test shots.  MODE 2 takes the data themselves
 * as adjoint sources (reverse-time migration). */
/* the per-shot part of grad_obj_psv.c:215-266 after the adjoint run:
 * ass_gradPSV, energy weights, the per-shot source taper, stacking and the
 * per-shot images */
static void shot_gradient(run *r, int ishot, double l2, int test, float **gs, float *ws, float *wr, float **g_sum)
{
    dnio_par *P = r->P;
    if (P->physics == 3 || P->physics == 4) {
        /* grad_obj_VTI.c:196-197 / grad_obj_TTI.c: ass_gradVTI is not called
         * (and assembles nothing): the images are the raw correlation sums */
        for (int k = 0; k < 3; k++) dn_ok(dn_get_gradient(r->ctx, k, gs[k]), "image");
    } else if (P->physics == 5) {
        /* grad_obj_sh.c:211-238: the raw shear and density sums of the shot;
         * FWI_SH.c assembles (ass_gradSH) after the stacking */
        for (int k = 1; k < 3; k++) dn_ok(dn_get_gradient(r->ctx, k, gs[k]), "gradient sums");
    } else {
        dn_ok(dn_assemble_gradient(r->ctx), "gradient");
        for (int k = 0; k < 3; k++) dn_ok(dn_get_gradient(r->ctx, 3 + k, gs[k]), "gradient");
    }
    /* grad_obj_ac.c: the acoustic problem has vp and rho gradients only,
     * grad_obj_sh.c vs and rho */
    const int ac = P->physics == 2, sh = P->physics == 5;
    if (r->eprecond) {
        /* grad_obj_psv.c:219-238, grad_obj_ac.c:223-234: energy weights of this shot */
        dn_ok(dn_get_energy(r->ctx, 0, ws), "energy");
        if (r->eprecond == 1) dn_ok(dn_get_energy(r->ctx, 1, wr), "energy");
        io_ok(dnfwi_energy_precond(r->eprecond, P->nx, P->ny, P->dh, ws, wr, r->c_vp, r->c_vs, r->c_rho,
                                   sh ? NULL : gs[0], ac ? NULL : gs[1], gs[2], NULL),
              "energy preconditioning");
    }
    if (P->sws_taper_circular_per_shot) {
        /* grad_obj_psv.c:241-250, taper_grad_shot.c: the source taper of this shot's source alone */
        float col[8], *coeff = xmalloc(sizeof(float) * r->ncell);
        for (int q = 0; q < 8; q++) col[q] = r->srcpos[(size_t)q * r->nsrc + (ishot - 1)];
        for (int k = 0; k < 3; k++) {
            if ((ac && k == 1) || (sh && k == 0)) continue;
            io_ok(dnfwi_taper_sources(P->nx, P->ny, P->dh, P->srtradius, P->srtshape, P->filtsize, 1, col, coeff,
                                      gs[k]),
                  "per-shot gradient taper");
        }
        free(coeff);
    }
    for (int k = 0; k < 3; k++) {
        if ((ac && k == 1) || (sh && k == 0)) continue;
        for (size_t q = 0; q < r->ncell; q++) g_sum[k][q] += gs[k][q];
        /* grad_obj_psv.c:263-266; RTM_AC_out_shot.c writes the P image only */
        if (P->rtm_shot && k < (P->physics == 2 ? 1 : 2)) {
            snprintf(r->name, sizeof r->name, "%s_%s_image_shot_%i", P->jacobian, k == 0 ? "P" : "S", ishot);
            write_plane(r->name, gs[k], P->nx, P->ny);
        }
    }
    printf("   L2 of shot %d: %.9e%s\n", ishot, l2, test ? "" : " (not a test shot)");
}

/* grad_obj_psv: stacked vp/vs/rho gradients over all shots; returns the
 * L2 misfit summed over the test shots.  MODE 2 takes the data themselves
 * as adjoint sources (reverse-time migration). */
static double gradient_pass(run *r, float **g_sum)
{
    dnio_par *P = r->P;
    float *gs[3], *ws = NULL, *wr = NULL;
    for (int k = 0; k < 3; k++) gs[k] = xmalloc(sizeof(float) * r->ncell);
    if (r->eprecond) {
        ws = xmalloc(sizeof(float) * r->ncell);
        wr = xmalloc(sizeof(float) * r->ncell);
    }
    for (int k = 0; k < 3; k++) memset(g_sum[k], 0, sizeof(float) * r->ncell);
    double misfit = 0.0;
    if (!r->batch) {
        dn_ok(dn_set_energy_weights(r->ctx, r->eprecond), "energy weights");
        for (int ishot = 1; ishot <= r->nshots; ishot++) {
            if (r->stf_now) stf_shot(r, ishot); /* grad_obj_psv.c:127-129 */
            shot_setup(r, ishot, NULL);
            run_forward(r->ctx, "forward run");
            const int test = is_test_shot(P, ishot);
            const double l2 = P->mode == 2 ? shot_rtm_sources(r, ishot, test) : shot_residuals(r, ishot, test);
            misfit += l2;
            run_adjoint(r->ctx);
            shot_gradient(r, ishot, l2, test, gs, ws, wr, g_sum);
        }
    } else {
        dn_ctx *own = r->ctx, **cs = xmalloc(sizeof(dn_ctx *) * (size_t)r->nshots);
        double *l2s = xmalloc(sizeof(double) * (size_t)r->nshots);
        for (int k0 = 1; k0 <= r->nshots;) {
            int n = 0;
            for (int ishot = k0; ishot <= r->nshots; ishot++) {
                dn_ctx *c = shot_ctx(r, 1);
                if (!c) break;
                cs[n++] = r->ctx = c;
                if (r->stf_now) stf_shot(r, ishot); /* grad_obj_psv.c:127-129 */
                shot_setup(r, ishot, NULL);
            }
            if (n == 0) die("device memory", "one gradient shot does not fit the device arena");
            engine_phase(r, cs, n, 0);
            for (int q = 0; q < n; q++) {
                const int ishot = k0 + q, test = is_test_shot(P, ishot);
                r->ctx = cs[q];
                l2s[q] = P->mode == 2 ? shot_rtm_sources(r, ishot, test) : shot_residuals(r, ishot, test);
                misfit += l2s[q];
            }
            engine_phase(r, cs, n, 1);
            for (int q = 0; q < n; q++) {
                const int ishot = k0 + q;
                r->ctx = cs[q];
                shot_gradient(r, ishot, l2s[q], is_test_shot(P, ishot), gs, ws, wr, g_sum);
                dn_ctx_destroy(cs[q]);
            }
            k0 += n;
        }
        r->ctx = own;
        free(cs);
        free(l2s);
    }
    for (int k = 0; k < 3; k++) free(gs[k]);
    free(ws);
    free(wr);
    return misfit;
}

/* precond_PSV (src/PSV/precond_PSV.c:13-60): per stacked gradient (vp, vs,
 * rho) the vertical taper (taper_grad.c:46-186, TFILE_coeff_vert.bin), the
 * horizontal taper (188-322, TFILE_coeff_hor.bin), the cylindrical taper
 * around the sources (328-577, TFILE_coeff_sources.bin) and the taper read
 * from TFILE.bin, TFILE_u.bin, TFILE_rho.bin (578-640), in that order */
static void precondition(run *r, float **g)
{
    dnio_par *P = r->P;
    if (!P->sws_taper_grad_vert && !P->sws_taper_grad_hor && !P->sws_taper_grad_sources && !P->sws_taper_file) return;
    float *coeff = xmalloc(sizeof(float) * r->ncell);
    for (int k = 0; k < 3; k++) {
        if (k == 1 && P->physics == 2) continue; /* precond_AC.c: no shear-modulus gradient */
        if (k == 0 && P->physics == 5) continue; /* precond_SH.c: vs and rho (taper files _u, _rho) */
        if (P->sws_taper_grad_vert) {
            io_ok(dnfwi_taper_vert(P->nx, P->ny, P->dh, P->gradt, r->nsrc, r->srcpos, r->ntr, r->recpos, coeff, g[k]),
                  "vertical gradient taper");
            snprintf(r->name, sizeof r->name, "%s_coeff_vert.bin", P->tfile);
            write_plane(r->name, coeff, P->nx, P->ny);
        }
        if (P->sws_taper_grad_hor) {
            io_ok(dnfwi_taper_hor(P->nx, P->ny, P->dh, P->gradt[0], P->gradt[1], P->exp_taper_grad_hor, coeff, g[k]),
                  "horizontal gradient taper");
            snprintf(r->name, sizeof r->name, "%s_coeff_hor.bin", P->tfile);
            write_plane(r->name, coeff, P->nx, P->ny);
        }
        if (P->sws_taper_grad_sources) {
            io_ok(dnfwi_taper_sources(P->nx, P->ny, P->dh, P->srtradius, P->srtshape, P->filtsize, r->nsrc, r->srcpos,
                                      coeff, g[k]),
                  "source gradient taper");
            snprintf(r->name, sizeof r->name, "%s_coeff_sources.bin", P->tfile);
            write_plane(r->name, coeff, P->nx, P->ny);
        }
        if (P->sws_taper_file) {
            const char *suffix[3] = {"", "_u", "_rho"};
            snprintf(r->name, sizeof r->name, "%s%s.bin", P->tfile, suffix[k]);
            FILE *f = fopen(r->name, "rb");
            if (!f) die("SWS_TAPER_FILE", r->name);
            const size_t got = fread(coeff, sizeof(float), r->ncell, f);
            fclose(f);
            if (got != r->ncell) die("SWS_TAPER_FILE", "taper file shorter than NX*NY floats");
            for (size_t q = 0; q < r->ncell; q++) g[k][q] *= coeff[q];
        }
    }
    free(coeff);
}

/* obj_psv: forward runs of the test shots on the current model, L2 sum */
typedef struct obj_env {
    run *r;
    const float *taus, *taup;
} obj_env;

static double misfit_of(void *user, const float *vp, const float *vs, const float *rho, int itest)
{
    obj_env *o = user;
    run *r = o->r;
    double l2 = 0.0;
    if (!r->batch) {
        dn_ok(dn_set_model(r->ctx, vp, vs, rho, o->taus, o->taup), "test model");
        for (int ishot = 1; ishot <= r->nshots; ishot++) {
            if (!is_test_shot(r->P, ishot)) continue;
            printf(" test forward modelling %d for shot %d\n", itest, ishot);
            shot_setup(r, ishot, NULL);
            run_forward(r->ctx, "forward run");
            l2 += shot_residuals(r, ishot, 1);
        }
        return l2;
    }
    /* obj_psv.c:124 per test shot, the test shots together through the engine */
    dn_ctx *own = r->ctx, **cs = xmalloc(sizeof(dn_ctx *) * (size_t)r->nshots);
    int *ids = xmalloc(sizeof(int) * (size_t)r->nshots);
    r->m_vp = vp;
    r->m_vs = vs;
    r->m_rho = rho;
    for (int k0 = 1; k0 <= r->nshots;) {
        int n = 0, ishot = k0;
        for (; ishot <= r->nshots; ishot++) {
            if (!is_test_shot(r->P, ishot)) continue;
            dn_ctx *c = shot_ctx(r, 0);
            if (!c) break;
            printf(" test forward modelling %d for shot %d\n", itest, ishot);
            ids[n] = ishot;
            cs[n++] = r->ctx = c;
            shot_setup(r, ishot, NULL);
        }
        if (n == 0 && ishot <= r->nshots) die("device memory", "one test shot does not fit the device arena");
        if (n) engine_phase(r, cs, n, 0);
        for (int q = 0; q < n; q++) {
            r->ctx = cs[q];
            l2 += shot_residuals(r, ids[q], 1);
            dn_ctx_destroy(cs[q]);
        }
        k0 = ishot;
    }
    r->ctx = own;
    free(cs);
    free(ids);
    return l2;
}

/* vs NULL: the acoustic run writes vp and rho (model_freq_out_AC.c,
 * model_it_out_AC.c, calc_mat_change_test_AC.c:168-180) */
static void write_models(const char *fmt_vp, const char *fmt_vs, const char *fmt_rho, const float *vp, const float *vs,
                         const float *rho, int nx, int ny)
{
    if (vp) write_plane(fmt_vp, vp, nx, ny); /* vp NULL: the SH run (model_it_out_SH.c) */
    if (vs) write_plane(fmt_vs, vs, nx, ny);
    write_plane(fmt_rho, rho, nx, ny);
}

/* MODE 1: src/PSV/FWI_PSV.c:381-836, and for PHYSICS 2 src/AC/FWI_AC.c:386-848,
 * whose iteration is the same with the parameters vp and rho only (vs NULL
 * below): two-class L-BFGS / PCG memories, calc_mat_change_test_AC, the
 * models and gradients of vp and rho, no MODEL_FILTER (commented out at
 * FWI_AC.c:784-787) and smooth_grad with the vp model (FWI_AC.c:676) */
static void fwi(run *r, const char *workflow, float *vp_in, float *vs_in, float *rho, const float *taus,
                const float *taup)
{
    dnio_par *P = r->P;
    const int nx = P->nx, ny = P->ny;
    /* PHYSICS 5: src/SH/FWI_SH.c:384-872 with vs and rho (vp NULL below):
     * per-shot raw sums, ass_gradSH on the stack, precond_SH, GRAD_FILTER
     * and smooth_grad in that order, the normalised SH L-BFGS and the SH
     * model update (no SCALERHO) */
    const int ac = P->physics == 2, sh = P->physics == 5;
    float *vs = ac ? NULL : vs_in, *vp = sh ? NULL : vp_in;
    int stagemax = 0;
    io_ok(dnio_count_stages(workflow, &stagemax), "workflow file");
    float *g[3];
    for (int k = 0; k < 3; k++) g[k] = xmalloc(sizeof(float) * r->ncell);
    float *nv = sh ? NULL : xmalloc(sizeof(float) * r->ncell), *ns = ac ? NULL : xmalloc(sizeof(float) * r->ncell);
    float *nr = xmalloc(sizeof(float) * r->ncell);
    dnfwi_lbfgs *lb = NULL;
    dnfwi_pcg *pc = NULL;
    if (P->grad_method == 2)
        io_ok(sh   ? dnfwi_lbfgs_create_sh(nx, ny, P->nlbfgs, &lb)
              : ac ? dnfwi_lbfgs_create_ac(nx, ny, P->nlbfgs, &lb)
                   : dnfwi_lbfgs_create(nx, ny, P->nlbfgs, &lb),
              "L-BFGS memory");
    else
        io_ok(dnfwi_pcg_create(nx, ny, &pc), "PCG memory");
    double *L2_hist = xmalloc(sizeof(double) * (size_t)(P->itermax + 2));
    dnfwi_linesearch ls;
    memset(&ls, 0, sizeof ls);
    float c_vp = 0.0f, c_vs = 0.0f, c_rho = 0.0f;
    int iter_true = 1, min_iter_help = P->min_iter;
    char name[3][4 * DNIO_STR];
    for (int nstage = 1; nstage <= stagemax; nstage++) {
        dnio_stage st;
        io_ok(dnio_read_stage(workflow, nstage, &st), "workflow file");
        check_supported(P, &st);
        dnfwi_update_par up = {P->vpupperlim, P->vplowerlim, P->vsupperlim, P->vslowerlim, P->rhoupperlim,
                               P->rholowerlim, st.scalerho, st.inv_vp_iter, st.inv_vs_iter, st.inv_rho_iter,
                               P->invmat1};
        r->n_order = st.n_order;
        /* FWI_PSV.c:413: FC = FC_END for the stage */
        r->time_filt = st.time_filt;
        r->order = st.order;
        r->fc = st.fc_end;
        r->fc_start = st.fc_start;
        /* grad_obj_psv.c:71, 219: the P/SV path acts on EPRECOND 1 and 3 only */
        r->eprecond = (st.eprecond == 1 || st.eprecond == 3) ? st.eprecond : 0;
        r->offset_mute = st.offset_mute;
        r->timewin = st.timewin;
        r->lnorm = st.lnorm;
        r->gamma = st.gamma;
        r->twlength_minus = st.twlength_minus;
        r->twlength_plus = st.twlength_plus;
        r->offsetc = st.offsetc;
        r->inv_stf = st.inv_stf;
        r->normalize = st.normalize;
        r->offsetc_stf = st.offsetc_stf;
        r->eps_stf = st.eps_stf;
        for (int iter = 1; iter <= P->itermax; iter++) {
            printf("\n TDFWI ITERATION %d of %d (stage %d of %d)\n", iter, P->itermax, nstage, stagemax);
            fflush(stdout);
            if (iter_true == 1) {
                /* FWI_PSV.c:540-580, FWI_AC.c:478-513: the covariances are the
                 * initial model's averages (C_vs stays 0 in the acoustic run) */
                c_vp = sh ? 0.0f : dnfwi_model_mean(vp, nx, ny);
                c_vs = ac ? 0.0f : dnfwi_model_mean(vs, nx, ny);
                c_rho = dnfwi_model_mean(rho, nx, ny);
                if (sh) /* FWI_SH.c:526-569: C_vs = Vs_avg, C_rho = rho_avg (printed as rho_max) */
                    printf(" Vs_avg = %e \t rho_max = %e\n", c_vs, c_rho);
                else if (ac)
                    printf(" Vp_avg = %.0f \t rho_avg = %.0f\n", c_vp, c_rho);
                else
                    printf(" Vp_avg = %.0f \t Vs_avg = %.0f \t rho_avg = %.0f\n", c_vp, c_vs, c_rho);
                r->c_vp = c_vp;
                r->c_vs = c_vs;
                r->c_rho = c_rho;
            }
            dn_ok(dn_set_model(r->ctx, vp, vs, rho, taus, taup), "model upload");
            r->m_vp = vp;
            r->m_vs = vs;
            r->m_rho = rho;
            r->stf_now = r->inv_stf && iter == 1; /* grad_obj_psv.c:127 */
            const double L2sum = gradient_pass(r, g);
            r->stf_now = 0;
            if (sh) {
                /* FWI_SH.c:645-665: ass_gradSH on the stack, precond_SH,
                 * GRAD_FILTER 1 (median of the gradients), then smooth_grad */
                io_ok(dnfwi_assemble_sh(nx, ny, P->dt, P->invmat1, P->grad_form, iter, st.inv_vs_iter, st.inv_rho_iter,
                                        vs, rho, g[1], g[2]),
                      "ass_gradSH");
                precondition(r, g);
                if (P->grad_filter == 1)
                    for (int k = 1; k < 3; k++)
                        if (dnfwi_median_filter(nx, ny, P->filt_size_grad, g[k])) die("GRAD_FILTER", dnio_last_error());
            }
            if (st.spatfilter == 3) {
                /* FWI_PSV.c:628-631: smooth_grad before precond_PSV; its width
                 * follows C_vs, which is 0 in the acoustic run (an empty filter) */
                int w = 0;
                for (int k = 0; k < 3; k++) {
                    if ((ac && k == 1) || (sh && k == 0)) continue;
                    io_ok(dnfwi_gauss_filter(nx, ny, P->dh, st.wd_damp, c_vs, st.fc_end, g[k], &w), "SPATFILTER 3");
                }
                printf(" gradients smoothed with a %d-point Gaussian (%.2f m)\n", w, w * P->dh);
            }
            if (st.spatfilter == 4) {
                /* gauss_filt_var.c with the current Vs model (FWI_PSV.c:629-631),
                 * with the Vp model in the acoustic run (FWI_AC.c:676-677) */
                for (int k = 0; k < 3; k++) {
                    if ((ac && k == 1) || (sh && k == 0)) continue;
                    io_ok(dnfwi_gauss_filter_var(nx, ny, P->dh, st.wd_damp, st.wd_damp1, st.fc_end, ac ? vp : vs,
                                                 g[k]),
                          "SPATFILTER 4");
                }
                printf(" gradients smoothed with the velocity-dependent Gaussian\n");
            }
            if (!sh) precondition(r, g); /* FWI_PSV.c:634: precond_PSV; FWI_AC.c:680: precond_AC */
            const char *suffix[3] = {"vp", "vs", "rho"};
            for (int k = 0; k < 3; k++) {
                if ((ac && k == 1) || (sh && k == 0)) continue;
                snprintf(r->name, sizeof r->name, "%s_grad_%s", P->jacobian, suffix[k]);
                write_plane(r->name, g[k], nx, ny);
            }
            float *gvs = ac ? NULL : g[1], *gvp = sh ? NULL : g[0];
            printf(" misfit (L2, test shots): %.9e\n", L2sum);
            /* FWI_PSV.c:636-676: PCG or L-BFGS search direction */
            if (lb) {
                io_ok(dnfwi_lbfgs_direction(lb, iter, c_vp, c_vs, c_rho, gvp, gvs, g[2], vp, vs, rho), "L-BFGS");
                /* extract_LBFGS_PSV.c:80-249 / extract_LBFGS_AC.c:60-178 /
                 * extract_LBFGS_SH.c:50-178: JACOBIAN_p*.old, _c*.old */
                io_ok(dnfwi_lbfgs_save(lb, P->jacobian, gvp, gvs, g[2]), "L-BFGS files");
            } else {
                if (sh) {
                    /* store_PCG_SH.c:33-64 writes the descent directions -g of
                     * vs and rho as JACOBIAN_p_u.old and _p_rho.old */
                    float *neg = xmalloc(sizeof(float) * r->ncell);
                    const char *nm[2] = {"p_u", "p_rho"};
                    for (int k = 1; k < 3; k++) {
                        for (size_t q = 0; q < r->ncell; q++) neg[q] = -g[k][q];
                        snprintf(r->name, sizeof r->name, "%s_%s.old", P->jacobian, nm[k - 1]);
                        io_ok(dnio_write_model(r->name, neg, nx, ny), "PCG files");
                    }
                    free(neg);
                }
                io_ok(dnfwi_pcg_direction(pc, P->pcg_beta, gvp, gvs, g[2], NULL), "PCG");
            }
            if (iter == 1) min_iter_help = P->min_iter; /* FWI_PSV.c:687-690 */
            obj_env oe = {r, taus, taup};
            io_ok(dnfwi_step_length(&up, iter, P->eps_scale, P->scalefac, P->stepmax, L2sum, nx, ny, vp, vs, rho, gvp,
                                    gvs, g[2], misfit_of, &oe, &ls),
                  "step length");
            /* FWI_PSV.c:700-708 */
            float eps_scale = ls.eps_scale, opteps_vp;
            if ((iter > min_iter_help) && (ls.step1 == 0)) {
                eps_scale = 0.0f;
                opteps_vp = 0.0f;
            } else {
                opteps_vp = eps_scale;
            }
            /* FWI_PSV.c:727-743: the misfit log */
            FILE *fl = fopen(P->misfit_log_file, iter_true == 1 ? "w" : "a");
            if (!fl) die("MISFIT_LOG_FILE", P->misfit_log_file);
            if (st.time_filt == 0)
                fprintf(fl, "%e \t %e \t %e \t %e \t %e \t %e \t %e \t %e \t %d \n", opteps_vp, ls.epst[0],
                        ls.epst[1], ls.epst[2], ls.L2t[0], ls.L2t[1], ls.L2t[2], L2sum, nstage);
            else
                fprintf(fl, "%e \t %e \t %e \t %e \t %e \t %e \t %e \t %e \t %f \t %f \t %d \n", opteps_vp,
                        ls.epst[0], ls.epst[1], ls.epst[2], ls.L2t[0], ls.L2t[1], ls.L2t[2], L2sum, st.fc_start,
                        r->fc, nstage);
            fclose(fl);
            printf(" step length %e (trials %e %e %e, misfits %e %e %e)\n", opteps_vp, ls.epst[0], ls.epst[1],
                   ls.epst[2], ls.L2t[0], ls.L2t[1], ls.L2t[2]);
            L2_hist[iter] = L2sum;
            /* FWI_PSV.c:750 / FWI_AC.c:782: the update, written as
             * INV_MODELFILE_{vp,vs,rho}.bin (vp and rho for the acoustic run) */
            dnfwi_model_update(&up, iter, eps_scale, nx, ny, vp, vs, rho, gvp, gvs, g[2], nv, ns, nr);
            if (vp) memcpy(vp, nv, sizeof(float) * r->ncell);
            if (vs) memcpy(vs, ns, sizeof(float) * r->ncell);
            memcpy(rho, nr, sizeof(float) * r->ncell);
            snprintf(name[0], sizeof name[0], "%s_vp.bin", P->inv_modelfile);
            snprintf(name[1], sizeof name[1], "%s_vs.bin", P->inv_modelfile);
            snprintf(name[2], sizeof name[2], "%s_rho.bin", P->inv_modelfile);
            write_models(name[0], name[1], name[2], vp, vs, rho, nx, ny);
            /* FWI_PSV.c:752, smooth_model.c:37: MODEL_FILTER 1 replaces the
             * updated models (not the files just written) by their medians */
            if (!ac && P->model_filter == 1 && P->filt_size != 0) {
                float *mods[3] = {vp, vs, rho};
                for (int k = 0; k < 3; k++)
                    if (mods[k] && dnfwi_median_filter(nx, ny, P->filt_size, mods[k]))
                        die("MODEL_FILTER", dnio_last_error());
                printf(" models smoothed with a %d-point median filter\n", abs(P->filt_size) | 1);
            }
            /* FWI_PSV.c:781-827: stage end on the misfit change over two
             * iterations (iterations before the first count as misfit 0,
             * which never ends a stage) or a failed step search */
            int stage_end = 0;
            if (iter > min_iter_help) {
                const double old = iter > 2 ? L2_hist[iter - 2] : 0.0;
                const float diff = fabs((old - L2_hist[iter]) / old);
                if ((diff <= st.pro) || (ls.step3 == 1)) stage_end = 1;
            }
            if (stage_end) {
                if (P->inv_mod_out == 0) {
                    snprintf(name[0], sizeof name[0], "%s_vp_stage_%d.bin", P->inv_modelfile, nstage);
                    snprintf(name[1], sizeof name[1], "%s_vs_stage_%d.bin", P->inv_modelfile, nstage);
                    snprintf(name[2], sizeof name[2], "%s_rho_stage_%d.bin", P->inv_modelfile, nstage);
                    write_models(name[0], name[1], name[2], vp, vs, rho, nx, ny);
                }
                min_iter_help = iter + P->min_iter;
                if (lb) dnfwi_lbfgs_reset(lb); /* FWI_PSV.c:802-811 */
                if (pc) dnfwi_pcg_reset(pc);
                printf(" %s: next FWI stage\n", ls.step3 ? "step-length search failed" : "abort criterion reached");
                break; /* FWI_PSV.c:824, before the iter_true++ of :835 */
            }
            if (P->inv_mod_out == 1) {
                snprintf(name[0], sizeof name[0], "%s_vp_stage_%d_it_%d.bin", P->inv_modelfile, nstage, iter);
                snprintf(name[1], sizeof name[1], "%s_vs_stage_%d_it_%d.bin", P->inv_modelfile, nstage, iter);
                snprintf(name[2], sizeof name[2], "%s_rho_stage_%d_it_%d.bin", P->inv_modelfile, nstage, iter);
                write_models(name[0], name[1], name[2], vp, vs, rho, nx, ny);
            }
            iter_true++;
        }
    }
    dnfwi_lbfgs_destroy(lb);
    dnfwi_pcg_destroy(pc);
    for (int k = 0; k < 3; k++) free(g[k]);
    free(nv);
    free(ns);
    free(nr);
    free(L2_hist);
}

/* MODE 2: src/PSV/RTM_PSV.c:265-385 -- workflow stage 1, one gradient pass
 * with the data as adjoint sources, the stacked P and S images
 * (RTM_PSV_out.c) or, with RTM_SHOT = 1, the per-shot images */
static void rtm(run *r, const dnio_stage *st)
{
    dnio_par *P = r->P;
    r->eprecond = (st->eprecond == 1 || st->eprecond == 3) ? st->eprecond : 0;
    r->offset_mute = st->offset_mute;
    r->timewin = st->timewin;
    r->gamma = st->gamma;
    r->twlength_minus = st->twlength_minus;
    r->twlength_plus = st->twlength_plus;
    r->offsetc = st->offsetc;
    if (P->physics == 3 || P->physics == 4) {
        r->c_vp = r->c_vs = r->c_rho = 1.0f; /* RTM_VTI.c:309-311, RTM_TTI.c:318-320 */
    } else {
        r->c_vp = dnfwi_model_mean(r->vp, P->nx, P->ny); /* RTM_PSV.c:316-352 */
        r->c_vs = dnfwi_model_mean(r->vs, P->nx, P->ny);
        r->c_rho = dnfwi_model_mean(r->rho, P->nx, P->ny);
    }
    r->time_filt = st->time_filt; /* RTM_PSV.c:283: FC = FC_END */
    r->order = st->order;
    r->fc = st->fc_end;
    r->fc_start = st->fc_start;
    float *g[3];
    for (int k = 0; k < 3; k++) g[k] = xmalloc(sizeof(float) * r->ncell);
    const double l2 = gradient_pass(r, g);
    precondition(r, g); /* RTM_PSV.c:381 */
    printf(" L2 of the data (test shots): %.9e\n", l2);
    if (P->rtm_shot == 0) {
        const char *img[2] = {"P", "S"};
        for (int k = 0; k < (P->physics == 2 ? 1 : 2); k++) { /* RTM_AC_out.c: the P image only */
            snprintf(r->name, sizeof r->name, "%s_%s_image", P->jacobian, img[k]);
            write_plane(r->name, g[k], P->nx, P->ny);
            printf(" wrote %s\n", r->name);
        }
    }
    for (int k = 0; k < 3; k++) free(g[k]);
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
    if (P.mode == 1 || P.mode == 2) {
        if (argc < 3) die("workflow", "MODE 1 and 2 need a workflow file as second argument");
        io_ok(dnio_read_stage(argv[2], 1, &stage), "workflow file");
        st = &stage;
    }
    check_supported(&P, st);
    if (P.nprocx * P.nprocy > 1)
        printf(" NPROCX x NPROCY = %d x %d: the whole grid runs as one domain on one GPU\n", P.nprocx, P.nprocy);

    run r;
    memset(&r, 0, sizeof r);
    r.P = &P;
    r.TS = P.ts;
    io_ok(dnio_read_sources(P.source_file, &r.nsrc, &r.srcpos, &r.TS), "SOURCE_FILE");
    r.nshots = P.run_multiple_shots ? r.nsrc : 1;
    printf(" %d source(s), %d shot(s); NX=%d NY=%d NT=%d DH=%g DT=%g FDORDER=%d L=%d\n", r.nsrc, r.nshots, P.nx, P.ny,
           P.nt, P.dh, P.dt, P.fdorder, P.L);

    r.ncell = (size_t)P.nx * P.ny;
    const int aniso = P.physics == 3 || P.physics == 4;
    float *vp = xmalloc(sizeof(float) * r.ncell), *vs = xmalloc(sizeof(float) * r.ncell);
    float *rho = xmalloc(sizeof(float) * r.ncell);
    const int elastic_only = aniso || P.physics == 2 || (P.physics == 5 && P.L == 0);
    float *taus = P.L && !elastic_only ? xmalloc(sizeof(float) * r.ncell) : NULL;
    float *taup = P.L && !elastic_only ? xmalloc(sizeof(float) * r.ncell) : NULL;
    float *cani[4] = {NULL, NULL, NULL, NULL}, *theta = NULL;
    if (aniso) {
        /* c11, c13, c33, c44 (+ theta for TTI); the VTI/TTI loops update the
         * stresses elastically whatever L is (VTI.c:137-149) */
        for (int k = 0; k < 4; k++) cani[k] = xmalloc(sizeof(float) * r.ncell);
        if (P.physics == 4) theta = xmalloc(sizeof(float) * r.ncell);
        io_ok(dnio_read_model_aniso(&P, rho, cani[0], cani[1], cani[2], cani[3], theta), "model");
    } else if (P.physics == 2) {
        /* vp (or the modulus for INVMAT1 3) and rho; the acoustic loop has
         * no relaxation mechanisms (ac.c) */
        io_ok(dnio_read_model_ac(&P, vp, rho), "model");
    } else if (P.physics == 5) {
        if (P.L > 0) /* vs (or mu), rho and taus from MFILE.qs */
            io_ok(dnio_read_model_sh_visc(&P, vs, rho, taus), "model");
        else
            io_ok(dnio_read_model_sh(&P, vs, rho), "model"); /* vs (or mu) and rho */
    } else {
        io_ok(dnio_read_model(&P, vp, vs, rho, taus, taup), "model");
    }
    r.vp = vp;
    r.vs = vs;
    r.rho = rho;

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
    dp.L = elastic_only ? 0 : P.L;
    dp.physics = P.physics;
    dp.seismo = P.mode >= 1 && P.seismo == 0 ? 1 : P.seismo;
    /* outseis_SHfor.c writes SH seismograms for SEISMO 1 only */
    if (P.physics == 5) dp.seismo = P.seismo == 1 || P.mode >= 1;
    dp.ndt = P.ndt;
    dp.invmat1 = P.invmat1;
    dp.grad_form = P.grad_form;
    dp.dtinv = P.dtinv > 0 ? P.dtinv : 1;
    dp.quelltypb = P.quelltypb;
    dp.mode = P.mode >= 1 ? 1 : 0; /* MODE 2 runs the gradient machinery with other adjoint sources */
    dp.dh = P.dh;
    dp.dt = P.dt;
    dp.damping = P.damping;
    dp.fpml = P.fpml;
    dp.npower = P.npower;
    dp.k_max_pml = P.k_max_pml;
    dp.ts = r.TS;
    for (int l = 0; l < P.L && l < 8; l++) dp.fl[l] = P.fl[l];

    r.dp = dp;
    r.m_vp = vp;
    r.m_vs = vs;
    r.m_rho = rho;
    r.m_taus = taus;
    r.m_taup = taup;
    r.batch = batching_applies(&r);
    dn_params dmain = dp;
    if (r.batch) dmain.mode = 0; /* the shots run in contexts of their own */
    dn_ok(dn_ctx_create(&dmain, 0, &r.ctx), "device context");
    if (aniso)
        dn_ok(dn_set_model_aniso(r.ctx, rho, cani[0], cani[1], cani[2], cani[3], theta), "model upload");
    else
        dn_ok(dn_set_model(r.ctx, vp, vs, rho, taus, taup), "model upload");
    if (P.seismo || P.mode >= 1) io_ok(dnio_read_receivers(&P, 1, &r.ntr, &r.recpos), "REC_FILE");
    r.n_order = st ? st->n_order : 0;
    if (P.mode >= 1 && P.trkill) r.kill = read_kill(P.trkill_file, r.nsrc, r.ntr);
    batch_init(&r);

    if (P.mode == 0)
        forward_pass(&r);
    else if (P.mode == 1)
        fwi(&r, argv[2], vp, vs, rho, taus, taup);
    else
        rtm(&r, st);

    printf(" time loops: %.6f s\n", g_loop_s);
    batch_free(&r);
    dn_ctx_destroy(r.ctx);
    free(r.kill);
    free(r.flag);
    free(r.picks);
    free(r.offs);
    dnio_free(r.recpos);
    dnio_free(r.srcpos);
    free(vp);
    free(vs);
    free(rho);
    free(taus);
    free(taup);
    for (int k = 0; k < 4; k++) free(cani[k]);
    free(theta);
    return 0;
}
