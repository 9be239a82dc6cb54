/*
 * Host file front end of the P/SV drop-in (include/denise_io.h).
 *
 * The formats are the reference's, so the readers keep its parsing rules
 * (scanf conversions, line counting, the one-sample shift of inseis) and the
 * writers its byte layout; each function cites the file it matches.
 */
#define _POSIX_C_SOURCE 200809L
#include "denise_io.h"

#include <complex.h>
#include <math.h>
#include <stdarg.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define PI_D 3.141592653589793 /* include/fd.h:22 */

static _Thread_local char g_err[512];

static int fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return -1;
}

const char *dnio_last_error(void) { return g_err; }
/* the error entry of the library's other host sources (dn_fwi.c) */
int dnio_fail_msg(const char *msg) { return fail("%s", msg); }
void dnio_free(void *p) { free(p); }

/* include/fd.h:17 -- iround(x) = (int)floor(x + 0.5) */
static int iround_d(double x) { return (int)floor(x + 0.5); }

/* ------------------------------------------------------------------------
 * .inp parameter file (src/read_par.c:14-456)
 *
 * A parameter line is one whose first character is not '#', found right
 * after a newline; that first character is consumed, then the line is read
 * with a "%s =<conversion>" pattern (key, '=', value).  Line k of the
 * parameter lines sets parameter k of the table below.
 * --------------------------------------------------------------------- */
enum kind { K_INT, K_FLOAT, K_STR, K_FL, K_REFREC, K_INT4, K_INT3, K_L };

struct entry {
    enum kind kind;
    size_t off;
};

#define E(k, field) {k, offsetof(dnio_par, field)}
static const struct entry k_table[] = {
    E(K_INT, mode), E(K_INT, physics), E(K_INT, nprocx), E(K_INT, nprocy), E(K_INT, fdorder),
    E(K_INT, maxrelerror), E(K_INT, nx), E(K_INT, ny), E(K_FLOAT, dh), E(K_FLOAT, time),
    E(K_FLOAT, dt), E(K_INT, quellart), E(K_STR, signal_file), E(K_FLOAT, ts), E(K_INT, srcrec),
    E(K_STR, source_file), E(K_INT, run_multiple_shots), E(K_FLOAT, fc_spike_1),
    E(K_FLOAT, fc_spike_2), E(K_INT, order_spike), E(K_INT, write_stf), E(K_INT, readmod),
    E(K_STR, mfile), E(K_INT, writemod), E(K_L, L), E(K_FL, fl), E(K_FLOAT, tau),
    E(K_INT, free_surf), E(K_INT, fw), E(K_FLOAT, damping), E(K_FLOAT, fpml), E(K_FLOAT, npower),
    E(K_FLOAT, k_max_pml), E(K_INT, boundary), E(K_INT, snap), E(K_INT, snap_shot),
    E(K_FLOAT, tsnap1), E(K_FLOAT, tsnap2), E(K_FLOAT, tsnapinc), E(K_INT, idx), E(K_INT, idy),
    E(K_INT, snap_format), E(K_STR, snap_file), E(K_INT, seismo), E(K_INT, readrec),
    E(K_STR, rec_file), E(K_REFREC, refrec), E(K_INT, n_streamer), E(K_FLOAT, rec_incr_x),
    E(K_FLOAT, rec_incr_y), E(K_INT, ndt), E(K_INT, seis_format), E(K_STR, seis_file_vx),
    E(K_STR, seis_file_vy), E(K_STR, seis_file_curl), E(K_STR, seis_file_div),
    E(K_STR, seis_file_p), E(K_STR, log_file), E(K_INT, log), E(K_INT, itermax),
    E(K_STR, jacobian), E(K_STR, data_dir), E(K_INT, taper), E(K_INT, taperlength),
    E(K_INT4, gradt), E(K_INT, invmat1), E(K_INT, grad_form), E(K_INT, quelltypb),
    E(K_INT3, testshot_start), E(K_INT, sws_taper_grad_vert), E(K_INT, sws_taper_grad_hor),
    E(K_FLOAT, exp_taper_grad_hor), E(K_INT, sws_taper_grad_sources),
    E(K_INT, sws_taper_circular_per_shot), E(K_INT, srtshape), E(K_FLOAT, srtradius),
    E(K_INT, filtsize), E(K_INT, sws_taper_file), E(K_STR, tfile), E(K_INT, inv_mod_out),
    E(K_STR, inv_modelfile), E(K_FLOAT, vpupperlim), E(K_FLOAT, vplowerlim),
    E(K_FLOAT, vsupperlim), E(K_FLOAT, vslowerlim), E(K_FLOAT, rhoupperlim),
    E(K_FLOAT, rholowerlim), E(K_FLOAT, qsupperlim), E(K_FLOAT, qslowerlim),
    E(K_INT, grad_method), E(K_INT, pcg_beta), E(K_INT, nlbfgs), E(K_INT, model_filter),
    E(K_INT, filt_size), E(K_INT, dtinv), E(K_FLOAT, eps_scale), E(K_INT, stepmax),
    E(K_FLOAT, scalefac), E(K_INT, trkill), E(K_STR, trkill_file), E(K_STR, picks_file),
    E(K_STR, misfit_log_file), E(K_INT, min_iter), E(K_INT, grad_filter),
    E(K_INT, filt_size_grad), E(K_INT, timelapse), E(K_STR, data_dir_t0), E(K_INT, rtmod),
    E(K_INT, gravity), E(K_INT, ngravb), E(K_INT, nzgrav), E(K_INT, grav_type),
    E(K_INT, back_density), E(K_STR, dfile), E(K_INT, rtm_shot),
};
#undef E
#define NPAR ((int)(sizeof k_table / sizeof k_table[0]))

/* defaults of the reference globals that the file may leave unset */
static void par_defaults(dnio_par *p)
{
    memset(p, 0, sizeof *p);
    p->ndt = 1;
    p->idx = p->idy = 1;
    p->dtinv = 1;
    p->invmat1 = 1;
    p->grad_form = 1;
    p->quelltypb = 1;
    p->seis_format = 1;
}

static int read_one(FILE *f, const struct entry *e, dnio_par *p)
{
    char key[1024];
    char *base = (char *)p + e->off;
    switch (e->kind) {
    case K_INT:
        return fscanf(f, "%1023s =%i", key, (int *)base);
    case K_FLOAT:
        return fscanf(f, "%1023s =%f", key, (float *)base);
    case K_STR:
        return fscanf(f, "%1023s =%255s", key, base);
    case K_L:
        return fscanf(f, "%1023s =%i", key, (int *)base);
    case K_FL: {
        /* FL[1] after the key, then FL[2..L] as bare numbers */
        float v = 0.0f;
        int r = fscanf(f, "%1023s =%f", key, &v);
        if (p->L >= 1) p->fl[0] = v;
        for (int l = 2; l <= p->L; l++) {
            if (fscanf(f, "%f", &v) != 1) break;
            if (l <= DNIO_MAXL) p->fl[l - 1] = v;
        }
        return r;
    }
    case K_REFREC:
        return fscanf(f, "%1023s =%f ,%f", key, &p->refrec[1], &p->refrec[2]);
    case K_INT4:
        return fscanf(f, "%1023s =%i, %i, %i, %i", key, &p->gradt[0], &p->gradt[1], &p->gradt[2],
                      &p->gradt[3]);
    case K_INT3:
        return fscanf(f, "%1023s =%i, %i, %i", key, &p->testshot_start, &p->testshot_end,
                      &p->testshot_incr);
    }
    return 0;
}

int dnio_read_par(const char *path, dnio_par *p)
{
    if (!path || !p) return fail("null argument");
    FILE *f = fopen(path, "r");
    if (!f) return fail("cannot open parameter file %s", path);
    par_defaults(p);
    int c, lineno = 0;
    while ((c = getc(f)) != EOF) {
        if (c != '\n') continue;
        if (getc(f) == '#') continue;
        lineno++;
        if (lineno <= NPAR) (void)read_one(f, &k_table[lineno - 1], p);
    }
    fclose(f);
    if (p->L < 0 || p->L > DNIO_MAXL) return fail("L = %d: at most %d relaxation mechanisms", p->L, DNIO_MAXL);
    if (p->fw < 0) p->fw = 0; /* read_par.c:186 */
    p->lines = lineno;
    if (p->dt > 0.0f) p->nt = iround_d((double)(p->time / p->dt)); /* FD_PSV.c:68 */
    return 0;
}

/* ------------------------------------------------------------------------
 * FWI workflow (src/read_par_inv.c:32-35)
 * --------------------------------------------------------------------- */
int dnio_read_stage(const char *path, int stage, dnio_stage *s)
{
    if (!path || !s || stage < 1) return fail("bad workflow request");
    FILE *f = fopen(path, "r");
    if (!f) return fail("cannot open workflow file %s", path);
    char w[256];
    for (int k = 0; k < 30; k++)
        if (fscanf(f, "%255s", w) != 1) {
            fclose(f);
            return fail("workflow file %s: header shorter than 30 words", path);
        }
    memset(s, 0, sizeof *s);
    for (int i = 1; i <= stage; i++) {
        int n = fscanf(f, "%f%i%f%f%i%i%f%f%f%i%i%i%i%i%f%f%i%i%i%i%f%f%i%i%f%f%f%i%f%i", &s->pro,
                       &s->time_filt, &s->fc_start, &s->fc_end, &s->order, &s->timewin, &s->gamma,
                       &s->twlength_minus, &s->twlength_plus, &s->inv_vp_iter, &s->inv_vs_iter,
                       &s->inv_rho_iter, &s->inv_qs_iter, &s->spatfilter, &s->wd_damp, &s->wd_damp1,
                       &s->eprecond, &s->lnorm, &s->rowi, &s->inv_stf, &s->offsetc_stf, &s->eps_stf,
                       &s->normalize, &s->offset_mute, &s->offsetc, &s->scalerho, &s->scaleqs, &s->env,
                       &s->gamma_grav, &s->n_order);
        if (n != 30) {
            fclose(f);
            return fail("workflow file %s: stage %d has %d of 30 values", path, i, n < 0 ? 0 : n);
        }
    }
    fclose(f);
    return 0;
}

int dnio_count_stages(const char *path, int *nstages)
{
    if (!path || !nstages) return fail("null argument");
    FILE *f = fopen(path, "r");
    if (!f) return fail("cannot open workflow file %s", path);
    char line[4096];
    int n = 0, first = 1;
    while (fgets(line, sizeof line, f)) {
        char *q = line;
        while (*q == ' ' || *q == '\t' || *q == '\r' || *q == '\n') q++;
        if (!*q) continue;
        if (first) {
            first = 0;
            continue;
        }
        n++;
    }
    fclose(f);
    *nstages = n;
    return 0;
}

/* ------------------------------------------------------------------------
 * sources (src/sources.c:9-127)
 * --------------------------------------------------------------------- */
int dnio_read_sources(const char *path, int *nsrc, float **srcpos, float *ts)
{
    if (!path || !nsrc || !srcpos) return fail("null argument");
    *srcpos = NULL;
    FILE *f = fopen(path, "r");
    if (!f) return fail("source file %s could not be opened", path);
    int n = 0;
    if (fscanf(f, "%d", &n) != 1 || n < 1) {
        fclose(f);
        return fail("source file %s: first line must give the number of sources", path);
    }
    rewind(f);
    char line[256];
    if (!fgets(line, 255, f)) {
        fclose(f);
        return fail("source file %s: unreadable", path);
    }
    float *t = calloc((size_t)8 * n, sizeof(float));
    if (!t) {
        fclose(f);
        return fail("out of memory");
    }
    float fcmax = 0.0f;
    for (int l = 0; l < n; l++) {
        float x = 0, z = 0, y = 0, tshift = 0, v[4] = {0, 0, 0, 0};
        int got = 0;
        if (fgets(line, 255, f)) got = sscanf(line, "%f%f%f%f%f%f%f%f", &x, &z, &y, &tshift, &v[0], &v[1], &v[2], &v[3]);
        if (got < 6) {
            free(t);
            fclose(f);
            return fail("Missing parameter in SOURCE_FILE %s (source %d has %d values, 6 needed)", path, l + 1,
                        got < 0 ? 0 : got);
        }
        if (got == 6) v[2] = 0.0f; /* azimuth */
        (void)z;
        t[0 * n + l] = x;
        t[1 * n + l] = y;
        t[2 * n + l] = 0.0f;
        t[3 * n + l] = tshift;
        t[4 * n + l] = v[0];
        t[5 * n + l] = v[1];
        t[6 * n + l] = v[2];
        t[7 * n + l] = v[3]; /* type; 0 when the column is missing */
    }
    fclose(f);
    fcmax = t[4 * n + n - 1]; /* sources.c: fc holds the last source's value first */
    for (int l = 0; l < n; l++)
        if (t[4 * n + l] > fcmax) fcmax = t[4 * n + l];
    if (!(fcmax > 0.0f)) {
        free(t);
        return fail("source file %s: no positive centre frequency", path);
    }
    *nsrc = n;
    *srcpos = t;
    if (ts) *ts = (float)(1.0 / fcmax);
    return 0;
}

/* splitsrc (src/splitsrc.c:9-60) with one subdomain (IENDX = NX,
 * IENDY = NY, POS = (0, 0)); C integer division and remainder. */
int dnio_split_sources(const dnio_par *p, const float *srcpos, int nsrc, int first, int count, float *loc,
                       int *nloc)
{
    if (!p || !srcpos || !loc || !nloc || first < 0 || count < 1 || first + count > nsrc)
        return fail("bad source selection");
    int k = 0;
    for (int j = first; j < first + count; j++) {
        const int ix = iround_d((double)(srcpos[0 * nsrc + j] / p->dh));
        const int iy = iround_d((double)(srcpos[1 * nsrc + j] / p->dh));
        if ((ix - 1) / p->nx != 0 || (iy - 1) / p->ny != 0) continue;
        loc[0 * count + k] = (float)(((ix - 1) % p->nx) + 1);
        loc[1 * count + k] = (float)(((iy - 1) % p->ny) + 1);
        loc[2 * count + k] = (float)(j + 1);
        for (int r = 3; r < 8; r++) loc[r * count + k] = srcpos[r * nsrc + j];
        k++;
    }
    *nloc = k;
    return 0;
}

/* ------------------------------------------------------------------------
 * receiver (src/receiver.c:9-121)
 * --------------------------------------------------------------------- */
int dnio_read_receivers(const dnio_par *p, int ishot, int *ntr, int **recpos)
{
    if (!p || !ntr || !recpos) return fail("null argument");
    *recpos = NULL;
    char name[2 * DNIO_STR];
    if (p->readrec == 2)
        snprintf(name, sizeof name, "%s_shot_%i.dat", p->rec_file, ishot);
    else
        snprintf(name, sizeof name, "%s.dat", p->rec_file);
    FILE *f = fopen(name, "r");
    if (!f) return fail("Receiver file %s could not be opened", name);
    int n = 0, c;
    while ((c = fgetc(f)) != EOF)
        if (c == '\n') ++n;
    rewind(f);
    int *r1 = malloc(sizeof(int) * 3 * (size_t)(n > 0 ? n : 1));
    if (!r1) {
        fclose(f);
        return fail("out of memory");
    }
    for (int k = 0; k < n; k++) {
        float x = 0.0f, y = 0.0f;
        if (fscanf(f, "%f%f\n", &x, &y) != 2) {
            free(r1);
            fclose(f);
            return fail("receiver file %s: line %d is not 'x y'", name, k + 1);
        }
        r1[0 * n + k] = iround_d((double)((x + p->refrec[1]) / p->dh));
        r1[1 * n + k] = iround_d((double)((y + p->refrec[2]) / p->dh));
        r1[2 * n + k] = iround_d((double)((0.0f + p->refrec[3]) / p->dh));
    }
    fclose(f);
    /* duplicates: later copies get a negative x index (receiver.c:73-79) */
    int flag = 0;
    for (int a = 0; a < n - 1; a++)
        for (int b = a + 1; b < n; b++)
            if (r1[a] == r1[b] && r1[n + a] == r1[n + b] && r1[2 * n + a] == r1[2 * n + b]) r1[b] = -(++flag);
    int m = 0;
    for (int a = 0; a < n; a++)
        if (r1[a] > 0) m++;
    int *r = malloc(sizeof(int) * 3 * (size_t)(m > 0 ? m : 1));
    if (!r) {
        free(r1);
        return fail("out of memory");
    }
    int k = 0;
    for (int a = 0; a < n; a++)
        if (r1[a] > 0) {
            r[0 * m + k] = r1[a];
            r[1 * m + k] = r1[n + a];
            r[2 * m + k] = r1[2 * n + a];
            k++;
        }
    free(r1);
    *ntr = m;
    *recpos = r;
    return 0;
}

/* ------------------------------------------------------------------------
 * wavelet (src/wavelet.c:10-121); float variables, double literals, as
 * the reference evaluates them
 * --------------------------------------------------------------------- */
static int read_signal_file(const char *name, int nt, float *out)
{
    FILE *f = fopen(name, "r");
    if (!f) return fail("Source file %s could not be opened", name);
    int n = 0, c;
    while ((c = fgetc(f)) != EOF)
        if (c == '\n') ++n;
    rewind(f);
    if (n < nt) {
        fclose(f);
        return fail("source signal file %s has %d samples, NT = %d", name, n, nt);
    }
    for (int i = 0; i < nt; i++)
        if (fscanf(f, "%e\n", &out[i]) != 1) {
            fclose(f);
            return fail("source signal file %s: bad sample %d", name, i + 1);
        }
    fclose(f);
    return 0;
}

int dnio_wavelet(const dnio_par *p, const float *loc, int nsrc, int ishot, float TS, int n_order, float *signals)
{
    if (!p || (nsrc > 0 && (!loc || !signals))) return fail("null argument");
    const int NT = p->nt;
    const float DT = p->dt;
    float *file = NULL;
    if (p->quellart == 3) {
        char name[2 * DNIO_STR];
        snprintf(name, sizeof name, "%s_shot_%i.dat", p->signal_file, ishot);
        file = malloc(sizeof(float) * (size_t)NT);
        if (!file) return fail("out of memory");
        if (read_signal_file(name, NT, file)) {
            free(file);
            return -1;
        }
    }
    if (p->quellart < 1 || p->quellart > 7) {
        free(file);
        return fail("Which source-wavelet ? (QUELLART = %d)", p->quellart);
    }
    float k1 = 0.0f, f0 = 0.0f;
    if (p->quellart == 7) {
        k1 = (p->fc_spike_2 - p->fc_spike_1) / TS;
        f0 = (float)((p->fc_spike_2 + p->fc_spike_1) / 2.0);
    }
    for (int nt = 1; nt <= NT; nt++) {
        const float t = (float)nt * DT;
        for (int k = 0; k < nsrc; k++) {
            const float tshift = loc[3 * nsrc + k], fc = loc[4 * nsrc + k], a = loc[5 * nsrc + k];
            float ts = (float)(1.0 / fc), ag, tau, amp = 0.0f, t1;
            switch (p->quellart) {
            case 1:
                tau = (float)(PI_D * (t - 1.5 * ts - tshift) / (1.5 * ts));
                amp = (float)((1.0 - 4.0 * tau * tau) * exp(-2.0 * tau * tau));
                break;
            case 2:
                if ((t < tshift) || (t > (tshift + ts)))
                    amp = 0.0f;
                else
                    amp = (float)(sin(2.0 * PI_D * (t - tshift) * fc) - 0.5 * sin(4.0 * PI_D * (t - tshift) * fc));
                break;
            case 3:
                amp = file[nt - 1];
                break;
            case 4:
                if ((t < tshift) || (t > (tshift + ts)))
                    amp = 0.0f;
                else
                    amp = (float)pow(sin(PI_D * (t + tshift) / ts), 3.0);
                break;
            case 5:
                ts = (float)(1.2 / fc);
                ag = (float)(PI_D * PI_D * fc * fc);
                amp = (float)(-2.0 * ag * (t - ts) * exp(-ag * (t - ts) * (t - ts)));
                break;
            case 6:
                amp = 0.0f;
                if (nt == 1 + iround_d((double)(tshift / DT))) amp = 1.0f;
                break;
            case 7: {
                ts = (float)(1.5 / p->fc_spike_1);
                t1 = t - ts - tshift;
                double complex e = cexp(0.0 + 2.0 * PI_D * f0 * t1 * I);
                amp = (float)creal(e * sin(PI_D * k1 * t1 * (TS - t1)) / (PI_D * k1 * t1));
                break;
            }
            }
            signals[(size_t)k * NT + (nt - 1)] = amp * a;
        }
    }
    free(file);
    for (int i = 1; i <= n_order; i++)
        for (int k = 0; k < nsrc; k++) {
            float *s = signals + (size_t)k * NT;
            s[0] = s[0] * DT;
            for (int nt = 2; nt <= NT; nt++) s[nt - 1] = s[nt - 2] + s[nt - 1] * DT;
        }
    return 0;
}

/* ------------------------------------------------------------------------
 * models (src/PSV/readmod_elastic_PSV.c:9-129, readmod_visc_PSV.c:9-162,
 * src/writemod.c, src/mergemod.c)
 * --------------------------------------------------------------------- */
static int write_floats(const char *name, const float *a, size_t n)
{
    FILE *f = fopen(name, "wb");
    if (!f) return fail("cannot write %s", name);
    size_t w = fwrite(a, sizeof(float), n, f);
    fclose(f);
    return w == n ? 0 : fail("short write to %s", name);
}

int dnio_write_model(const char *name, const float *a, int nx, int ny)
{
    if (!name || !a || nx < 1 || ny < 1) return fail("bad model write request");
    char part[2 * DNIO_STR + 8];
    snprintf(part, sizeof part, "%s.0.0", name);
    if (write_floats(part, a, (size_t)nx * ny)) return -1;
    return write_floats(name, a, (size_t)nx * ny);
}

static int read_model_file(const char *base, const char *ext, float *out, size_t n)
{
    char name[2 * DNIO_STR];
    snprintf(name, sizeof name, "%s.%s", base, ext);
    FILE *f = fopen(name, "rb");
    if (!f) return fail("Could not open model file %s", name);
    size_t r = out ? fread(out, sizeof(float), n, f) : n;
    fclose(f);
    if (r != n) return fail("model file %s holds %zu of %zu values", name, r, n);
    return 0;
}

int dnio_read_model(const dnio_par *p, float *vp, float *vs, float *rho, float *taus, float *taup)
{
    if (!p || !vp || !vs || !rho) return fail("null argument");
    if (!p->readmod) return fail("READMOD = 0 (built-in model functions) is not supported; give model files");
    if (p->invmat1 != 1 && p->invmat1 != 3) return fail("INVMAT1 = %d: 1 or 3 expected", p->invmat1);
    const size_t n = (size_t)p->nx * p->ny;
    const char *e1 = p->invmat1 == 1 ? "vp" : "lam", *e2 = p->invmat1 == 1 ? "vs" : "mu";
    if (read_model_file(p->mfile, e1, vp, n) || read_model_file(p->mfile, e2, vs, n) ||
        read_model_file(p->mfile, "rho", rho, n))
        return -1;
    char name[2 * DNIO_STR];
    if (p->L > 0) {
        if (!taus || !taup) return fail("L > 0 needs taus and taup outputs");
        /* .qp / .qs must exist; their values are replaced by Q = 30
         * (readmod_visc_PSV.c:106-107) */
        if (read_model_file(p->mfile, "qp", NULL, n) || read_model_file(p->mfile, "qs", NULL, n)) return -1;
        const float q = 30.0f;
        for (size_t k = 0; k < n; k++) {
            taus[k] = (float)(2.0 / q);
            taup[k] = (float)(2.0 / q);
        }
        const char *ext[3] = {"fdveps.pi", "fdveps.mu", "fdveps.rho"};
        const float *arr[3] = {vp, vs, rho};
        for (int k = 0; k < 3; k++) {
            snprintf(name, sizeof name, "%s.%s", p->mfile, ext[k]);
            if (dnio_write_model(name, arr[k], p->nx, p->ny)) return -1;
        }
    } else if (p->writemod) {
        const char *ext[3] = {"denise.pi", "denise.mu", "denise.rho"};
        const float *arr[3] = {vp, vs, rho};
        for (int k = 0; k < 3; k++) {
            snprintf(name, sizeof name, "%s.%s", p->mfile, ext[k]);
            if (dnio_write_model(name, arr[k], p->nx, p->ny)) return -1;
        }
    }
    return 0;
}

/* VTI (PHYSICS 3) and TTI (PHYSICS 4) models: readmod_elastic_VTI.c:27-79 /
 * readmod_elastic_TTI.c:27-85 read MFILE.c11, .c13, .c33, .c44, (.theta),
 * .rho in the same x-major order and write the copies MFILE.denise.c11,
 * .c13, .c33, .c44 (VTI: merged file and the one-rank part .0.0,
 * readmod_elastic_VTI.c:96-118) or, for TTI, also .theta in radians, with
 * the parts removed after the merge (readmod_elastic_TTI.c:96-151). */
int dnio_read_model_aniso(const dnio_par *p, float *rho, float *c11, float *c13, float *c33, float *c44,
                          float *theta)
{
    if (!p || !rho || !c11 || !c13 || !c33 || !c44) return fail("null argument");
    if (p->physics != 3 && p->physics != 4) return fail("PHYSICS = %d: 3 (VTI) or 4 (TTI) expected", p->physics);
    if (p->physics == 4 && !theta) return fail("TTI needs a theta output");
    if (!p->readmod) return fail("READMOD = 0 (built-in model functions) is not supported; give model files");
    const size_t n = (size_t)p->nx * p->ny;
    if (read_model_file(p->mfile, "c11", c11, n) || read_model_file(p->mfile, "c13", c13, n) ||
        read_model_file(p->mfile, "c33", c33, n) || read_model_file(p->mfile, "c44", c44, n) ||
        (p->physics == 4 && read_model_file(p->mfile, "theta", theta, n)) ||
        read_model_file(p->mfile, "rho", rho, n))
        return -1;
    const char *ext[5] = {"denise.c11", "denise.c13", "denise.c33", "denise.c44", "denise.theta"};
    float *rad = NULL;
    if (p->physics == 4) {
        rad = (float *)malloc(sizeof(float) * n);
        if (!rad) return fail("out of host memory");
        for (size_t k = 0; k < n; k++) rad[k] = (float)(theta[k] * 3.141592653589793 / 180.0);
    }
    const float *arr[5] = {c11, c13, c33, c44, rad};
    char name[2 * DNIO_STR], part[2 * DNIO_STR + 8];
    int rc = 0;
    for (int k = 0; k < (p->physics == 4 ? 5 : 4) && !rc; k++) {
        snprintf(name, sizeof name, "%s.%s", p->mfile, ext[k]);
        rc = dnio_write_model(name, arr[k], p->nx, p->ny);
        if (!rc && p->physics == 4) {
            snprintf(part, sizeof part, "%s.0.0", name);
            remove(part);
        }
    }
    free(rad);
    return rc;
}

/* Acoustic models (PHYSICS 2): readmod_AC.c:27-93 reads MFILE.vp (INVMAT1 1)
 * or MFILE.lam (INVMAT1 3) and MFILE.rho in the same x-major order and
 * always writes the copies MFILE.fdveps.pi and .fdveps.rho (merged file and
 * the one-rank part .0.0). */
int dnio_read_model_ac(const dnio_par *p, float *vp, float *rho)
{
    if (!p || !vp || !rho) return fail("null argument");
    if (p->physics != 2) return fail("PHYSICS = %d: 2 (acoustic) expected", p->physics);
    if (!p->readmod) return fail("READMOD = 0 (built-in model functions) is not supported; give model files");
    if (p->invmat1 != 1 && p->invmat1 != 3) return fail("INVMAT1 = %d: 1 or 3 expected", p->invmat1);
    const size_t n = (size_t)p->nx * p->ny;
    if (read_model_file(p->mfile, p->invmat1 == 1 ? "vp" : "lam", vp, n) || read_model_file(p->mfile, "rho", rho, n))
        return -1;
    const char *ext[2] = {"fdveps.pi", "fdveps.rho"};
    const float *arr[2] = {vp, rho};
    char name[2 * DNIO_STR];
    for (int k = 0; k < 2; k++) {
        snprintf(name, sizeof name, "%s.%s", p->mfile, ext[k]);
        if (dnio_write_model(name, arr[k], p->nx, p->ny)) return -1;
    }
    return 0;
}

/* SH models (PHYSICS 5): readmod_elastic_SH.c:21-98 reads MFILE.vs (INVMAT1
 * 1) or MFILE.mu (INVMAT1 3) and MFILE.rho in the same x-major order and,
 * with WRITEMOD, writes MFILE.denise.mu and .denise.rho (merged file and the
 * one-rank part .0.0). */
int dnio_read_model_sh(const dnio_par *p, float *vs, float *rho)
{
    if (!p || !vs || !rho) return fail("null argument");
    if (p->physics != 5) return fail("PHYSICS = %d: 5 (SH) expected", p->physics);
    if (!p->readmod) return fail("READMOD = 0 (built-in model functions) is not supported; give model files");
    if (p->invmat1 != 1 && p->invmat1 != 3) return fail("INVMAT1 = %d: 1 or 3 expected", p->invmat1);
    const size_t n = (size_t)p->nx * p->ny;
    if (read_model_file(p->mfile, p->invmat1 == 1 ? "vs" : "mu", vs, n) || read_model_file(p->mfile, "rho", rho, n))
        return -1;
    if (!p->writemod) return 0;
    const char *ext[2] = {"denise.mu", "denise.rho"};
    const float *arr[2] = {vs, rho};
    char name[2 * DNIO_STR];
    for (int k = 0; k < 2; k++) {
        snprintf(name, sizeof name, "%s.%s", p->mfile, ext[k]);
        if (dnio_write_model(name, arr[k], p->nx, p->ny)) return -1;
    }
    return 0;
}

int dnio_read_model_sh_visc(const dnio_par *p, float *vs, float *rho, float *taus)
{
    if (!p || !taus) return fail("null argument");
    if (dnio_read_model_sh(p, vs, rho)) return -1;
    /* readmod_visc_SH.c:45-84: MFILE.qs, taus = 2 / Qs */
    const size_t n = (size_t)p->nx * p->ny;
    if (read_model_file(p->mfile, "qs", taus, n)) return -1;
    for (size_t k = 0; k < n; k++) taus[k] = 2.0 / taus[k];
    return 0;
}

/* ------------------------------------------------------------------------
 * seismograms (src/outseis_glob.c:8-303, src/inseis.c:8-110,
 * include/segy.h:429-710)
 *
 * The SU trace header is 240 bytes, native endian, SEG-Y field offsets; the
 * fields outseis_glob sets are written at their offsets, all others are 0.
 * --------------------------------------------------------------------- */
static void put_i32(unsigned char *h, int off, int32_t v) { memcpy(h + off, &v, 4); }
static void put_i16(unsigned char *h, int off, int16_t v) { memcpy(h + off, &v, 2); }
static void put_u16(unsigned char *h, int off, uint16_t v) { memcpy(h + off, &v, 2); }
static void put_f32(unsigned char *h, int off, float v) { memcpy(h + off, &v, 4); }

enum {
    SU_TRACL = 0, SU_EP = 16, SU_CDP = 20, SU_TRID = 28, SU_OFFSET = 36, SU_GELEV = 40, SU_SDEPTH = 48,
    SU_SWDEP = 60, SU_SCALEL = 68, SU_SCALCO = 70, SU_SX = 72, SU_SY = 76, SU_GX = 80, SU_GY = 84,
    SU_NS = 114, SU_DT = 116, SU_D1 = 180
};

int dnio_write_seis(const char *path, const dnio_par *p, const float *section, int ntr, int nt,
                    const int *recpos, float XS, float YS)
{
    if (!path || !p || !section || !recpos || ntr < 1 || nt < 1) return fail("bad seismogram write request");
    const int NDT = p->ndt > 0 ? p->ndt : 1;
    const int ns2 = (int)floor((double)(nt / NDT) + 0.5); /* iround(nt/NDT), integer division inside */
    FILE *f = fopen(path, p->seis_format == 2 ? "w" : "wb");
    if (!f) return fail("cannot write seismograms to %s", path);
    const float DH = p->dh;
    const float xshift = 800.0f, yshift = 800.0f;
    const float drec = ntr > 1 ? (float)((float)recpos[1] * DH - (float)recpos[0] * DH) : 0.0f;
    unsigned char h[240];
    float *buf = malloc(sizeof(float) * (size_t)(ns2 > 0 ? ns2 : 1));
    if (!buf) {
        fclose(f);
        return fail("out of memory");
    }
    int werr = 0;
    for (int k = 0; k < ntr && !werr; k++) {
        const float *s = section + (size_t)k * nt;
        int lsamp = 0;
        for (int j = 0; j < ns2; j++) {
            lsamp += NDT;
            buf[j] = s[lsamp - 1];
        }
        if (p->seis_format == 1) {
            const float xr = (float)recpos[k] * DH, yr = (float)recpos[ntr + k] * DH;
            const float x = xr - p->refrec[1], y = yr - p->refrec[2];
            memset(h, 0, sizeof h);
            put_i32(h, SU_TRACL, recpos[2 * ntr + k]);
            put_i32(h, SU_EP, 0);
            /* 100 + iround(|xr+XS|/2 / drec); a single trace has no spacing */
            const int cdp = drec != 0.0f ? 100 + (int)floor((fabs((double)(xr + XS)) / 2.0) / drec + 0.5) : 100;
            put_i32(h, SU_CDP, cdp);
            put_i16(h, SU_TRID, 1);
            put_i32(h, SU_OFFSET, iround_d(sqrt((double)((XS - xr) * (XS - xr) + (YS - yr) * (YS - yr))) * 1000.0));
            put_i32(h, SU_GELEV, iround_d((double)yr * 1000.0));
            put_i32(h, SU_SDEPTH, iround_d((double)YS * 1000.0));
            put_i32(h, SU_SWDEP, iround_d(((360.0 / (2.0 * PI_D)) * atan2((double)(x - xshift), (double)(y - yshift))) * 1000.0));
            put_i16(h, SU_SCALEL, -1000);
            put_i16(h, SU_SCALCO, -1000);
            put_i32(h, SU_SX, iround_d((double)XS * 1000.0));
            put_i32(h, SU_GX, iround_d((double)xr * 1000.0));
            put_u16(h, SU_NS, (uint16_t)ns2);
            const uint16_t dtus = (uint16_t)iround_d((double)((float)NDT * p->dt) * 1.0e6);
            put_u16(h, SU_DT, dtus);
            put_f32(h, SU_D1, (float)((float)dtus * 1.0e-6));
            put_i32(h, SU_SY, iround_d((double)YS * 1000.0));
            put_i32(h, SU_GY, iround_d((double)yr * 1000.0));
            if (fwrite(h, 1, 240, f) != 240 || fwrite(buf, sizeof(float), (size_t)ns2, f) != (size_t)ns2) werr = 1;
        } else if (p->seis_format == 2) {
            for (int j = 0; j < ns2; j++)
                if (fprintf(f, "%e\n", buf[j]) < 0) werr = 1;
        } else if (p->seis_format == 3) {
            if (fwrite(buf, sizeof(float), (size_t)ns2, f) != (size_t)ns2) werr = 1;
        } else {
            werr = 2;
        }
    }
    free(buf);
    if (fclose(f) != 0 && !werr) werr = 1;
    if (werr == 2) return fail("Don't know data format %d for seismograms", p->seis_format);
    if (werr) return fail("write error on %s", path);
    return 0;
}

int dnio_read_su(const char *path, int ntr, int ns, float *section)
{
    if (!path || !section || ntr < 1 || ns < 1) return fail("bad SU read request");
    FILE *f = fopen(path, "rb");
    if (!f) return fail("Field data file %s could not be opened", path);
    unsigned char h[240];
    float *buf = malloc(sizeof(float) * (size_t)ns);
    if (!buf) {
        fclose(f);
        return fail("out of memory");
    }
    for (int k = 0; k < ntr; k++) {
        if (fread(h, 1, 240, f) != 240 || fread(buf, sizeof(float), (size_t)ns, f) != (size_t)ns) {
            free(buf);
            fclose(f);
            return fail("%s: trace %d of %d incomplete (%d samples expected)", path, k + 1, ntr, ns);
        }
        float *s = section + (size_t)k * ns;
        s[0] = 0.0f;
        for (int j = 1; j < ns; j++) s[j] = buf[j];
    }
    free(buf);
    fclose(f);
    return 0;
}

int dnio_seis_name(const dnio_par *p, int comp, int ishot, int iter, char *out, int n)
{
    if (!p || !out || comp < 0 || comp > 4) return fail("bad seismogram name request");
    const char *base[5] = {p->seis_file_vx, p->seis_file_vy, p->seis_file_p, p->seis_file_div, p->seis_file_curl};
    if (p->mode > 0)
        snprintf(out, (size_t)n, "%s.shot%d.it%d", base[comp], ishot, iter);
    else
        snprintf(out, (size_t)n, "%s.shot%d", base[comp], ishot);
    return 0;
}

int dnio_data_name(const dnio_par *p, int comp, int ishot, char *out, int n)
{
    if (!p || !out || comp < 0 || comp > 2) return fail("bad data name request");
    const char *suf[3] = {"_x", "_y", "_p"};
    snprintf(out, (size_t)n, "%s%s.su.shot%d", p->data_dir, suf[comp], ishot);
    return 0;
}

/* ------------------------------------------------------------------------
 * Wavefield snapshots (src/snap_ssg.c:10-184, src/writedsk.c:15-43)
 * ---------------------------------------------------------------------- */
static int snap_put(FILE *f, float amp, int format)
{
    if (format == 2) return fprintf(f, "%e\n", amp) < 0;
    return fwrite(&amp, sizeof(float), 1, f) != 1;
}

int dnio_write_snap(const dnio_par *p, int nsnap, const float *vx, const float *vy, const float *sxx,
                    const float *syy, const float *ppi, const float *pu, const float *hc)
{
    if (!p || !vx || !vy || !sxx || !syy || !ppi || !pu || !hc) return fail("dnio_write_snap: null argument");
    if (p->snap_format != 2 && p->snap_format != 3)
        return fail(" Sorry, SU-format for snapshots not implemented yet. (SNAP_FORMAT = %d)", p->snap_format);
    if (p->snap < 1 || p->snap > 4) return fail("dnio_write_snap: SNAP = %d", p->snap);
    const int NX = p->nx, NY = p->ny, nd = p->fdorder / 2 + 1, fdoh = p->fdorder / 2;
    const int idx = p->idx > 0 ? p->idx : 1, idy = p->idy > 0 ? p->idy : 1;
    const size_t ncol = (size_t)NX + 2 * nd;
#define NR(a, j, i) (a)[((size_t)((j) + nd - 1)) * ncol + (size_t)((i) + nd - 1)]
#define MOD(a, j, i) (a)[(size_t)((i)-1) * NY + (size_t)((j)-1)]
    const char *ext = p->snap_format == 2 ? ".asc" : ".bin";
    const char *mode = nsnap == 1 ? "w" : "a";
    char name[2][2 * DNIO_STR + 32];
    const float dhi = 1.0 / p->dh;
    int bad = 0;
    if (p->snap == 1 || p->snap == 4) {
        snprintf(name[0], sizeof name[0], "%s%s.x.0.0", p->snap_file, ext);
        snprintf(name[1], sizeof name[1], "%s%s.y.0.0", p->snap_file, ext);
        FILE *fx = fopen(name[0], mode), *fy = fopen(name[1], mode);
        if (!fx || !fy) {
            if (fx) fclose(fx);
            if (fy) fclose(fy);
            return fail("cannot open snapshot file %s", !fx ? name[0] : name[1]);
        }
        for (int i = 1; i <= NX; i += idx)
            for (int j = 1; j <= NY; j += idy) {
                bad |= snap_put(fx, NR(vx, j, i), p->snap_format);
                bad |= snap_put(fy, NR(vy, j, i), p->snap_format);
            }
        fclose(fx);
        fclose(fy);
    }
    if (p->snap == 2) {
        snprintf(name[0], sizeof name[0], "%s%s.p.0.0", p->snap_file, ext);
        FILE *fp = fopen(name[0], mode);
        if (!fp) return fail("cannot open snapshot file %s", name[0]);
        for (int i = 1; i <= NX; i += idx)
            for (int j = 1; j <= NY; j += idy) {
                const float amp = -NR(sxx, j, i) - NR(syy, j, i);
                bad |= snap_put(fp, amp, p->snap_format);
            }
        fclose(fp);
    }
    if (p->snap == 3 || p->snap == 4) {
        /* curl and divergence (Dougherty and Stephen, PAGEOPH 1988), snap_ssg.c:120-178 */
        snprintf(name[0], sizeof name[0], "%s%s.div.0.0", p->snap_file, ext);
        snprintf(name[1], sizeof name[1], "%s%s.rot.0.0", p->snap_file, ext);
        FILE *fd = fopen(name[0], mode), *fr = fopen(name[1], mode);
        if (!fd || !fr) {
            if (fd) fclose(fd);
            if (fr) fclose(fr);
            return fail("cannot open snapshot file %s", !fd ? name[0] : name[1]);
        }
        for (int i = 1; i <= NX; i += idx)
            for (int j = 1; j <= NY; j += idy) {
                float vyx = 0, vxy = 0;
                for (int m = 1; m <= fdoh; m++) {
                    vyx += hc[m] * (NR(vy, j, i + m) - NR(vy, j, i - m + 1));
                    vxy += hc[m] * (NR(vx, j + m, i) - NR(vx, j - m + 1, i));
                }
                vyx *= dhi;
                vxy *= dhi;
                const float curl = (vxy - vyx) * sqrt(MOD(pu, j, i));
                bad |= snap_put(fr, curl, p->snap_format);
            }
        for (int i = 1; i <= NX; i += idx)
            for (int j = 1; j <= NY; j += idy) {
                float vxx = 0, vyy = 0;
                for (int m = 1; m <= fdoh; m++) {
                    vxx += hc[m] * (NR(vx, j, i + m - 1) - NR(vx, j, i - m));
                    vyy += hc[m] * (NR(vy, j + m - 1, i) - NR(vy, j - m, i));
                }
                vxx *= dhi;
                vyy *= dhi;
                const float div = (vxx + vyy) * sqrt(MOD(ppi, j, i));
                bad |= snap_put(fd, div, p->snap_format);
            }
        fclose(fd);
        fclose(fr);
    }
#undef NR
#undef MOD
    return bad ? fail("write error in snapshot files %s*", p->snap_file) : 0;
}
