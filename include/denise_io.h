/*
 * denise_io.h -- host-side file front end of the P/SV drop-in (libdnio.so):
 * the reference's parameter, workflow, model, source and receiver readers,
 * the source wavelets, and the SU seismogram / binary model writers, so that
 * the executable denise_hip reads and writes the same files as DENISE.
 *
 * Plain C11, no GPU code.  Every function returns 0 on success or -1 with a
 * message in dnio_last_error() (thread-local); nothing exits the process.
 * Arrays are caller-owned unless a function documents a malloc'd result,
 * which the caller releases with dnio_free().
 */
#ifndef DENISE_IO_H
#define DENISE_IO_H

#ifdef __cplusplus
extern "C" {
#endif

#define DNIO_STR 256
#define DNIO_MAXL 8

/* The parameters of the .inp file, in the order src/read_par.c:96-447 reads
 * them (one value per non-comment line).  nt is derived (FD_PSV.c:68). */
typedef struct dnio_par {
    int mode, physics, nprocx, nprocy, fdorder, maxrelerror, nx, ny;
    float dh, time, dt;
    int quellart;
    char signal_file[DNIO_STR];
    float ts;
    int srcrec;
    char source_file[DNIO_STR];
    int run_multiple_shots;
    float fc_spike_1, fc_spike_2;
    int order_spike, write_stf, readmod;
    char mfile[DNIO_STR];
    int writemod, L;
    float fl[DNIO_MAXL];
    float tau;
    int free_surf, fw;
    float damping, fpml, npower, k_max_pml;
    int boundary, snap, snap_shot;
    float tsnap1, tsnap2, tsnapinc;
    int idx, idy, snap_format;
    char snap_file[DNIO_STR];
    int seismo, readrec;
    char rec_file[DNIO_STR];
    float refrec[4];          /* [1], [2] from the file; [3] stays 0 */
    int n_streamer;
    float rec_incr_x, rec_incr_y;
    int ndt, seis_format;
    char seis_file_vx[DNIO_STR], seis_file_vy[DNIO_STR], seis_file_curl[DNIO_STR];
    char seis_file_div[DNIO_STR], seis_file_p[DNIO_STR], log_file[DNIO_STR];
    int log, itermax;
    char jacobian[DNIO_STR], data_dir[DNIO_STR];
    int taper, taperlength, gradt[4], invmat1, grad_form, quelltypb;
    int testshot_start, testshot_end, testshot_incr;
    int sws_taper_grad_vert, sws_taper_grad_hor;
    float exp_taper_grad_hor;
    int sws_taper_grad_sources, sws_taper_circular_per_shot, srtshape;
    float srtradius;
    int filtsize, sws_taper_file;
    char tfile[DNIO_STR];
    int inv_mod_out;
    char inv_modelfile[DNIO_STR];
    float vpupperlim, vplowerlim, vsupperlim, vslowerlim, rhoupperlim, rholowerlim;
    float qsupperlim, qslowerlim;
    int grad_method, pcg_beta, nlbfgs, model_filter, filt_size, dtinv;
    float eps_scale;
    int stepmax;
    float scalefac;
    int trkill;
    char trkill_file[DNIO_STR], picks_file[DNIO_STR], misfit_log_file[DNIO_STR];
    int min_iter, grad_filter, filt_size_grad, timelapse;
    char data_dir_t0[DNIO_STR];
    int rtmod, gravity, ngravb, nzgrav, grav_type, back_density;
    char dfile[DNIO_STR];
    int rtm_shot;
    /* derived */
    int nt;                   /* iround(TIME/DT) */
    int lines;                /* parameter lines found (115 in a full file) */
} dnio_par;

/* One stage of the FWI workflow file (src/read_par_inv.c:32-35). */
typedef struct dnio_stage {
    float pro;
    int time_filt;
    float fc_start, fc_end;
    int order, timewin;
    float gamma, twlength_minus, twlength_plus;
    int inv_vp_iter, inv_vs_iter, inv_rho_iter, inv_qs_iter, spatfilter;
    float wd_damp, wd_damp1;
    int eprecond, lnorm, rowi, inv_stf;
    float offsetc_stf, eps_stf;
    int normalize, offset_mute;
    float offsetc, scalerho, scaleqs;
    int env;
    float gamma_grav;
    int n_order;
} dnio_stage;

const char *dnio_last_error(void);
void dnio_free(void *p);

/* read_par (src/read_par.c:14-456): the same line-counting rule (a line is
 * a parameter line when it does not start with '#'; the first line of the
 * file is never one) and the same per-line conversions.  Sets nt. */
int dnio_read_par(const char *path, dnio_par *p);

/* read_par_inv (src/read_par_inv.c:32-35): skips the 30-word header, reads
 * stages 1..stage and keeps the last. */
int dnio_read_stage(const char *path, int stage, dnio_stage *s);

/* Number of stages (non-empty lines after the header) in a workflow file. */
int dnio_count_stages(const char *path, int *nstages);

/* sources (src/sources.c:9-127): srcpos is a malloc'd 8 x nsrc table
 * (row-major, rows = srcpos[1..8] of the reference: x, y, 0, tshift, fc,
 * amp, azimuth, type; coordinates in metres).  ts gets 1/max(fc). */
int dnio_read_sources(const char *path, int *nsrc, float **srcpos, float *ts);

/* splitsrc (src/splitsrc.c:9-60) on one rank: sources first..first+count-1
 * (0-based columns of srcpos) that fall on the NX x NY grid, converted to
 * grid indices; loc is 8 x count (row-major, count = columns), *nloc the
 * number kept (they are packed in the first *nloc columns). */
int dnio_split_sources(const dnio_par *p, const float *srcpos, int nsrc, int first, int count,
                       float *loc, int *nloc);

/* receiver (src/receiver.c:9-121): REC_FILE.dat (READREC 1) or
 * REC_FILE_shot_<ishot>.dat (READREC 2); positions to grid indices with
 * REFREC, duplicates and receivers at x index <= 0 removed.  recpos is a
 * malloc'd 3 x ntr table (x, y, z indices). */
int dnio_read_receivers(const dnio_par *p, int ishot, int *ntr, int **recpos);

/* wavelet (src/wavelet.c:10-121): signals (nsrc x nt, row-major) for
 * QUELLART 1 Ricker, 2 Fuchs-Mueller, 3 SIGNAL_FILE_shot_<ishot>.dat,
 * 4 sin^3, 5 Gaussian derivative, 6 spike, 7 Klauder; then N_ORDER time
 * integrations.  srcpos_loc as from dnio_split_sources; ts is TS. */
int dnio_wavelet(const dnio_par *p, const float *srcpos_loc, int nsrc, int ishot, float ts, int n_order,
                 float *signals);

/* readmod_elastic_PSV / readmod_visc_PSV (src/PSV/readmod_*_PSV.c): model
 * files MFILE.vp/.vs/.rho (INVMAT1 1) or .lam/.mu/.rho (3), plus .qp/.qs for
 * L > 0 (values replaced by Q = 30 as the reference does; taus = taup =
 * 2/Q).  Outputs NX*NY x-major; taus/taup may be NULL when L = 0.  Writes
 * the copies the reference writes (WRITEMOD: MFILE.denise.{pi,mu,rho};
 * always for L > 0: MFILE.fdveps.{pi,mu,rho}). */
int dnio_read_model(const dnio_par *p, float *vp, float *vs, float *rho, float *taus, float *taup);
/* VTI / TTI model files (PHYSICS 3 / 4): MFILE.c11, .c13, .c33, .c44, .rho and,
 * for TTI, .theta in degrees (src/VTI/readmod_elastic_VTI.c:27-118,
 * src/TTI/readmod_elastic_TTI.c:27-151, with their model copies); NX*NY
 * x-major each, theta NULL for VTI. */
int dnio_read_model_aniso(const dnio_par *p, float *rho, float *c11, float *c13, float *c33, float *c44,
                          float *theta);
/* acoustic model files (PHYSICS 2): MFILE.vp (INVMAT1 1) or MFILE.lam (3)
 * and MFILE.rho, NX*NY x-major; writes MFILE.fdveps.pi and .fdveps.rho as
 * src/AC/readmod_AC.c:27-93 does. */
int dnio_read_model_ac(const dnio_par *p, float *vp, float *rho);
/* SH model files (PHYSICS 5): MFILE.vs (INVMAT1 1) or MFILE.mu (3) and
 * MFILE.rho, NX*NY x-major; with WRITEMOD writes MFILE.denise.mu and
 * .denise.rho as src/SH/readmod_elastic_SH.c:21-98 does. */
int dnio_read_model_sh(const dnio_par *p, float *vs, float *rho);
/* readmod_visc_SH (src/SH/readmod_visc_SH.c:9-123), PHYSICS 5 with L > 0:
 * as dnio_read_model_sh plus MFILE.qs, returned as taus = 2 / Qs. */
int dnio_read_model_sh_visc(const dnio_par *p, float *vs, float *rho, float *taus);

/* writemod + mergemod on one rank (src/writemod.c, src/mergemod.c): a
 * float32 x-major NX x NY file `name` and its per-rank copy `name.0.0`. */
int dnio_write_model(const char *name, const float *a_xmajor, int nx, int ny);

/* outseis_glob (src/outseis_glob.c:8-303), sws = 1: ntr traces of section
 * (ntr x nt, every NDT-th sample written), SEIS_FORMAT 1 SU (240-byte
 * native-endian header + float32 samples), 2 ASCII, 3 binary.  recpos is
 * 3 x ntr grid indices; xs, ys the shot's source coordinates in metres. */
int dnio_write_seis(const char *path, const dnio_par *p, const float *section, int ntr, int nt,
                    const int *recpos, float xs, float ys);

/* inseis (src/inseis.c:8-110): ntr SU traces of ns samples into section
 * (ntr x ns): sample k (1-based) of the file goes to index k, the first
 * output sample is 0 (the reference's one-sample shift). */
int dnio_read_su(const char *path, int ntr, int ns, float *section);

/* File names the reference uses: saveseis_glob (src/saveseis_glob.c:35-50)
 * comp 0 vx, 1 vy, 2 p, 3 div, 4 curl; observed data of inseis sws 1/2/11
 * (src/inseis.c:21-61) comp 0 vx, 1 vy, 2 p. */
/* One wavefield snapshot (src/snap_ssg.c:10-184, src/writedsk.c): SNAP 1
 * vx, vy; 2 -sxx-syy; 3 div*sqrt(ppi) and curl*sqrt(pu) of the velocity by
 * the FD stencil hc; 4 both 1 and 3; every IDX-th column and IDY-th row;
 * SNAP_FORMAT 2 ASCII "%e", 3 raw float32 (1 is refused, as by writedsk).
 * Files SNAP_FILE<.asc|.bin>.{x,y,p,div,rot}.0.0 (one rank's POS), created
 * at nsnap = 1 and appended to after.  vx, vy, sxx, syy: the padded NR layout
 * of dn_get_field; ppi, pu: the model as read (x-major NX x NY; vp and vs
 * for INVMAT1 = 1); hc: hc[0..6] (dn_get_hc). */
int dnio_write_snap(const dnio_par *p, int nsnap, const float *vx, const float *vy, const float *sxx,
                    const float *syy, const float *ppi, const float *pu, const float *hc);
/* Butterworth filtering of ntr traces of ns samples in place
 * (src/apply_tdfilt.c, src/timedomain_filt.c, the vendored libcseife's
 * seife_lpb / seife_hpb): low-pass with corner fc2 and, when fc1 > 0,
 * high-pass with corner fc1, both of the given order; each pass filters in
 * double and stores float.  Used for the spike source (QUELLART 6: ORDER_SPIKE,
 * FC_SPIKE_2, FC_SPIKE_1) and for TIME_FILT stages (ORDER, FC_END, FC_START). */
int dnio_tdfilt(float *data, int ntr, int ns, float dt, int order, float fc2, float fc1);
int dnio_seis_name(const dnio_par *p, int comp, int ishot, int iter, char *out, int n);
int dnio_data_name(const dnio_par *p, int comp, int ishot, char *out, int n);

#ifdef __cplusplus
}
#endif
#endif /* DENISE_IO_H */
