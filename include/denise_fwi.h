/*
 * denise_fwi.h -- host side of DENISE's FWI iteration over the stacked
 * gradients the device library returns: the limited-memory BFGS search
 * direction, the parabolic step-length search and the bounded model update.
 * Part of libdnio.so (plain C, no device code).
 *
 * Vectors are the caller's plain float arrays, nx*ny values each in the
 * reference's model-file order (x-major, y fastest: a[(i-1)*ny + (j-1)]),
 * which is the order of the reference's loops `for i, for j` over [j][i].
 * Every float operation follows the reference's order and precision,
 * including its single-precision sequential dot products, so a run gives
 * the reference's directions, step lengths and models bit for bit.
 *
 * Return codes: 0 ok, -1 with the message in dnio_last_error() otherwise.
 */
#ifndef DENISE_FWI_H
#define DENISE_FWI_H

#ifdef __cplusplus
extern "C" {
#endif

/* L-BFGS memory: NLBFGS (s, y) pairs over vectors of 3*nx*ny floats laid
 * out [rho | vs | vp], plus the previous iteration's normalised gradients
 * and model.  Replaces y_LBFGS, s_LBFGS, q_LBFGS, r_LBFGS, rho_LBFGS,
 * alpha_LBFGS and the JACOBIAN_p*.old files of the reference
 * (src/PSV/FWI_PSV.c:282-300 allocation, src/PSV/store_LBFGS_PSV.c:91-205,
 * src/PSV/extract_LBFGS_PSV.c:80-249). */
typedef struct dnfwi_lbfgs dnfwi_lbfgs;

int dnfwi_lbfgs_create(int nx, int ny, int nlbfgs, dnfwi_lbfgs **out);
/* The acoustic problem's memory (src/AC/FWI_AC.c:269-283, NLBFGS_class = 2):
 * vectors of 2*nx*ny floats [rho | vp] (src/AC/store_LBFGS_AC.c:63-150,
 * src/AC/extract_LBFGS_AC.c:28-57).  dnfwi_lbfgs_direction and
 * dnfwi_lbfgs_save then ignore the vs arguments (they may be NULL) and
 * dnfwi_lbfgs_save writes the vp and rho files only
 * (src/AC/extract_LBFGS_AC.c:60-178). */
int dnfwi_lbfgs_create_ac(int nx, int ny, int nlbfgs, dnfwi_lbfgs **out);
/* The SH problem's memory (src/SH/FWI_SH.c, NLBFGS_class = 2): vectors
 * [vs | rho] (src/SH/store_LBFGS_SH.c:36-134) whose model differences are of
 * the normalised models vs / C_vs and rho / C_rho (C_vs_min = C_rho_min = 0,
 * FWI_SH.c:572-573), and a direction left in normalised units
 * (src/SH/extract_LBFGS_SH.c:27-48 scales nothing back).
 * dnfwi_lbfgs_direction and dnfwi_lbfgs_save ignore the vp arguments (they
 * may be NULL); the files are the vs ones (_p_vs.old holding the normalised
 * model, _p_u.old, _c_u.old) and the rho ones (_p_mrho.old normalised too,
 * _p_rho.old, _c_rho.old). */
int dnfwi_lbfgs_create_sh(int nx, int ny, int nlbfgs, dnfwi_lbfgs **out);
void dnfwi_lbfgs_destroy(dnfwi_lbfgs *lb);
/* zero_LBFGS (src/zero_LBFGS.c) and LBFGS_pointer = 1 (src/PSV/FWI_PSV.c:807-811):
 * the memory is cleared at the start of each workflow stage after the first. */
void dnfwi_lbfgs_reset(dnfwi_lbfgs *lb);
/* the slot (1..NLBFGS) the last dnfwi_lbfgs_direction stored into */
int dnfwi_lbfgs_pointer(const dnfwi_lbfgs *lb);

/* One iteration's search direction (GRAD_METHOD = 2):
 *   the slot pointer step of src/PSV/FWI_PSV.c:452-466,
 *   store_LBFGS_PSV (src/PSV/store_LBFGS_PSV.c:37-205: normalise the
 *     gradients by C_vp, C_vs, C_rho; for iter > 1 store y = gradient
 *     change, s = model change, q = gradient),
 *   LBFGS (src/LBFGS.c:26-95: the two-loop recursion over the stored slots),
 *   extract_LBFGS_PSV (src/PSV/extract_LBFGS_PSV.c:31-249: for iter > 1 the
 *     direction H*q, scaled by C_vp, C_vs, C_rho; then keep the model and the
 *     normalised gradients for the next iteration).
 * g_vp, g_vs, g_rho: the stacked gradients on entry, the direction on return
 * (at iter = 1 the normalised gradients).  vp, vs, rho: the current model. */
int dnfwi_lbfgs_direction(dnfwi_lbfgs *lb, int iter, float c_vp, float c_vs, float c_rho, float *g_vp,
                          float *g_vs, float *g_rho, const float *vp, const float *vs, const float *rho);

/* The files the reference's extract_LBFGS_PSV writes after each direction
 * (src/PSV/extract_LBFGS_PSV.c:80-249, read back by store_LBFGS_PSV.c:91-160),
 * in its order, each as <jacobian>_<name>.old.0.0 (the one rank's part) and
 * <jacobian>_<name>.old (mergemod's whole-model file, identical on one rank),
 * float32 NX*NY in model-file order:
 *   _p_vp.old / _p_vs.old / _p_mrho.old   the model the direction was taken at,
 *   _p.old / _p_u.old / _p_rho.old        its normalised gradients C*g,
 *   _c.old / _c_u.old / _c_rho.old        the direction d_vp, d_vs, d_rho that
 *                                          dnfwi_lbfgs_direction returned.
 * Call after dnfwi_lbfgs_direction. */
int dnfwi_lbfgs_save(const dnfwi_lbfgs *lb, const char *jacobian, const float *d_vp, const float *d_vs,
                     const float *d_rho);

/* Preconditioned conjugate gradients (GRAD_METHOD = 1): PCG_old and PCG_dir
 * of 3*nx*ny floats [vp | vs | rho] (src/PSV/FWI_PSV.c:304-316), zero at
 * the start and after each stage (zero_PCG, FWI_PSV.c:802-805). */
typedef struct dnfwi_pcg dnfwi_pcg;
int dnfwi_pcg_create(int nx, int ny, dnfwi_pcg **out);
void dnfwi_pcg_destroy(dnfwi_pcg *pc);
void dnfwi_pcg_reset(dnfwi_pcg *pc);

/* One iteration's direction (FWI_PSV.c:637-662): descent (-g),
 * store_PCG_PSV, PCG (src/PCG.c:11-105: per class beta from single-precision
 * sums, PCG_BETA 1 Fletcher-Reeves, 2 Polak-Ribiere with beta >= 0,
 * 3 Hestenes-Stiefel, 4 Dai-Yuan; beta = 0 when its denominator is 0;
 * dir = -g + beta dir), extract_PCG_PSV, and descent again: the gradients
 * g_vp, g_vs, g_rho are replaced by -dir, the direction the model update
 * subtracts.  beta_out (may be NULL) gets the three betas.  g_vs may be
 * NULL (the acoustic problem, src/AC/FWI_AC.c:683-706 with store_PCG_AC and
 * extract_PCG_AC): the vs class then stays 0 and its beta is 0; the vp and
 * rho classes, each with its own beta, are those of the two-class run.
 * Likewise g_vp may be NULL (the SH problem, src/SH/FWI_SH.c:667-677 with
 * store_PCG_SH and extract_PCG_SH: classes vs and rho). */
int dnfwi_pcg_direction(dnfwi_pcg *pc, int pcg_beta, float *g_vp, float *g_vs, float *g_rho, float *beta_out);

/* Source-time-function inversion of one shot (INV_STF 1: src/PSV/stf_psv.c:13-66
 * calling src/stf.c:13-262): the Wiener deconvolution that turns the spike
 * wavelet into the source wavelet that best explains the observed vy traces.
 *   obs, syn: ntr x ns observed and synthetic vy traces (trace-major), the
 *     synthetics from a forward run with `spike` as source signal; both are
 *     modified in place: traces with mute[i] != 0 (TRKILL, OFFSET_MUTE,
 *     OFFSETC_STF; mute may be NULL) are zeroed, then NORMALIZE 1 divides
 *     each trace by its largest |sample| and NORMALIZE 2 scales the observed
 *     trace to the synthetic one's (samples 2..ns, src/util.c:489-579);
 *   spike: ns samples, the first source's signal of the synthetic run;
 *   eps_stf: EPS_STF, the water level relative to the mean energy;
 *   stf_out: ns samples of the wavelet, which the reference writes as
 *     SIGNAL_FILE_shot_<N>.dat and reads back with QUELLART 3.
 * Transforms of length dnfwi_stf_npad(ns) = 2^(ceil(log2 ns) + 2), in double
 * precision (the build's own FFT; equal to FFTW's up to double rounding). */
int dnfwi_stf(float *obs, float *syn, int ntr, int ns, const int *mute, int normalize, float eps_stf,
              const float *spike, float *stf_out);
int dnfwi_stf_npad(int ns);

/* Model averages C_vp, C_vs, C_rho of src/PSV/FWI_PSV.c:540-580: a
 * single-precision sequential sum in file order divided by NX*NY. */
float dnfwi_model_mean(const float *a, int nx, int ny);

/* Limits and switches of the model update (src/PSV/calc_mat_change_test_PSV.c). */
typedef struct dnfwi_update_par {
    float vpupperlim, vplowerlim, vsupperlim, vslowerlim, rhoupperlim, rholowerlim;
    float scalerho;                          /* SCALERHO of the workflow stage */
    int inv_vp_iter, inv_vs_iter, inv_rho_iter;
    int invmat1;                             /* 1: vp, vs, rho with limits; 3: no limits */
} dnfwi_update_par;

/* calc_mat_change_test_PSV (src/PSV/calc_mat_change_test_PSV.c:10-251):
 * new = old - eps * direction with eps = eps_scale * max(model) /
 * max|direction| per parameter (times SCALERHO for rho; 0 before the
 * parameter's INV_*_ITER), cells leaving the limits or turning negative
 * keep the old value.  Returns EPSILON_u, the reference's "eps_true".
 * With vs = NULL (d_vs and vs_new are then not used) it is
 * calc_mat_change_test_AC (src/AC/calc_mat_change_test_AC.c:9-189): vp and
 * rho only, and the value returned is EPSILON, that routine's eps_true.
 * With vp = NULL (d_vp and vp_new not used) it is calc_mat_change_test_SH
 * (src/SH/calc_mat_change_test_SH.c:10-189): vs and rho, the rho step
 * WITHOUT SCALERHO, and the value returned is EPSILON_u. */
float dnfwi_model_update(const dnfwi_update_par *u, int iter, float eps_scale, int nx, int ny, const float *vp,
                         const float *vs, const float *rho, const float *d_vp, const float *d_vs, const float *d_rho,
                         float *vp_new, float *vs_new, float *rho_new);

/* calc_opt_step (src/calc_opt_step.c:14-84) with solvelin's Gauss
 * elimination (src/solvelin.c:19-48): the vertex of the parabola through
 * (epst[k], L2t[k]), k = 0..2, kept within [epst[0], epst[2]]. */
float dnfwi_opt_step(const double *L2t, const float *epst);

/* Misfit of a test model: the sum of the L2 misfits of the test shots
 * (src/PSV/obj_psv.c:38-140).  itest is 2 or 3, the slot of L2t it fills. */
typedef double (*dnfwi_misfit_fn)(void *user, const float *vp, const float *vs, const float *rho, int itest);

typedef struct dnfwi_linesearch {
    float eps_scale;   /* the step length to apply (calc_opt_step's, or the last trial) */
    float epst[3];     /* epst1[1..3] */
    double L2t[3];     /* L2t[1..3] */
    int step1, step3;  /* step1 = 1: a step that lowers the misfit was found; step3 = 1: search failed */
    int nevals;        /* test forward modellings run */
} dnfwi_linesearch;

/* step_length_est_psv (src/PSV/step_length_est_psv.c:37-182): from
 * EPS_SCALE, divide the trial step by SCALEFAC until the misfit falls below
 * L2t0 (the current model's), then grow it by 1 + 1/SCALEFAC until the
 * misfit rises again (at most STEPMAX tries each), then fit the parabola.
 * Test models are the model update of (vp, vs, rho) along d at the trial step.
 * `out` also carries epst and L2t from one iteration to the next, as the
 * reference's arrays do (a failed search leaves slot 3 as it was): zero it
 * before the first iteration and pass the same record every iteration.
 * vs = NULL runs the acoustic search (src/AC/step_length_est_ac.c:37-181,
 * the same steps with calc_mat_change_test_AC): the misfit callback then
 * gets vs = NULL.  vp = NULL runs the SH search (src/SH/step_length_est_sh.c,
 * with calc_mat_change_test_SH): the callback gets vp = NULL. */
int dnfwi_step_length(const dnfwi_update_par *u, int iter, float eps_scale0, float scalefac, int stepmax,
                      double L2t0, int nx, int ny, const float *vp, const float *vs, const float *rho,
                      const float *d_vp, const float *d_vs, const float *d_rho, dnfwi_misfit_fn misfit,
                      void *user, dnfwi_linesearch *out);

/* Energy preconditioning of one shot's gradients (src/eprecond1.c:9-60 and
 * src/PSV/grad_obj_psv.c:219-238): the weights We = sqrt(Ws*Wr)
 * (EPRECOND 1) or sqrt(Ws) * (asinh((xmax-x)/y) - asinh((xmin-x)/y)) with
 * x = i*DH, y = j*DH, xmin = DH, xmax = NX*DH (EPRECOND 3, the forward field
 * and an approximate receiver Green's function, Plessix and Mulder 2004),
 * regularised by 0.005 * max(We); then g_vp /= We*C_vp*C_vp, g_vs /=
 * We*C_vs*C_vs (0 when C_vs = 0), g_rho /= We*C_rho*C_rho.  Ws, Wr, g_*:
 * x-major NX x NY (Wr only for EPRECOND 1); We (may be NULL) gets the
 * weights.  g_vs may be NULL (src/AC/grad_obj_ac.c:225-233: vp and rho
 * gradients only), and g_vp (src/SH/grad_obj_sh.c:211-219: the shear and
 * density sums of the SH run, divided by We C_vs^2 and We C_rho^2). */
int dnfwi_energy_precond(int eprecond, int nx, int ny, float dh, const float *Ws, const float *Wr, float c_vp,
                         float c_vs, float c_rho, float *g_vp, float *g_vs, float *g_rho, float *We);

/* Horizontal gradient taper (src/taper_grad.c:188-322, SWS_TAPER_GRAD_HOR):
 * 0 above depth index GRADT1, a Gaussian ramp exp(-(3 (k - L)/(L/2))^2 / 2),
 * k = 1..L+1, L = GRADT2 - GRADT1, from GRADT1 to GRADT2, and
 * (j*DH)^EXP_TAPER_GRAD_HOR below GRADT2; coeff (NX x NY, x-major) gets the
 * taper, and g (may be NULL) is multiplied by it. */
int dnfwi_taper_hor(int nx, int ny, float dh, int gradt1, int gradt2, float exp_taper, float *coeff, float *g);

/* Gaussian gradient filter (src/gauss_filt.c:9-208, SPATFILTER 3, applied by
 * FWI_PSV.c:628-631 through smooth_grad.c to the three stacked gradients
 * before the tapers): the width is round(WD_DAMP * (C_vs / FC_END) / DH)
 * points, an even width widened by one away from 0, half width h = |width|/2;
 * the kernel exp(-(di^2 / s) - (dj^2 / s)) with s = 2 sigma^2, sigma = h / 2
 * in integer division, divided by its sum (the reference starts that sum
 * uninitialised; here it starts at 0); the x-major nx x ny array g is
 * replaced by its convolution with the kernel, edges repeated.  Width 0
 * leaves g unchanged.  *width (may be NULL) gets the width used. */
int dnfwi_gauss_filter(int nx, int ny, float dh, float wd_damp, float c_vs, float fc_end, float *g, int *width);

/* Global-correlation residuals of one component of one shot (LNORM 5,
 * src/calc_res.c:130-157, 330-360, 465-489): per trace, with d the
 * observed and s the synthetic trace (GRAD_FORM 1: their running integrals
 * DT * sum; GRAD_FORM 2: the traces, the first observed sample replaced
 * by the synthetic one as calc_res.c:414 does, after the norms), the norms
 * |d|, |s| and the product d.s as float sums, the adjoint source
 * s (d.s) / (|s|^3 |d|) - d / (|s| |d|) time-reversed into diff (ntr x ns,
 * the order of dn_set_adjoint_sources), and, for a test shot, *l2 (in/out,
 * running like the reference's L2 over components) decreased by
 * d s / (|d| |s|) per sample.  Traces with skip[i] = 1 (may be NULL) get
 * zeros and no misfit. */
int dnfwi_residual_gc(const float *data, const float *syn, int ntr, int ns, float dt, int grad_form, int testshot,
                      const int *skip, float *diff, double *l2);

/* Time windows of one shot's residuals, in place on diff (ntr traces of ns
 * time-reversed samples: diff[i * ns + k] belongs to time (ns - k) DT, the
 * order of dn_get_adjoint_sources), as src/calc_res.c:447-450 (TIMEWIN 5)
 * and src/time_window.c:7-146 (TIMEWIN 1, 2, 4) apply them.
 *   TIMEWIN 1, 4: samples at t >= pick + TWLENGTH_PLUS are multiplied by
 *     exp(-GAMMA (t - pick - TWLENGTH_PLUS)^2); picks: ntr times in s.
 *   TIMEWIN 2: samples at t >= TWLENGTH_PLUS are multiplied by
 *     exp(-GAMMA (t + DT)^2) (the reference's constant window).
 *   TIMEWIN 5: surface waves damped, samples at t >= t0 = |offset /
 *     TWLENGTH_MINUS| multiplied by exp(-GAMMA (t - t0)^2); offset: ntr
 *     source-receiver distances in m.
 * Other values leave diff unchanged. */
int dnfwi_time_window(int timewin, int ntr, int ns, float dt, float gamma, float twlength_minus,
                      float twlength_plus, const float *picks, const float *offset, float *diff);

/* Velocity-dependent Gaussian gradient filter (src/gauss_filt_var.c:12-240,
 * SPATFILTER 4, with the current Vs model as vel, x-major like g): at each
 * cell, with lam = vel / FC_END, widths wx = round(WD_DAMP lam / DH) and
 * wz = round(WD_DAMP1 lam / DH); if wz > 0 both are made odd by adding one
 * to an even width and then raised to at least 6, h = w/2, sigma = h/2 in
 * integer division, s = 2 sigma^2, and the cell becomes
 * sum g k / sum k over the (2hx+1) x (2hz+1) window with
 * k = exp(-(di^2/sx) - (dj^2/sz)), edges repeated; cells with wz <= 0 are
 * unchanged.  The reference pads the edges by the half width of the maximum
 * velocity's WD_DAMP window only, so a wider WD_DAMP1 window reads outside
 * its array; here the edge cells repeat there too. */
int dnfwi_gauss_filter_var(int nx, int ny, float dh, float wd_damp, float wd_damp1, float fc_end, const float *vel,
                           float *g);

/* Median model filter (src/smooth_model.c:37-196 with src/median2d.c,
 * MODEL_FILTER 1, applied by FWI_PSV.c:752 to vp, vs and rho after each
 * model update): every cell of the x-major nx x ny array a is replaced by
 * the median of the |filt_size| x |filt_size| window around it (an even
 * size is widened by one), the array extended at its edges by repeating the
 * edge cells.  filt_size 0 leaves a unchanged. */
int dnfwi_median_filter(int nx, int ny, int filt_size, float *a);

/* Vertical gradient taper (src/taper_grad.c:46-186, SWS_TAPER_GRAD_VERT):
 * along x, 0 up to column GRADT1 and from GRADT4, Gaussian ramps
 * exp(-(3 (k - L1)/(L1/2))^2 / 2) up from GRADT1 (L1 = GRADT2 - GRADT1) and
 * down to GRADT4 (L2 = GRADT4 - GRADT3), 1 between; coeff (x-major) gets the
 * taper; g (may be NULL) is multiplied by it and then set to 0 at the grid
 * cells of all nsrc sources (srcpos: the 8 x nsrc table of
 * dnio_read_sources, x and y in metres, cell iround(x/DH), iround(y/DH)) and
 * of the ntr receivers (recpos: 2 x ntr, 1-based grid indices). */
int dnfwi_taper_vert(int nx, int ny, float dh, const int *gradt, int nsrc, const float *srcpos, int ntr,
                     const int *recpos, float *coeff, float *g);

/* Cylindrical gradient taper around the sources (src/taper_grad.c:328-577,
 * SWS_TAPER_GRAD_SOURCES): a stencil of ceil(2 SRTRADIUS/DH) (at least 5)
 * points with erf(r/rmax) (SRTSHAPE 1) or log(r) (2), shifted to 0, scaled
 * by its edge maximum and clipped at 1, placed at every source's cell; the
 * minimum over the sources, rescaled to [0, 1], is the taper (coeff,
 * x-major), then 0 in the (2 FILTSIZE + 1)^2 box around each source.  g (may
 * be NULL) is multiplied by the taper before the boxes are set. */
int dnfwi_taper_sources(int nx, int ny, float dh, float srtradius, int srtshape, int filtsize, int nsrc,
                        const float *srcpos, float *coeff, float *g);

/* ass_gradSH (src/SH/ass_gradSH.c:10-98) on the shot-stacked SH sums (FWI_SH.c
 * assembles after the stacking): with mu' = -DT g_u, g_u becomes the vs
 * gradient (2 / rho * vs^3) mu' (INVMAT1 1, GRAD_FORM 1, where vs > 0, as the
 * reference writes it), 2 rho vs mu' (INVMAT1 1, GRAD_FORM 2) or mu'
 * (INVMAT1 3); g_rho becomes 1/(vs^2 rho^2) mu' - DT g_rho (INVMAT1 1,
 * GRAD_FORM 1, vs > 0), vs^2 mu' - DT g_rho (GRAD_FORM 2) or -DT g_rho
 * (INVMAT1 3); 0 before INV_VS_ITER / INV_RHO_ITER.  vs: the model's vs (or
 * mu for INVMAT1 3), rho; x-major NX x NY. */
int dnfwi_assemble_sh(int nx, int ny, float dt, int invmat1, int grad_form, int iter, int inv_vs_iter,
                      int inv_rho_iter, const float *vs, const float *rho, float *g_u, float *g_rho);

/* Adjoint source of reverse-time migration (MODE 2), the MODE 2 branch of
 * calc_res with LNORM = 2 (src/calc_res.c:416-446): per trace the first data
 * sample is replaced by the synthetic one, then GRAD_FORM 1 integrates
 * DT * data in time (single precision), GRAD_FORM 2 takes the data, and the
 * result is stored time-reversed.  data, syn, diff: ntr x ns (trace-major).
 * l2 (may be NULL) gets sum 0.5*diff^2 in the reference's order
 * (src/calc_res.c:563-569), or 0 when testshot is 0. */
int dnfwi_rtm_adjoint_source(const float *data, const float *syn, int ntr, int ns, float dt, int grad_form,
                             int testshot, float *diff, double *l2);

#ifdef __cplusplus
}
#endif
#endif
