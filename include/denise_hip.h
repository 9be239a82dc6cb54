/*
 * denise_hip.h -- C ABI of the MI355X (gfx950) P/SV time-stepping library
 * (libdnhip.so).
 *
 * This is the drop-in seam for DENISE Black Edition's per-shot P/SV path.
 * The reference crosses it with
 *     psv(struct wavePSV*, struct wavePSV_PML*, struct matPSV*, struct fwiPSV*,
 *         struct mpiPSV*, struct seisPSV*, struct seisPSVfwi*, struct acq*,
 *         float *hc, int ishot, int nshots, int nsrc_loc, int ns, int ntr,
 *         float **Ws, float **Wr, int hin, int *DTINV_help, int mode, ...)
 *     (src/PSV/psv.c:17-19, prototype include/fd.h:274-276)
 * called from src/PSV/FD_PSV.c:351 (forward), src/PSV/grad_obj_psv.c:173
 * (forward with snapshot store) and :211 (adjoint with correlation), plus the
 * host set-up it relies on (PML_pro, holbergcoeff, av_mue/av_rho/av_tau,
 * prepare_update_s_visc_PSV, calc_res, ass_gradPSV).
 *
 * Conventions (all entry points):
 *   - plain pointers and sizes; caller owns host buffers; the context owns
 *     device memory;
 *   - models and gradients are NX*NY float32 in the reference's file order
 *     (x-major, y fastest: src/PSV/readmod_elastic_PSV.c:72-76,
 *     src/PSV/extract_LBFGS_PSV.c:103-110);
 *   - seismograms / adjoint sources are trace-major ntr x NT float32
 *     (the reference's section[itr][nt], src/seismo_ssg.c:32-35);
 *   - grid indices are the reference's 1-based (i = 1..NX, j = 1..NY);
 *   - every call returns DN_OK (0) or a negative dn_status; nothing exits the
 *     process; dn_last_error() gives the message (thread-local).
 * Contexts are independent; one context must not be used from two threads at
 * once.  Each context owns one HIP stream.
 */
#ifndef DENISE_HIP_H
#define DENISE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DN_ABI_VERSION 3

typedef enum dn_status {
    DN_OK = 0,
    DN_ERR_ARG = -1,        /* invalid argument / configuration            */
    DN_ERR_HIP = -2,        /* HIP runtime error                           */
    DN_ERR_OOM = -3,        /* device allocation failed                    */
    DN_ERR_UNSTABLE = -4,   /* NaN probe tripped (src/PSV/psv.c:99-103)     */
    DN_ERR_STATE = -5,      /* call out of order (e.g. adjoint before store)*/
    DN_ERR_CFL = -6         /* DT above the stability limit
                               (src/checkfd_ssg_elastic.c:127-128)          */
} dn_status;

/* Run parameters: the reference globals the P/SV path reads
 * (include/globvar.h; values as parsed by src/read_par.c:92-449). */
typedef struct dn_params {
    int nx, ny, nt;          /* NX, NY, NT = iround(TIME/DT)                  */
    int fdorder;             /* FDORDER 2..12                                  */
    int maxrelerror;         /* MAXRELERROR 0..4                               */
    int fw;                  /* FW (CPML width in cells)                       */
    int free_surf;           /* FREE_SURF                                      */
    int boundary;            /* BOUNDARY (must be 0)                           */
    int L;                   /* relaxation mechanisms, 0 = elastic (<= 8)      */
    int seismo;              /* SEISMO 0..4                                    */
    int ndt;                 /* NDT                                            */
    int invmat1;             /* INVMAT1 1 (vp,vs,rho) or 3 (lambda,mu,rho)     */
    int grad_form;           /* GRAD_FORM 1 or 2                               */
    int dtinv;               /* DTINV                                          */
    int quelltypb;           /* QUELLTYPB 1..7                                 */
    int mode;                /* MODE: 0 forward only, 1 FWI (store snapshots)  */
    float dh, dt;            /* DH, DT                                         */
    float damping, fpml, npower, k_max_pml;   /* CPML (src/PML_pro.c)          */
    float ts;                /* TS = 1/max(fc) (src/sources.c:92)              */
    float fl[8];             /* FL[1..L]                                       */
    int physics;             /* PHYSICS: 1 (or 0) isotropic P/SV, 2 acoustic,
                                3 VTI, 4 TTI, 5 SH (src/denise.c:116-141);
                                2-4 with L = 0, 5 with L = 0..3 (L > 0:
                                forward runs, FD_SH.c:171-206); 5 records
                                vz as component 1 (SEISMO 0 or 1);
                                2, 3, 4 and elastic 5 also adjoint runs
                                (mode 1: the RTM passes of RTM_AC.c,
                                RTM_VTI.c, RTM_TTI.c, TTI with GRAD_FORM 1;
                                the gradients of FWI_AC.c and FWI_SH.c; SH
                                runs the vz residual, calc_res_SH.c)        */
} dn_params;

typedef struct dn_ctx dn_ctx;

/* ids for dn_get_field: padded NR layout (rows -nd+1..NY+nd, columns
 * -nd+1..NX+nd, nd = FDORDER/2+1, x fastest), i.e. the reference's
 * float** arrays (src/util.c:177-203). */
enum {
    DN_F_VX = 0, DN_F_VY, DN_F_SXX, DN_F_SYY, DN_F_SXY, DN_F_VXP1, DN_F_VYP1,
    DN_F_UX, DN_F_UY, DN_F_UXY
};

/* Library / device information. */
int dn_abi_version(void);
const char *dn_last_error(void);
int dn_device_count(int *count);

/* Context.  Replaces FD_PSV.c:170-249: allocation (alloc_PSV), CPML profiles
 * (src/PML_pro.c:65-407), FD coefficients (src/holbergcoeff.c:21-90) and the
 * stability check (src/checkfd_ssg_elastic.c:9-145 / checkfd_ssg_visc.c).
 * device: HIP device ordinal. */
int dn_ctx_create(const dn_params *p, int device, dn_ctx **out);
int dn_ctx_destroy(dn_ctx *ctx);

/* Model upload + material preparation.  Replaces readmod_*_PSV (in-memory
 * arrays instead of files), matcopy_elastic_PSV / matcopy_PSV (one-rank
 * periodic ghosts), av_mue, av_rho, av_tau, prepare_update_s_visc_PSV
 * (src/PSV/FD_PSV.c:203-249).  vp/vs/rho (INVMAT1=3: lambda/mu/rho) and, for
 * L>0, taus/taup: NX*NY x-major.  Runs checkfd_ssg_*: DN_ERR_CFL if DT is
 * above the stability limit.  Acoustic contexts (params.physics 2) take vp
 * (INVMAT1=3: the modulus) and rho only, vs/taus/taup ignored (may be
 * NULL): readmod_AC, matcopy_acoustic_AC, av_rho and checkfd_acoustic
 * (src/AC/FD_AC.c:184-207).  SH contexts (params.physics 5) take vs
 * (INVMAT1=3: mu) and rho, and for L > 0 taus (2/Qs), vp/taup ignored:
 * readmod_elastic_SH / readmod_visc_SH, matcopy_SH, av_mu_SH, inv_rho_SH,
 * av_tau, prepare_update_s_visc_SH and checkfd_elast_SH / checkfd_visc_SH
 * (src/SH/FD_SH.c:171-206). */
int dn_set_model(dn_ctx *ctx, const float *vp, const float *vs, const float *rho,
                 const float *taus, const float *taup);

/* Anisotropic model upload (params.physics 3 or 4).  Replaces
 * readmod_elastic_VTI (src/VTI/readmod_elastic_VTI.c:60-79: MFILE.c11, .c13,
 * .c33, .c44, .rho) with matcopy_elastic_VTI, av_harm (c44 at the shear
 * points) and av_rho (src/VTI/FD_VTI.c:180-197), and for TTI
 * readmod_elastic_TTI (src/TTI/readmod_elastic_TTI.c:64-85, plus MFILE.theta
 * in degrees), rot_el_tensor_TTI and av_harm of the rotated d15, d35, d55
 * (src/TTI/FD_TTI.c:180-205).  All NX*NY x-major; theta is NULL for VTI.
 * VTI runs checkfd_ssg_VTI (src/VTI/checkfd_ssg_VTI.c:33-125: DN_ERR_CFL);
 * the reference's TTI driver runs no stability check (FD_TTI.c:187). */
int dn_set_model_aniso(dn_ctx *ctx, const float *rho, const float *c11, const float *c13,
                       const float *c33, const float *c44, const float *theta);

/* Sources of the shot.  srcpos: 8 x nsrc rows as srcpos_loc
 * (src/splitsrc.c:28-36): [0] x grid index, [1] y grid index, [2] source
 * number, [3] tshift, [4] fc, [5] amp, [6] azimuth (deg), [7] type
 * (QUELLTYP 1 explosive, 2/3 force x/y, 4 rotated force, 5 moment tensor).
 * signals: nsrc x NT (src/wavelet.c:10-121 output). */
int dn_set_sources(dn_ctx *ctx, int nsrc, const float *srcpos, const float *signals);

/* Receivers: recpos 2 x ntr, [0..ntr) x grid indices, [ntr..2ntr) y grid
 * indices (src/splitrec.c:29-31).  Positions must be distinct (the reference
 * removes duplicates in src/receiver.c:73-86). */
int dn_set_receivers(dn_ctx *ctx, int ntr, const int *recpos);

/* psv() mode 0: nsteps (<= NT, 0 = NT) forward time steps on the device.
 * With params.mode > 0 the forward fields are stored every DTINV steps
 * (src/PSV/psv.c:240-309). Blocks until done. */
int dn_run_forward(dn_ctx *ctx, int nsteps);

/* Seismograms of the last forward run, comp 0 vx, 1 vy, 2 p, 3 div, 4 curl,
 * ntr x NT (src/seismo_ssg.c:25-106). */
int dn_get_seismogram(dn_ctx *ctx, int comp, float *out);

/* L2 data residuals as adjoint sources (LNORM=2, GRAD_FORM=1): time-integrated
 * synthetic minus observed, time reversed (src/calc_res.c:296-452,
 * src/PSV/calc_res_PSV.c:8-184, src/inseis.c:96-103).  obs_*: ntr x NT SU
 * samples (NULL for a component QUELLTYPB does not use).  l2 (may be NULL)
 * gets the misfit 0.5*sum(residual^2) when testshot != 0, else 0. */
int dn_calc_residuals(dn_ctx *ctx, const float *obs_vx, const float *obs_vy,
                      const float *obs_p, int testshot, double *l2);

/* dn_calc_residuals that allocates nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and adjoint sources as dn_calc_residuals, but the
 * device scratch of the call -- the observed traces of one component staged
 * on the device and the misfit with its partial sums -- lives in one fixed
 * static device array of libdnhip.so (about 32 MiB per GPU, held until the
 * process ends) instead of the context's buffers; no hipMalloc, hipFree or
 * allocating library call.  ntr * NT must not exceed DN_NOALLOC_MAX_SAMPLES
 * (else DN_ERR_ARG).  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (from several threads) race. */
#define DN_NOALLOC_MAX_SAMPLES 8388608
int dn_calc_residuals_noalloc(dn_ctx *ctx, const float *obs_vx, const float *obs_vy,
                              const float *obs_p, int testshot, double *l2);

/* Component weights of the residuals (COMP_WEIGHT, src/calc_res.c:100-123,
 * set by src/PSV/calc_res_PSV.c:65-67 and 118-120 for QUELLTYPB 5, 6, 7):
 * dn_calc_residuals multiplies the synthetic and the observed samples of
 * vx by w_vx and of vy by w_vy (float products) before it forms their
 * residuals; the reference's weight is C_vp*C_rho of the weighted
 * component, 1 otherwise.  Default 1, 1 (no change). */
int dn_set_residual_weights(dn_ctx *ctx, float w_vx, float w_vy);

/* Adjoint sources supplied directly (ntr x NT each, NULL = zero):
 * sectionvxdiff / sectionvydiff / sectionpdiff of the reference. */
int dn_set_adjoint_sources(dn_ctx *ctx, const float *dvx, const float *dvy, const float *dp);
int dn_get_adjoint_sources(dn_ctx *ctx, int comp, float *out);

/* psv() mode 1: adjoint propagation from the receivers with the
 * cross-correlation against the stored forward fields
 * (src/PSV/psv.c:313-387). Requires a preceding dn_run_forward with
 * params.mode > 0 over the full NT. */
int dn_run_adjoint(dn_ctx *ctx, int nsteps);

/* ass_gradPSV (src/PSV/ass_gradPSV.c:10-155), for acoustic contexts
 * ass_gradAC (src/AC/ass_gradAC.c:12-93: vp or lambda and rho, no shear
 * term), for SH contexts ass_gradSH (src/SH/ass_gradSH.c:10-98: vs or mu
 * and rho, vp gradient 0) of this shot, iteration 1 with INV_*_ITER=0
 * (the SH inversion assembles the shot stack instead: dnfwi_assemble_sh in
 * denise_fwi.h).  VTI/TTI contexts refuse it
 * (DN_ERR_STATE): their images are the raw correlation sums
 * (grad_obj_VTI.c:196-197 does not assemble). */
int dn_assemble_gradient(dn_ctx *ctx);

/* which: 0..2 raw correlation sums (waveconv_shot, waveconv_u_shot,
 * waveconv_rho_shot; VTI/TTI: the P image, 0, the density image,
 * src/VTI/VTI.c:297-368; SH: 0, the shear sum, the density sum,
 * src/SH/sh.c:186-223), 3..5 per-shot vp / vs / rho gradients.  NX*NY
 * x-major. */
int dn_get_gradient(dn_ctx *ctx, int which, float *out);

/* Diagnostics used by the tests. */
int dn_get_dims(dn_ctx *ctx, int *nd, int *ntdtinv);
int dn_get_field(dn_ctx *ctx, int id, float *out);
/* CPML memory: 0 psi_sxx_x 1 psi_sxy_x 2 psi_vxx 3 psi_vyx 4 psi_vxxs
 * (NY x 2FW), 5 psi_syy_y 6 psi_sxy_y 7 psi_vyy 8 psi_vxy (2FW x NX). */
int dn_get_psi(dn_ctx *ctx, int id, float *out);
/* memory variables which 0 r, 1 p, 2 q; l = 1..L; padded NR layout */
int dn_get_memvar(dn_ctx *ctx, int which, int l, float *out);
/* CPML profiles (20 vectors of 2FW, PML_pro argument order) and hc[0..6] */
int dn_get_pml(dn_ctx *ctx, int which, float *out);
int dn_get_hc(dn_ctx *ctx, float *out7);

/* Tile shape of the fused sweeps: blocks of tbx x by threads, each thread
 * marching ty rows of vw adjacent columns (vw 1, or 2 / 4 with vector
 * loads).  sweep: 0 velocity, 1 stress, -1 both (set only).  Compiled
 * shapes (tbx,by,ty,vw): (64,4,1,1), (128,2,1,1), (256,1,1,1), (64,4,2,1),
 * (128,2,4,1), (64,8,2,1), (64,4,1,4), (64,4,2,4), (64,2,4,4), (64,4,1,2),
 * (128,2,2,2), (64,4,2,2); VTI/TTI stress sweeps take vw = 1 only.
 * Results do not depend on the shapes.  No reference
 * counterpart: the reference has no GPU launch shapes. */
int dn_set_tile(dn_ctx *ctx, int sweep, int tbx, int by, int ty, int vw);
int dn_get_tile(dn_ctx *ctx, int sweep, int *tbx, int *by, int *ty, int *vw);

/* Cache policy of a sweep: 0 plain; 1 streaming (nontemporal) loads of the
 * planes a step reads once (material and coefficient planes) and streaming
 * stores of the memory variables, so that they do not displace from the
 * caches the wavefields the next sweep reads.  sweep 0 velocity, 1 stress,
 * -1 both (set only).  Results do not depend on it.  No reference
 * counterpart. */
int dn_set_store_hint(dn_ctx *ctx, int sweep, int streaming);
int dn_get_store_hint(dn_ctx *ctx, int sweep, int *streaming);

/* Tile autotuner: times nsteps forward steps of this context's own
 * configuration with every compiled shape (HIP events around each launch)
 * and keeps the fastest shape of each sweep independently, then the pair
 * of cache policies (dn_set_store_hint) with the fastest step; vel_shape /
 * str_shape (4 ints:
 * tbx, by, ty, vw; may be NULL) receive the shapes.  Leaves
 * the wavefields zeroed (a following dn_run_forward starts from rest).  No
 * reference counterpart. */
int dn_autotune_tile(dn_ctx *ctx, int nsteps, int *vel_shape, int *str_shape);

/* Timing of the last run: total device milliseconds (hipEvents around the
 * time loop on the context stream) and, when kernel sampling is on, the
 * average duration of the velocity and stress kernels. every = 0 turns
 * sampling off; every = k times one launch of each in k steps. */
int dn_set_kernel_sampling(dn_ctx *ctx, int every);

/* Energy preconditioning (EPRECOND, src/eprecond.c, psv.c:300-303, 380-383):
 * with eprecond 1 or 3 the runs sum vx^2 + vy^2 over the grid, from zero at
 * each forward run: which 0 (Ws) the forward field at its snapshot-store
 * steps, which 1 (Wr, eprecond 1 only) the adjoint field at its
 * correlation steps.  dn_get_energy returns NX x NY, x-major.  The weights
 * (src/eprecond1.c) and the gradient scaling are the host's
 * (dnfwi_energy_precond).  0 switches it off.  The engine sums them too. */
int dn_set_energy_weights(dn_ctx *ctx, int eprecond);
int dn_get_energy(dn_ctx *ctx, int which, float *out);

/* Wavefield snapshots during dn_run_forward (src/PSV/psv.c:75-77, 225-231):
 * after time step `first` and every `inc` steps after it, up to `last`,
 * the run waits for the device and calls fn with the step, the snapshot
 * count of this run (from 1) and vx, vy, sxx, syy in the dn_get_field
 * layout (valid during the call).  inc <= 0 gives one snapshot; fn = NULL
 * switches snapshots off.  The multi-shot engine takes none. */
typedef void (*dn_snap_fn)(void *user, int nt, int nsnap, const float *vx, const float *vy, const float *sxx,
                           const float *syy);
int dn_set_snapshots(dn_ctx *ctx, int first, int inc, int last, dn_snap_fn fn, void *user);
int dn_get_timing(dn_ctx *ctx, double *run_ms, double *v_kernel_ms, double *s_kernel_ms,
                  int *samples);

/* Device bytes held by the context. */
int dn_get_footprint(dn_ctx *ctx, size_t *bytes);

/* Free and total device memory of a HIP device (hipMemGetInfo). */
int dn_device_memory(int device, size_t *free_bytes, size_t *total_bytes);

/* Device bytes a context with these parameters, nsrc sources and ntr
 * receivers will hold (wavefields, material, CPML, seismograms and, for
 * params.mode > 0, the forward snapshots), computed without a device: the
 * admission test of a scheduler that packs shots into HBM. */
int dn_estimate_footprint(const dn_params *p, int nsrc, int ntr, size_t *bytes);

/* ---- Device-memory arena ---------------------------------------------------
 * No reference counterpart (the reference allocates per process with
 * malloc).  One hipMalloc of `bytes`, carved into the device buffers of the
 * contexts created in it (best fit, 256-byte aligned, free neighbours
 * coalesced).  Creating and destroying such contexts costs host bookkeeping
 * only: hipMalloc / hipFree cost time that grows with the size and serialise
 * with the process's other HIP calls, which stalls a server that admits
 * gradient jobs of tens of GB while others run.  Thread-safe.  A context
 * whose buffers do not fit fails with DN_ERR_OOM and leaves the arena as it
 * was. */
typedef struct dn_arena dn_arena;
int dn_arena_create(int device, size_t bytes, dn_arena **out);
/* DN_ERR_STATE while contexts still hold arena memory. */
int dn_arena_destroy(dn_arena *a);
int dn_arena_info(dn_arena *a, size_t *free_bytes, size_t *largest_free, size_t *total);
/* *fits = 1 when the buffers of a context with these parameters, nsrc
 * sources and ntr receivers would fit in the arena's free blocks now. */
int dn_arena_fits(dn_arena *a, const dn_params *p, int nsrc, int ntr, int *fits);
/* dn_ctx_create on the arena's device, with every device buffer (including
 * those of dn_set_sources / dn_set_receivers) taken from the arena. */
int dn_ctx_create_in(const dn_params *p, dn_arena *a, dn_ctx **out);

/* ---- Multi-shot engine ----------------------------------------------------
 * No reference counterpart: the reference runs the shots of one process one
 * after another (src/PSV/FD_PSV.c:271-354).  The engine advances many
 * prepared contexts (model, sources and receivers set) together: each time
 * step is one launch per sweep for all shots of the same FD order,
 * relaxation count, direction and tile shape, so shots of any grid size and
 * length fill the GPU together.  Results equal dn_run_forward /
 * dn_run_adjoint of each context alone, energy weights (dn_set_energy_weights)
 * included.  While attached, a context must not be used through the other
 * entry points.  Moment-tensor shots (type 5) and VTI/TTI shots are not
 * accepted. */
typedef struct dn_engine dn_engine;
int dn_engine_create(int device, dn_engine **out);
int dn_engine_destroy(dn_engine *e);
/* phase 0: forward run (stores snapshots when params.mode > 0); phase 1:
 * adjoint run with correlation, after a finished phase 0 and
 * dn_calc_residuals (or dn_set_adjoint_sources). */
int dn_engine_attach(dn_engine *e, dn_ctx *ctx, int phase);
/* Advance every attached shot by up to nsteps time steps; shots that reach
 * their NT are detached.  *running (may be NULL) = shots still attached.
 * Same as dn_engine_launch followed by dn_engine_wait. */
int dn_engine_step(dn_engine *e, int nsteps, int *running);
/* Enqueue up to nsteps time steps of every attached shot and return at once.
 * Until dn_engine_wait, the host may attach further contexts (they join the
 * next launch) and read or release contexts that are not attached, so that
 * host work overlaps the GPU run; no second launch. */
int dn_engine_launch(dn_engine *e, int nsteps);
/* Wait for the launched run; shots that reached their NT are detached. */
int dn_engine_wait(dn_engine *e, int *running);
/* 1 running, 0 finished, DN_ERR_UNSTABLE if its NaN probe tripped. */
int dn_engine_status(dn_engine *e, dn_ctx *ctx, int *status);
/* Drop the engine's record of a finished context (before destroying it). */
int dn_engine_forget(dn_engine *e, dn_ctx *ctx);
/* Replay runs of `steps` time steps from a captured hipGraph (0: plain
 * launches).  The graph is re-captured when the set of shots changes. */
int dn_engine_set_graph(dn_engine *e, int steps);

#ifdef __cplusplus
}
#endif
#endif /* DENISE_HIP_H */
