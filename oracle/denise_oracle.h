/*
 * CPU oracle for the DENISE Black Edition 2D P/SV time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * It restates, in plain C, the arithmetic of the reference's P/SV path
 * (operation order and the float/double promotion points of the reference's
 * C literals included).  Built twice from denise_oracle.c:
 *   liboracle_f32.so : storage float, reference promotions kept
 *                      (the "reference FP32 CPU path")
 *   liboracle_f64.so : storage and arithmetic double ("FP64 oracle")
 * Every function cites the reference file:line it follows.
 *
 * Indexing follows the reference: grid j=1..NY (depth), i=1..NX, arrays padded
 * by nd = FDORDER/2+1 cells on every side (src/PSV/alloc_PSV.c:21-36).
 * Model inputs are x-major / y-fastest float32 (the reference's model-file
 * order, src/PSV/readmod_elastic_PSV.c:72-76).
 */
#ifndef DENISE_ORACLE_H
#define DENISE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Run parameters: the subset of the reference's globals (include/globvar.h)
 * that the P/SV path reads. */
typedef struct orc_params {
    int nx, ny, nt;
    int fdorder, maxrelerror;
    int fw, free_surf, boundary;
    int L;
    int seismo, ndt;
    int invmat1, grad_form, dtinv, quelltypb;
    int mode;           /* global MODE: 0 forward, 1 FWI (snapshot store on) */
    float dh, dt;
    float damping, fpml, npower, k_max_pml;
    float ts;           /* TS (1/max source fc after sources.c:92) */
    float fl[8];        /* relaxation frequencies FL[1..L] */
} orc_params;

typedef struct orc_ctx orc_ctx;

/* field ids for orc_get_field (padded NR layout, rows -nd+1..NY+nd,
 * columns -nd+1..NX+nd, x fastest) */
enum {
    ORC_VX = 0, ORC_VY, ORC_SXX, ORC_SYY, ORC_SXY, ORC_VXP1, ORC_VYP1,
    ORC_UX, ORC_UY, ORC_UXY,
    ORC_RIP, ORC_RJP, ORC_UIPJP, ORC_PRHO, ORC_PPI, ORC_PU,
    ORC_FIPJP, ORC_F, ORC_G,
    /* anisotropic material: VTI c11, c13, c33, harmonic c44; TTI rotated
     * d11, d13, d15, d15h, d33, d35, d35h, d55h */
    ORC_C11, ORC_C13, ORC_C33, ORC_C44H,
    ORC_D11, ORC_D13, ORC_D15, ORC_D15H, ORC_D33, ORC_D35, ORC_D35H, ORC_D55H,
    /* SH shear moduli at the sxz / syz points */
    ORC_UIP, ORC_UJP,
    ORC_NFIELD
};

orc_ctx *orc_create(const orc_params *p, const float *vp, const float *vs,
                    const float *rho, const float *taus, const float *taup);
/* Anisotropic elastic forward modelling (L = 0, MODE 0): physics 3 VTI
 * (rho, c11, c13, c33, c44; theta unused) or 4 TTI (theta: tilt in degrees,
 * as in the model file MFILE.theta).  All NX*NY x-major. */
orc_ctx *orc_create_aniso(const orc_params *p, int physics, const float *rho, const float *c11,
                          const float *c13, const float *c33, const float *c44, const float *theta);
/* Acoustic forward modelling (PHYSICS 2, L = 0, MODE 0): vp and rho, NX*NY
 * x-major; the pressure is field ORC_SXX. */
orc_ctx *orc_create_ac(const orc_params *p, const float *vp, const float *rho);
/* SH forward modelling (PHYSICS 5, L = 0, MODE 0, SEISMO 0 or 1): vs (INVMAT1
 * 3: mu) and rho, NX*NY x-major; vz is field ORC_VX, sxz ORC_SXX, syz
 * ORC_SYY, 1/rho ORC_RIP; vz is recorded as seismogram component 1. */
orc_ctx *orc_create_sh(const orc_params *p, const float *vs, const float *rho);
orc_ctx *orc_create_sh_visc(const orc_params *p, const float *vs, const float *rho, const float *taus);
void orc_destroy(orc_ctx *c);
const char *orc_error(void);
void orc_set_threads(int n);

int orc_set_sources(orc_ctx *c, int nsrc, const float *srcpos, const float *signals);
int orc_set_receivers(orc_ctx *c, int ntr, const int *recpos);

/* psv() mode 0 for nsteps (<= NT) time steps. */
int orc_forward(orc_ctx *c, int nsteps);
/* one time step on the current state; zero all wavefields; overwrite a
 * wavefield (ids ORC_VX..ORC_UXY, padded layout, element type real) */
int orc_step(orc_ctx *c, int nt, int mode);
int orc_zero(orc_ctx *c);
int orc_set_field(orc_ctx *c, int id, const void *in);
/* calc_res() LNORM=2 on vx / vy; returns the L2 term; fills the adjoint sources. */
double orc_calc_res(orc_ctx *c, const float *obs_vx, const float *obs_vy,
                    const float *obs_p, int testshot);
int orc_set_res_weights(orc_ctx *c, float w_vx, float w_vy);
int orc_set_adjoint(orc_ctx *c, const float *dvx, const float *dvy, const float *dp);
/* psv() mode 1 (adjoint + correlation) for nsteps time steps. */
int orc_adjoint(orc_ctx *c, int nsteps);
/* ass_gradPSV() */
int orc_assemble(orc_ctx *c);

int orc_dims(const orc_ctx *c, int *nd, int *ntdtinv);
int orc_get_field(const orc_ctx *c, int id, void *out);
int orc_get_psi(const orc_ctx *c, int id, void *out);
int orc_get_mem(const orc_ctx *c, int which, int l, void *out);
int orc_get_seis(const orc_ctx *c, int comp, void *out);
int orc_get_adjsrc(const orc_ctx *c, int comp, void *out);
int orc_get_grad(const orc_ctx *c, int which, void *out);
int orc_get_pml(const orc_ctx *c, int which, void *out);
int orc_get_hc(const orc_ctx *c, void *out);
int orc_real_size(void);

/* Pure helpers (no context). */
int orc_holberg(int fdorder, int maxrelerror, float *hc7);
void orc_ricker(int nt, float dt, float tshift, float fc, float amp, float *out);

#ifdef __cplusplus
}
#endif
#endif
