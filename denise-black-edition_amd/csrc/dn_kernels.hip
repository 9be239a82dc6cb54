// HIP kernels of the P/SV time step for gfx950.
//
// One thread per grid cell; a 64-wide wavefront covers 64 consecutive x
// cells of one row, so every field access is a coalesced row segment.  All
// arithmetic keeps the reference's operation order and float/double
// promotion points; the library is compiled with -ffp-contract=off so no
// multiply-add is fused and results match the reference's FP32 CPU path.
#include <hip/hip_runtime.h>

#include "dn_internal.h"

namespace dn {

namespace {

constexpr int BX = 64;  // threads along x (one wavefront per row segment)
constexpr int BY = 4;   // rows per block

__device__ __forceinline__ float cpml(float &psi, float b, float a, float K, float d)
{
    psi = b * psi + a * d;
    return d / K + psi;
}

// ---------------------------------------------------------------------------
// velocity sweep: src/PSV/update_v_PML_PSV.c:14-786
// SW = 0 forward (vxp1 = b*dsigma/DH when write_vp1), SW = 1 adjoint
// (vxp1 += DT*v, GRAD_FORM 1; vxp1 = v, GRAD_FORM 2).
// ---------------------------------------------------------------------------
template <int FDOH, int SW>
__global__ __launch_bounds__(BX *BY) void k_update_v(KArgs a, int write_vp1)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    if (i > a.nx || j > a.ny) return;
    const long P = a.pitch;
    const long o = (long)(j - 1) * P + (i - 1);
    const float *sxx = a.sxx + o, *syy = a.syy + o, *sxy = a.sxy + o;
    float sxx_x = a.hc[1] * (sxx[1] - sxx[0]);
    float sxy_x = a.hc[1] * (sxy[0] - sxy[-1]);
    float sxy_y = a.hc[1] * (sxy[0] - sxy[-P]);
    float syy_y = a.hc[1] * (syy[P] - syy[0]);
#pragma unroll
    for (int m = 2; m <= FDOH; ++m) {
        sxx_x = sxx_x + a.hc[m] * (sxx[m] - sxx[1 - m]);
        sxy_x = sxy_x + a.hc[m] * (sxy[m - 1] - sxy[-m]);
        sxy_y = sxy_y + a.hc[m] * (sxy[(m - 1) * P] - sxy[-m * P]);
        syy_y = syy_y + a.hc[m] * (syy[m * P] - syy[(1 - m) * P]);
    }
    if (!a.boundary && i <= a.fw) {
        const long q = (long)(j - 1) * a.fw2 + (i - 1);
        sxx_x = cpml(a.psi_sxx_x[q], a.b_xh[i], a.a_xh[i], a.K_xh[i], sxx_x);
        sxy_x = cpml(a.psi_sxy_x[q], a.b_x[i], a.a_x[i], a.K_x[i], sxy_x);
    }
    if (!a.boundary && i >= a.nx - a.fw + 1) {
        const int h = i - a.nx + 2 * a.fw;
        const long q = (long)(j - 1) * a.fw2 + (h - 1);
        sxx_x = cpml(a.psi_sxx_x[q], a.b_xh[h], a.a_xh[h], a.K_xh[h], sxx_x);
        sxy_x = cpml(a.psi_sxy_x[q], a.b_x[h], a.a_x[h], a.K_x[h], sxy_x);
    }
    if (!a.free_surf && j <= a.fw) {
        const long q = (long)(j - 1) * a.nx + (i - 1);
        syy_y = cpml(a.psi_syy_y[q], a.b_yh[j], a.a_yh[j], a.K_yh[j], syy_y);
        sxy_y = cpml(a.psi_sxy_y[q], a.b_y[j], a.a_y[j], a.K_y[j], sxy_y);
    }
    if (j >= a.ny - a.fw + 1) {
        const int h = j - a.ny + 2 * a.fw;
        const long q = (long)(h - 1) * a.nx + (i - 1);
        syy_y = cpml(a.psi_syy_y[q], a.b_yh[h], a.a_yh[h], a.K_yh[h], syy_y);
        sxy_y = cpml(a.psi_sxy_y[q], a.b_y[h], a.a_y[h], a.K_y[h], sxy_y);
    }
    const float rip = a.rip[o], rjp = a.rjp[o];
    const float vx = a.vx[o], vy = a.vy[o];
    if (SW == 0) {
        if (write_vp1) {
            a.vxp1[o] = rip * (sxx_x + sxy_y) / a.dh;
            a.vyp1[o] = rjp * (sxy_x + syy_y) / a.dh;
        }
    } else {
        if (a.grad_form == 1) {
            a.vxp1[o] = a.vxp1[o] + a.dt * vx;
            a.vyp1[o] = a.vyp1[o] + a.dt * vy;
        } else if (a.grad_form == 2) {
            a.vxp1[o] = vx;
            a.vyp1[o] = vy;
        }
    }
    a.vx[o] = vx + a.dt * rip * (sxx_x + sxy_y) / a.dh;
    a.vy[o] = vy + a.dt * rjp * (sxy_x + syy_y) / a.dh;
}

// stress-sweep derivatives with CPML: src/PSV/update_s_elastic_PML_PSV.c:360-433
template <int FDOH>
__device__ __forceinline__ void stress_derivs(const KArgs &a, int i, int j, long o, float dhi,
                                              float &vxx, float &vyy, float &vxy, float &vyx)
{
    const long P = a.pitch;
    const float *vx = a.vx + o, *vy = a.vy + o;
    vxx = a.hc[1] * (vx[0] - vx[-1]);
    vyx = a.hc[1] * (vy[1] - vy[0]);
    vxy = a.hc[1] * (vx[P] - vx[0]);
    vyy = a.hc[1] * (vy[0] - vy[-P]);
#pragma unroll
    for (int m = 2; m <= FDOH; ++m) {
        vxx = vxx + a.hc[m] * (vx[m - 1] - vx[-m]);
        vyx = vyx + a.hc[m] * (vy[m] - vy[1 - m]);
        vxy = vxy + a.hc[m] * (vx[m * P] - vx[(1 - m) * P]);
        vyy = vyy + a.hc[m] * (vy[(m - 1) * P] - vy[-m * P]);
    }
    vxx = vxx * dhi;
    vyx = vyx * dhi;
    vxy = vxy * dhi;
    vyy = vyy * dhi;
    if (!a.boundary && i <= a.fw) {
        const long q = (long)(j - 1) * a.fw2 + (i - 1);
        vxx = cpml(a.psi_vxx[q], a.b_x[i], a.a_x[i], a.K_x[i], vxx);
        vyx = cpml(a.psi_vyx[q], a.b_xh[i], a.a_xh[i], a.K_xh[i], vyx);
    }
    if (!a.boundary && i >= a.nx - a.fw + 1) {
        const int h = i - a.nx + 2 * a.fw;
        const long q = (long)(j - 1) * a.fw2 + (h - 1);
        vxx = cpml(a.psi_vxx[q], a.b_x[h], a.a_x[h], a.K_x[h], vxx);
        vyx = cpml(a.psi_vyx[q], a.b_xh[h], a.a_xh[h], a.K_xh[h], vyx);
    }
    if (!a.free_surf && j <= a.fw) {
        const long q = (long)(j - 1) * a.nx + (i - 1);
        vyy = cpml(a.psi_vyy[q], a.b_y[j], a.a_y[j], a.K_y[j], vyy);
        vxy = cpml(a.psi_vxy[q], a.b_yh[j], a.a_yh[j], a.K_yh[j], vxy);
    }
    if (j >= a.ny - a.fw + 1) {
        const int h = j - a.ny + 2 * a.fw;
        const long q = (long)(h - 1) * a.nx + (i - 1);
        vyy = cpml(a.psi_vyy[q], a.b_y[h], a.a_y[h], a.K_y[h], vyy);
        vxy = cpml(a.psi_vxy[q], a.b_yh[h], a.a_yh[h], a.K_yh[h], vxy);
    }
}

// elastic stress sweep: src/PSV/update_s_elastic_PML_PSV.c:13-742
// f = rho*vs^2, g = rho*(vp^2 - 2 vs^2) precomputed with the same float ops;
// the normal stresses are accumulated in double like the reference's 2.0*f
// literal promotion.
template <int FDOH>
__global__ __launch_bounds__(BX *BY) void k_update_s_el(KArgs a, int mode0_gf2)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    if (i > a.nx || j > a.ny) return;
    const long o = (long)(j - 1) * a.pitch + (i - 1);
    float vxx, vyy, vxy, vyx;
    stress_derivs<FDOH>(a, i, j, o, a.dhi_el, vxx, vyy, vxy, vyx);
    const float fipjp = a.muipjp[o], f = a.f[o], g = a.g[o];
    const float gd = g * (vxx + vyy);
    const double f2 = 2.0 * (double)f;
    if (mode0_gf2) {
        a.ux[o] = (float)(((double)gd + f2 * (double)vxx) / (double)a.dt);
        a.uy[o] = (float)(((double)gd + f2 * (double)vyy) / (double)a.dt);
        a.uxy[o] = fipjp * (vyx + vxy) / a.dt;
    }
    a.sxy[o] = a.sxy[o] + fipjp * (vyx + vxy);
    a.sxx[o] = (float)((double)a.sxx[o] + ((double)gd + f2 * (double)vxx));
    a.syy[o] = (float)((double)a.syy[o] + ((double)gd + f2 * (double)vyy));
}

// viscoelastic stress + memory variables: src/PSV/update_s_visc_PML_PSV.c:13-825
// LT > 0: L known at compile time; LT == 0: runtime a.L.
template <int FDOH, int LT>
__global__ __launch_bounds__(BX *BY) void k_update_s_vis(KArgs a, int mode0_gf2)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    if (i > a.nx || j > a.ny) return;
    const int L = LT > 0 ? LT : a.L;
    const long o = (long)(j - 1) * a.pitch + (i - 1);
    float vxx, vyy, vxy, vyx;
    stress_derivs<FDOH>(a, i, j, o, a.dhi_vis, vxx, vyy, vxy, vyx);
    float sumr = 0.0f, sump = 0.0f, sumq = 0.0f;
#pragma unroll
    for (int l = 0; l < (LT > 0 ? LT : MAXL); ++l) {
        if (LT == 0 && l >= L) break;
        sumr += a.rmem[l * a.plane + o];
        sump += a.pmem[l * a.plane + o];
        sumq += a.qmem[l * a.plane + o];
    }
    const float fipjp = a.muipjp[o], f = a.f[o], g = a.g[o];
    const float gd = g * (vxx + vyy);
    const double f2 = 2.0 * (double)f;
    a.sxy[o] = a.sxy[o] + ((fipjp * (vxy + vyx)) + (a.dthalbe * sumr));
    a.sxx[o] = (float)((double)a.sxx[o] + (((double)gd - f2 * (double)vyy) + (double)(a.dthalbe * sump)));
    a.syy[o] = (float)((double)a.syy[o] + (((double)gd - f2 * (double)vxx) + (double)(a.dthalbe * sumq)));
    sumr = sump = sumq = 0.0f;
    const float vsum_xy = vxy + vyx, vsum = vxx + vyy;
#pragma unroll
    for (int l = 0; l < (LT > 0 ? LT : MAXL); ++l) {
        if (LT == 0 && l >= L) break;
        const long ol = l * a.plane + o;
        const float r = a.bip[l + 1] * (a.rmem[ol] * a.cip[l + 1] - (a.dipvis[ol] * vsum_xy));
        const double d2 = 2.0 * (double)a.dvis[ol];
        const float ev = a.evis[ol] * vsum;
        const float p = (float)((double)a.bjm[l + 1] * ((double)(a.pmem[ol] * a.cjm[l + 1] - ev) + d2 * (double)vyy));
        const float q = (float)((double)a.bjm[l + 1] * ((double)(a.qmem[ol] * a.cjm[l + 1] - ev) + d2 * (double)vxx));
        a.rmem[ol] = r;
        a.pmem[ol] = p;
        a.qmem[ol] = q;
        sumr += r;
        sump += p;
        sumq += q;
    }
    a.sxy[o] = a.sxy[o] + (a.dthalbe * sumr);
    a.sxx[o] = a.sxx[o] + (a.dthalbe * sump);
    a.syy[o] = a.syy[o] + (a.dthalbe * sumq);
    if (mode0_gf2) {
        a.ux[o] = sump;
        a.uy[o] = sumq;
        a.uxy[o] = sumr;
    }
}

// free surface at row j=1 by imaging: src/PSV/surface_elastic_PML_PSV.c:10-105
// and src/PSV/surface_visc_PML_PSV.c:10-173.  One thread per column.
template <int FDOH>
__global__ __launch_bounds__(256) void k_surface(KArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i > a.nx) return;
    const long P = a.pitch;
    const long o = (long)(i - 1);  // row 1
    a.syy[o] = 0.0f;
    if (a.L > 0)
        for (int l = 0; l < a.L; ++l) a.qmem[l * a.plane + o] = 0.0f;
    float vxx = 0.0f, vyy = 0.0f;
#pragma unroll
    for (int m = 1; m <= FDOH; ++m) {
        a.syy[o - m * P] = -a.syy[o + m * P];
        a.sxy[o - m * P] = -a.sxy[o + (m - 1) * P];
        vxx += a.hc[m] * (a.vx[o + m - 1] - a.vx[o - m]);
        vyy += a.hc[m] * (a.vy[o + (m - 1) * P] - a.vy[o - m * P]);
    }
    vxx *= a.dhi_vis;
    vyy *= a.dhi_vis;
    if (!a.boundary && i <= a.fw) vxx = cpml(a.psi_vxxs[(i - 1)], a.b_x[i], a.a_x[i], a.K_x[i], vxx);
    if (!a.boundary && i >= a.nx - a.fw + 1) {
        const int h = i - a.nx + 2 * a.fw;
        vxx = cpml(a.psi_vxxs[(h - 1)], a.b_x[h], a.a_x[h], a.K_x[h], vxx);
    }
    if (a.L == 0) {
        const float g = a.g[o];
        const float fjm = a.f[o] * 2.0f;
        a.sxx[o] = a.sxx[o] + -a.dt * ((g * g) / (g + fjm) * vxx + g * vyy);
    } else {
        const float fjm = a.fs_fjm[i], g = a.fs_g[i], dme = a.fs_dme[i];
        float sump = 0.0f;
        for (int l = 0; l < a.L; ++l) sump += a.pmem[l * a.plane + o];
        a.sxx[o] = a.sxx[o] + (-(a.dt * (g - fjm) * (g - fjm) * vxx / g) - (a.dt * (g - fjm) * vyy) -
                               (a.dthalbe * sump));
        sump = 0.0f;
        const double fg1 = (double)(fjm / g) - 1.0;
        for (int l = 0; l < a.L; ++l) {
            const long ol = l * a.plane + o;
            const float pv = (float)((double)a.pmem[ol] +
                                     (double)a.fs_bjm[l + 1] * ((((double)dme * fg1) * (double)vxx) - (double)(dme * vyy)));
            a.pmem[ol] = pv;
            sump += pv;
        }
        a.sxx[o] = a.sxx[o] + (a.dthalbe * sump);
    }
}

// src/psource.c:23-38 / src/msource.c:20-40 / rotated forces of
// src/PSV/update_v_PML_PSV.c:739-752.  Sources are applied one after another
// by a single thread so coinciding positions accumulate in source order.
__device__ __forceinline__ float src_amp(const float *s, int nt, int NT, float dt)
{
    float amp = 0.0f;
    if (nt == 1) amp = s[1] / dt;
    if (nt > 1 && nt < NT) amp = (s[nt] - s[nt - 2]) / dt;
    if (nt == NT) amp = -s[nt - 2] / dt;
    return amp;
}

__global__ void k_force_src(KArgs a, DevAcq q, int nt)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int l = 0; l < q.nsrc; ++l) {
        const long o = (long)(q.src_iy[l] - 1) * a.pitch + (q.src_ix[l] - 1);
        const float s = q.signals[(long)l * q.nt + (nt - 1)];
        const int type = q.src_type[l];
        if (type == 2) a.vx[o] = a.vx[o] + s;
        if (type == 3) a.vy[o] = a.vy[o] + s;
        if (type == 4) {
            a.vx[o] = (float)((double)a.vx[o] + (-q.src_sin[l]) * (double)s);
            a.vy[o] = (float)((double)a.vy[o] + q.src_cos[l] * (double)s);
        }
    }
}

__global__ void k_psource(KArgs a, DevAcq q, int nt)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int l = 0; l < q.nsrc; ++l) {
        const long o = (long)(q.src_iy[l] - 1) * a.pitch + (q.src_ix[l] - 1);
        const float amp = src_amp(q.signals + (long)l * q.nt, nt, q.nt, a.dt);
        a.sxx[o] = a.sxx[o] + amp;
        a.syy[o] = a.syy[o] + amp;
    }
}

__global__ void k_msource(KArgs a, DevAcq q, int nt)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const long P = a.pitch;
    for (int l = 0; l < q.nsrc; ++l) {
        const long o = (long)(q.src_iy[l] - 1) * P + (q.src_ix[l] - 1);
        const float amp = src_amp(q.signals + (long)l * q.nt, nt, q.nt, a.dt);
        a.sxx[o] = a.sxx[o] + amp * 0.0f;
        a.syy[o] = a.syy[o] + amp * 0.0f;
        const double add = 0.25 * (double)amp * 1.0;
        a.sxy[o] = (float)((double)a.sxy[o] + add);
        a.sxy[o - 1] = (float)((double)a.sxy[o - 1] + add);
        a.sxy[o - P] = (float)((double)a.sxy[o - P] + add);
        a.sxy[o - P - 1] = (float)((double)a.sxy[o - P - 1] + add);
    }
}

// adjoint forces at the receivers: src/PSV/update_v_PML_PSV.c:755-768;
// receiver cells are distinct, so one thread per receiver.
__global__ void k_adj_force(KArgs a, DevAcq q, int nt, int qb)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= q.ntr) return;
    const long o = (long)(q.rec_iy[l] - 1) * a.pitch + (q.rec_ix[l] - 1);
    const float sx = q.adj[0][(long)l * q.nt + (nt - 1)];
    const float sy = q.adj[1][(long)l * q.nt + (nt - 1)];
    if (a.grad_form == 1 || a.grad_form == 2) {
        if (qb == 1) {
            a.vx[o] = a.vx[o] + sx;
            a.vy[o] = a.vy[o] + sy;
        }
        if (qb == 2 || qb == 6 || qb == 7) a.vy[o] = a.vy[o] + sy;
        if (qb == 3 || qb == 5 || qb == 7) a.vx[o] = a.vx[o] + sx;
    }
}

// adjoint pressure source: src/psource.c:41-52
__global__ void k_adj_psource(KArgs a, DevAcq q, int nt)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= q.ntr) return;
    const long o = (long)(q.rec_iy[l] - 1) * a.pitch + (q.rec_ix[l] - 1);
    const float s = q.adj[2][(long)l * q.nt + (nt - 1)];
    a.sxx[o] = a.sxx[o] + s;
    a.syy[o] = a.syy[o] + s;
}

// receiver sampling (src/seismo_ssg.c:8-108) and the NaN probe of
// src/PSV/psv.c:99-103 (thread 0 looks at vy[NY/2][NX/2] after step nt, which
// is what the reference tests at the start of step nt+1).
template <int FDOH>
__global__ void k_end_step(KArgs a, DevAcq q, int nt, int seismo, int invmat1, int probe)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0 && probe) {
        const long oc = (long)(a.ny / 2 - 1) * a.pitch + (a.nx / 2 - 1);
        if (__builtin_isnan(a.vy[oc]) && *q.nan_step == 0) *q.nan_step = nt + 1;
    }
    if (!seismo || t >= q.ntr) return;
    const long P = a.pitch;
    const long o = (long)(q.rec_iy[t] - 1) * P + (q.rec_ix[t] - 1);
    const long so = (long)t * q.nt + (nt - 1);
    if (seismo == 1 || seismo == 4) {
        q.sec[0][so] = a.vx[o];
        q.sec[1][so] = a.vy[o];
    }
    if (seismo == 2 || seismo == 4) q.sec[2][so] = -a.sxx[o] - a.syy[o];
    if (seismo == 3 || seismo == 4) {
        float vxx = 0.0f, vyy = 0.0f, vyx = 0.0f, vxy = 0.0f;
        const float *vx = a.vx + o, *vy = a.vy + o;
#pragma unroll
        for (int m = 1; m <= FDOH; ++m) {
            vxx += a.hc[m] * (vx[m - 1] - vx[-m]);
            vyy += a.hc[m] * (vy[(m - 1) * P] - vy[-m * P]);
            vyx += a.hc[m] * (vy[m] - vy[1 - m]);
            vxy += a.hc[m] * (vx[m * P] - vx[(1 - m) * P]);
        }
        vxx *= a.dhi_el;
        vyy *= a.dhi_el;
        vyx *= a.dhi_el;
        vxy *= a.dhi_el;
        q.sec[3][so] = (float)((double)(vxx + vyy) * q.rec_sqlam[t]);
        q.sec[4][so] = (float)((double)(vxy - vyx) * q.rec_sqmu[t]);
    }
}

// forward snapshot store (src/PSV/psv.c:240-309); snapshot k of each field is
// a compact NX*NY row-major plane.
__global__ __launch_bounds__(BX *BY) void k_store(KArgs a, float *f0, float *f1, float *f2, float *f3,
                                                  float *f4, long k)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    if (i > a.nx || j > a.ny) return;
    const long o = (long)(j - 1) * a.pitch + (i - 1);
    const long s = k * (long)a.nx * a.ny + (long)(j - 1) * a.nx + (i - 1);
    f0[s] = a.vxp1[o];
    f1[s] = a.vyp1[o];
    if (a.grad_form == 1) {
        f2[s] = a.sxx[o];
        f3[s] = a.syy[o];
        f4[s] = a.sxy[o];
    } else {
        f2[s] = a.ux[o];
        f3[s] = a.uy[o];
        f4[s] = a.uxy[o];
    }
}

// adjoint cross-correlation: src/PSV/psv.c:313-387
__global__ __launch_bounds__(BX *BY) void k_correlate(KArgs a, const float *f0, const float *f1,
                                                      const float *f2, const float *f3, const float *f4,
                                                      long k, float *wc, float *wcu, float *wcr,
                                                      int invmat1)
{
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    if (i > a.nx || j > a.ny) return;
    const long o = (long)(j - 1) * a.pitch + (i - 1);
    const long s = k * (long)a.nx * a.ny + (long)(j - 1) * a.nx + (i - 1);
    const long g = (long)(j - 1) * a.nx + (i - 1);
    const float frx = f0[s], fry = f1[s], fx = f2[s], fy = f3[s], fu = f4[s];
    const float axx = a.sxx[o], ayy = a.syy[o], axy = a.sxy[o];
    wcr[g] = wcr[g] + ((a.vxp1[o] * frx) + (a.vyp1[o] * fry));
    wc[g] = wc[g] + (fx + fy) * (axx + ayy);
    float muss, lamss;
    if (invmat1 == 3) {
        muss = a.vs[o];
        lamss = a.vp[o];
    } else {
        muss = a.rho[o] * a.vs[o] * a.vs[o];
        lamss = (float)((double)(a.rho[o] * a.vp[o] * a.vp[o]) - 2.0 * (double)muss);
    }
    if (muss > 0.0f) {
        const double mm = (double)(muss * muss);
        const double lm = (double)((lamss + muss) * (lamss + muss));
        double add;
        if (a.grad_form == 1) {
            add = ((1.0 / mm) * (double)(fu * axy)) + ((0.25 * (double)((fx + fy) * (axx + ayy))) / lm) +
                  ((0.25 * (double)((fx - fy) * (axx - ayy))) / mm);
        } else {
            const double lmd = (double)(lamss + muss), mud = (double)muss;
            const float tmp = (float)((1.0 / (4.0 * lmd * lmd)) - (1.0 / (4.0 * mud * mud)));
            const float tmp1 = (float)((1.0 / (4.0 * lmd * lmd)) + (1.0 / (4.0 * mud * mud)));
            add = ((1.0 / mm) * (double)(fu * axy)) + (double)(tmp1 * (fx * axx + fy * ayy)) +
                  (double)(tmp * (fx * ayy + fy * axx));
        }
        wcu[g] = (float)((double)wcu[g] + add);
    }
}

// src/PSV/ass_gradPSV.c:10-155 (gradients stored x-major like the files)
__global__ __launch_bounds__(256) void k_assemble(KArgs a, const float *wc, const float *wcu,
                                                  const float *wcr, float *gvp, float *gvs, float *grho,
                                                  int invmat1)
{
    const long n = (long)a.nx * a.ny;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t / a.ny) + 1, j = (int)(t % a.ny) + 1;
    const long o = (long)(j - 1) * a.pitch + (i - 1);
    const long g = (long)(j - 1) * a.nx + (i - 1);
    const float ppi = a.vp[o], pu = a.vs[o], prho = a.rho[o];
    const float DT = a.dt;
    float lam = -DT * wc[g];
    float vpg, vsg, rhog;
    const float mu = -DT * wcu[g];
    const float rho_s = -DT * wcr[g];
    if (invmat1 == 3) {
        const float muss = pu, lamss = ppi;
        if (a.grad_form == 1 && (muss > 0.0f || lamss > 0.0f)) {
            const double lm = (double)(lamss + muss);
            lam = (float)((1.0 / (4.0 * lm * lm)) * (double)lam);
        }
        vpg = lam;
        vsg = mu;
        rhog = rho_s;
    } else {
        const float muss = prho * pu * pu;
        const float lamss = (float)((double)(prho * ppi * ppi) - 2.0 * (double)muss);
        if (muss > 0.0f || lamss > 0.0f) {
            const double lm = (double)(lamss + muss);
            lam = (float)((1.0 / (4.0 * lm * lm)) * (double)lam);
        }
        vpg = (float)(2.0 * (double)ppi * (double)prho * (double)lam);
        vsg = (float)((-4.0 * (double)prho * (double)pu * (double)lam) + 2.0 * (double)prho * (double)pu * (double)mu);
        rhog = (float)(((((double)(ppi * ppi) - (2.0 * (double)pu * (double)pu)) * (double)lam) +
                        (double)(pu * pu * mu)) + (double)rho_s);
    }
    gvp[t] = vpg;
    gvs[t] = vsg;
    grho[t] = rhog;
}

// time-integrated, time-reversed L2 residual of one trace
// (src/calc_res.c:406-452 with src/inseis.c:96-103): one thread per trace.
__global__ void k_residual(DevAcq q, int comp, const float *obs, float dt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= q.ntr) return;
    const long ns = q.nt;
    const float *syn = q.sec[comp] + (long)t * ns;
    const float *ob = obs + (long)t * ns;
    float *diff = q.adj[comp == 2 ? 2 : comp] + (long)t * ns;
    float intseis = 0.0f;
    for (long jj = 1; jj <= ns; ++jj) {
        const float d = jj == 1 ? syn[0] : ob[jj - 1];
        intseis += dt * (syn[jj - 1] - d);
        diff[ns - jj] = intseis;
    }
}

// misfit 0.5*sum(residual^2) in the reference's summation order
// (src/calc_res.c:553-556): one sequential chain in double.
__global__ void k_l2(DevAcq q, int comp)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const long ns = q.nt;
    double L2 = *q.l2;
    for (int t = 0; t < q.ntr; ++t) {
        const float *diff = q.adj[comp] + (long)t * ns;
        for (long jj = ns; jj >= 1; --jj) L2 += 0.5 * (double)diff[jj - 1] * (double)diff[jj - 1];
    }
    *q.l2 = L2;
}

inline dim3 grid2d(const KArgs &a) { return dim3((a.nx + BX - 1) / BX, (a.ny + BY - 1) / BY); }

template <int FDOH>
void update_v_fd(const KArgs &a, int sw, int write_vp1, hipStream_t s)
{
    if (sw == 0)
        hipLaunchKernelGGL((k_update_v<FDOH, 0>), grid2d(a), dim3(BX, BY), 0, s, a, write_vp1);
    else
        hipLaunchKernelGGL((k_update_v<FDOH, 1>), grid2d(a), dim3(BX, BY), 0, s, a, write_vp1);
}

template <int FDOH>
void update_s_fd(const KArgs &a, int mode0_gf2, hipStream_t s)
{
    switch (a.L) {
    case 0: hipLaunchKernelGGL((k_update_s_el<FDOH>), grid2d(a), dim3(BX, BY), 0, s, a, mode0_gf2); break;
    case 1: hipLaunchKernelGGL((k_update_s_vis<FDOH, 1>), grid2d(a), dim3(BX, BY), 0, s, a, mode0_gf2); break;
    case 2: hipLaunchKernelGGL((k_update_s_vis<FDOH, 2>), grid2d(a), dim3(BX, BY), 0, s, a, mode0_gf2); break;
    case 3: hipLaunchKernelGGL((k_update_s_vis<FDOH, 3>), grid2d(a), dim3(BX, BY), 0, s, a, mode0_gf2); break;
    default: hipLaunchKernelGGL((k_update_s_vis<FDOH, 0>), grid2d(a), dim3(BX, BY), 0, s, a, mode0_gf2); break;
    }
}

}  // namespace

#define DN_FD_SWITCH(fdoh, CALL)                 \
    switch (fdoh) {                              \
    case 1: CALL(1); break;                      \
    case 2: CALL(2); break;                      \
    case 3: CALL(3); break;                      \
    case 4: CALL(4); break;                      \
    case 5: CALL(5); break;                      \
    default: CALL(6); break;                     \
    }

void launch_update_v(const KArgs &a, int fdoh, int sw, int write_vp1, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
#define CALL(N) update_v_fd<N>(a, sw, write_vp1, s)
    DN_FD_SWITCH(fdoh, CALL)
#undef CALL
}

void launch_update_s(const KArgs &a, int fdoh, int mode0_gf2, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
#define CALL(N) update_s_fd<N>(a, mode0_gf2, s)
    DN_FD_SWITCH(fdoh, CALL)
#undef CALL
}

void launch_surface(const KArgs &a, int fdoh, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    dim3 g((a.nx + 255) / 256);
#define CALL(N) hipLaunchKernelGGL((k_surface<N>), g, dim3(256), 0, s, a)
    DN_FD_SWITCH(fdoh, CALL)
#undef CALL
}

void launch_sources_fwd(const KArgs &a, const DevAcq &q, int nt, int qtype, int has_force, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    (void)has_force;
    if (qtype == 1) hipLaunchKernelGGL(k_psource, dim3(1), dim3(64), 0, s, a, q, nt);
    if (qtype == 5) hipLaunchKernelGGL(k_msource, dim3(1), dim3(64), 0, s, a, q, nt);
}

void launch_force_sources(const KArgs &a, const DevAcq &q, int nt, void *stream)
{
    hipLaunchKernelGGL(k_force_src, dim3(1), dim3(64), 0, (hipStream_t)stream, a, q, nt);
}

void launch_sources_adj(const KArgs &a, const DevAcq &q, int nt, int quelltypb, void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    if (q.ntr <= 0) return;
    dim3 g((q.ntr + 255) / 256);
    hipLaunchKernelGGL(k_adj_force, g, dim3(256), 0, s, a, q, nt, quelltypb);
}

void launch_adj_psource(const KArgs &a, const DevAcq &q, int nt, void *stream)
{
    if (q.ntr <= 0) return;
    dim3 g((q.ntr + 255) / 256);
    hipLaunchKernelGGL(k_adj_psource, g, dim3(256), 0, (hipStream_t)stream, a, q, nt);
}

void launch_end_step(const KArgs &a, const DevAcq &q, int nt, int seismo, int fdoh, int invmat1, int probe,
                     void *stream)
{
    hipStream_t s = (hipStream_t)stream;
    const int n = (seismo ? q.ntr : 0) > 1 ? q.ntr : 1;
    if (!probe && !(seismo && q.ntr > 0)) return;
    dim3 g((n + 255) / 256);
#define CALL(N) hipLaunchKernelGGL((k_end_step<N>), g, dim3(256), 0, s, a, q, nt, seismo, invmat1, probe)
    DN_FD_SWITCH(fdoh, CALL)
#undef CALL
}

void launch_store(const KArgs &a, float *const fwd[5], long k, void *stream)
{
    hipLaunchKernelGGL(k_store, grid2d(a), dim3(BX, BY), 0, (hipStream_t)stream, a, fwd[0], fwd[1], fwd[2],
                       fwd[3], fwd[4], k);
}

void launch_correlate(const KArgs &a, const float *const fwd[5], long k, float *wc, float *wcu, float *wcr,
                      int invmat1, void *stream)
{
    hipLaunchKernelGGL(k_correlate, grid2d(a), dim3(BX, BY), 0, (hipStream_t)stream, a, fwd[0], fwd[1],
                       fwd[2], fwd[3], fwd[4], k, wc, wcu, wcr, invmat1);
}

void launch_assemble(const KArgs &a, const float *wc, const float *wcu, const float *wcr, float *gvp,
                     float *gvs, float *grho, int invmat1, void *stream)
{
    const long n = (long)a.nx * a.ny;
    dim3 g((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_assemble, g, dim3(256), 0, (hipStream_t)stream, a, wc, wcu, wcr, gvp, gvs, grho,
                       invmat1);
}

void launch_residual(const DevAcq &q, int comp, const float *obs, float dt, void *stream)
{
    if (q.ntr <= 0) return;
    dim3 g((q.ntr + 63) / 64);
    hipLaunchKernelGGL(k_residual, g, dim3(64), 0, (hipStream_t)stream, q, comp, obs, dt);
}

void launch_l2(const DevAcq &q, int comp, void *stream)
{
    hipLaunchKernelGGL(k_l2, dim3(1), dim3(64), 0, (hipStream_t)stream, q, comp);
}

}  // namespace dn
