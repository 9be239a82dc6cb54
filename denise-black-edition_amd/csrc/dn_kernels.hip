// Auxiliary HIP kernels of the P/SV path for gfx950 (moment-tensor source,
// stand-alone snapshot store, gradient assembly, residuals, misfit) and the
// dispatch to the fused sweeps compiled per tile shape.
//
// Numerics: reference operation order, no contraction (see dn_sweeps.inc).
#include <hip/hip_runtime.h>

#include "dn_internal.h"

namespace dn {

namespace {

__device__ __forceinline__ float src_amp(const float *s, int nt, int NT, float dt)
{
    float amp = 0.0f;
    if (nt == 1) amp = s[1] / dt;
    if (nt > 1 && nt < NT) amp = (s[nt] - s[nt - 2]) / dt;
    if (nt == NT) amp = -s[nt - 2] / dt;
    return amp;
}

// src/msource.c:9-42 (M11=0, M12=1, M22=0), after the stress sweep; kept as
// its own launch because it writes four cells around the source, including
// a ghost row when the source sits in row 1.
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

// forward snapshot store as a separate pass (used with the moment-tensor
// source, whose shear-stress increments land after the stress sweep)
__global__ __launch_bounds__(256) void k_store(KArgs a, float *f0, float *f1, float *f2, float *f3, float *f4,
                                               long k)
{
    const long n = (long)a.nx * a.ny;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t % a.nx) + 1, j = (int)(t / a.nx) + 1;
    const long o = (long)(j - 1) * a.pitch + (i - 1);
    const long s = k * n + t;
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

// src/eprecond.c:9-18: W += vx*vx + vy*vy over the grid
__global__ __launch_bounds__(256) void k_energy(KArgs a, float *W)
{
    const long n = (long)a.nx * a.ny;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t % a.nx) + 1, j = (int)(t / a.nx) + 1;
    const long o = (long)(j - 1) * a.pitch + (i - 1);
    const float vx = a.vx[o], vy = a.vy[o];
    W[t] = W[t] + ((vx * vx) + (vy * vy));
}

// src/PSV/ass_gradPSV.c:10-155, for acoustic contexts src/AC/ass_gradAC.c:12-93
// (gradients stored x-major like the files)
__global__ __launch_bounds__(256) void k_assemble(KArgs a, const float *wc, const float *wcu,
                                                  const float *wcr, float *gvp, float *gvs, float *grho,
                                                  int invmat1, int physics)
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
    if (physics == 5) {
        // ass_gradSH.c:19-98 (INV_*_ITER 0): where the INVMAT1 1 / GRAD_FORM 1
        // branches write nothing (vs = 0) the raw sums stay
        float gu = wcu[g], gr = wcr[g];
        if (invmat1 == 1) {
            if (a.grad_form == 1) {
                if (pu > 0.0f) {
                    const float vs3 = pu * pu * pu;
                    gu = (float)((2.0 / (double)prho * (double)vs3) * (double)mu);
                    const float vs2rho2 = (float)(1.0 / (double)(pu * pu * prho * prho));
                    gr = (vs2rho2 * mu) + rho_s;
                }
            } else {
                gu = (float)(2.0 * (double)prho * (double)pu * (double)mu);
                gr = (pu * pu * mu) + rho_s;
            }
        } else {
            gu = mu;
            gr = rho_s;
        }
        gvp[t] = 0.0f;
        gvs[t] = gu;
        grho[t] = gr;
        return;
    }
    if (physics == 2) {
        const float lamss = invmat1 == 3 ? ppi : prho * ppi * ppi;
        if (lamss > 0.0f) {
            if (a.grad_form == 1)
                lam = (float)((1.0 / (4.0 * (double)lamss * (double)lamss)) * (double)lam);
            else
                lam = (float)(-(1.0 / (double)(lamss * lamss)) * (double)lam);
        }
        vpg = invmat1 == 3 ? lam : (float)(2.0 * (double)ppi * (double)prho * (double)lam);
        vsg = 0.0f;
        rhog = invmat1 == 1 ? ((ppi * ppi) * lam) + rho_s : rho_s;
    } else if (invmat1 == 3) {
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
// (src/calc_res.c:406-452 with src/inseis.c:96-103).  The integration is a
// sequential float chain per trace, so each lane owns one trace; a block of
// 64 lanes stages 64-sample chunks of its 64 traces in LDS so that global
// reads and writes run along the traces (coalesced) while each lane walks
// its own trace in the reference's order.  w: the component weight of
// src/calc_res.c:100-123 (COMP_WEIGHT, QUELLTYPB 5-7), applied to the
// synthetic and the observed samples before the residual as there (float
// products); 1 leaves every sample unchanged.
constexpr int RB = 64;  // traces per block = samples per chunk
__global__ __launch_bounds__(RB) void k_residual(DevAcq q, int comp, const float *obs, float dt, float w)
{
    __shared__ float S[RB][RB + 1], O[RB][RB + 1];
    const int lane = threadIdx.x, t0 = blockIdx.x * RB;
    const long ns = q.nt;
    const int nr = q.ntr - t0 < RB ? q.ntr - t0 : RB;
    const float *syn = q.sec[comp];
    float *diff = q.adj[comp];
    const float syn0 = lane < nr ? syn[(long)(t0 + lane) * ns] * w : 0.0f;
    float intseis = 0.0f;
    for (long c0 = 0; c0 < ns; c0 += RB) {
        const int nc = ns - c0 < RB ? (int)(ns - c0) : RB;
        for (int r = 0; r < nr; ++r) {
            const long g = (long)(t0 + r) * ns + c0 + lane;
            if (lane < nc) {
                S[r][lane] = syn[g] * w;
                O[r][lane] = obs[g] * w;
            }
        }
        __syncthreads();
        if (lane < nr) {
            for (int l = 0; l < nc; ++l) {
                const long jj = c0 + l + 1;
                const float d = jj == 1 ? syn0 : O[lane][l];
                intseis += dt * (S[lane][l] - d);
                S[lane][l] = intseis;  // sample ns - jj of the residual
            }
        }
        __syncthreads();
        for (int r = 0; r < nr; ++r)
            if (lane < nc) diff[(long)(t0 + r) * ns + (ns - 1 - c0 - lane)] = S[r][lane];
        __syncthreads();
    }
}

// misfit 0.5*sum(residual^2) of src/calc_res.c:553-556 in double.  Every
// term 0.5*d*d of a float d is exact in double; the terms are summed in a
// fixed tree order (L2_PARTS block ranges of 256 strided lanes each, then a
// 256-lane tree over the block sums), so the result is reproducible run to
// run and differs from the reference's single sequential chain only by
// double rounding.
constexpr int L2_PARTS = 256;
__device__ __forceinline__ double tree256(double *sh, double v)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + h];
        __syncthreads();
    }
    return sh[0];
}

__global__ __launch_bounds__(256) void k_l2_part(DevAcq q, int comp)
{
    __shared__ double sh[256];
    const long n = (long)q.ntr * q.nt;
    const long lo = n * blockIdx.x / L2_PARTS, hi = n * (blockIdx.x + 1) / L2_PARTS;
    const float *diff = q.adj[comp];
    double s = 0.0;
    for (long i = lo + threadIdx.x; i < hi; i += 256) s += 0.5 * (double)diff[i] * (double)diff[i];
    const double tot = tree256(sh, s);
    if (threadIdx.x == 0) q.l2[1 + blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_l2_sum(DevAcq q)
{
    __shared__ double sh[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < L2_PARTS; b += 256) s += q.l2[1 + b];
    const double tot = tree256(sh, s);
    if (threadIdx.x == 0) q.l2[0] = q.l2[0] + tot;
}

}  // namespace

// Tile shapes compiled into the library: X(TBX, BY, TY, name), each in its
// own dn_sweep_<name>.hip translation unit.
// Compiled sweep shapes (tbx, by, ty, vw): one translation unit each
// (dn_sweep_t*.hip scalar columns, dn_sweepv_t*.hip vector columns).
#define DN_TILES(X)                                                                              \
    X(64, 4, 1, 1, t64x4x1) X(128, 2, 1, 1, t128x2x1) X(256, 1, 1, 1, t256x1x1)                   \
    X(64, 4, 2, 1, t64x4x2) X(128, 2, 4, 1, t128x2x4) X(64, 4, 1, 4, t64x4x1v4)                   \
    X(64, 4, 2, 4, t64x4x2v4) X(64, 2, 4, 4, t64x2x4v4) X(64, 4, 1, 2, t64x4x1v2)                 \
    X(128, 2, 2, 2, t128x2x2v2) X(64, 8, 2, 1, t64x8x2) X(64, 4, 2, 2, t64x4x2v2)

#define X(B, Y, T, W, N)                                                                    \
    void launch_vel_##N(const KArgs &, const StepCtl &, const DevAcq &, int, int, int, void *); \
    void launch_str_##N(const KArgs &, const StepCtl &, const DevAcq &, int, int, void *);
DN_TILES(X)
#undef X

int tile_list(Tile *out, int max)
{
    int n = 0;
#define X(B, Y, T, W, N)           \
    if (n < max) {                 \
        out[n].tbx = B;            \
        out[n].by = Y;             \
        out[n].ty = T;             \
        out[n].vw = W;             \
    }                              \
    ++n;
    DN_TILES(X)
#undef X
    return n;
}

#define DN_SAME(B, Y, T, W) (t.tbx == B && t.by == Y && t.ty == T && t.vw == W)

bool tile_supported(const Tile &t)
{
#define X(B, Y, T, W, N) \
    if (DN_SAME(B, Y, T, W)) return true;
    DN_TILES(X)
#undef X
    return false;
}

// The step control carries the tile grid of the sweep being launched.
static StepCtl with_grid(const StepCtl &c0, const Tile &t, const KArgs &a)
{
    StepCtl c = c0;
    c.ntx = t.ntx(a.nx);
    c.ntiles = c.ntx * t.nty(a.ny);
    return c;
}

void launch_vel(const KArgs &a, const StepCtl &c0, const DevAcq &q, const Tile &t, int fdoh, int sw, int ntl,
                void *stream)
{
    const StepCtl c = with_grid(c0, t, a);
#define X(B, Y, T, W, N) \
    if (DN_SAME(B, Y, T, W)) return launch_vel_##N(a, c, q, fdoh, sw, ntl, stream);
    DN_TILES(X)
#undef X
}

void launch_str(const KArgs &a, const StepCtl &c0, const DevAcq &q, const Tile &t, int fdoh, int ntl, void *stream)
{
    const StepCtl c = with_grid(c0, t, a);
#define X(B, Y, T, W, N) \
    if (DN_SAME(B, Y, T, W)) return launch_str_##N(a, c, q, fdoh, ntl, stream);
    DN_TILES(X)
#undef X
}

void launch_msource(const KArgs &a, const DevAcq &q, int nt, void *stream)
{
    hipLaunchKernelGGL(k_msource, dim3(1), dim3(64), 0, (hipStream_t)stream, a, q, nt);
}

void launch_store(const KArgs &a, float *const fwd[5], long k, void *stream)
{
    const long n = (long)a.nx * a.ny;
    hipLaunchKernelGGL(k_store, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a, fwd[0],
                       fwd[1], fwd[2], fwd[3], fwd[4], k);
}

void launch_energy(const KArgs &a, float *W, void *stream)
{
    const long n = (long)a.nx * a.ny;
    hipLaunchKernelGGL(k_energy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, a, W);
}

void launch_assemble(const KArgs &a, const float *wc, const float *wcu, const float *wcr, float *gvp,
                     float *gvs, float *grho, int invmat1, int physics, void *stream)
{
    const long n = (long)a.nx * a.ny;
    dim3 g((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_assemble, g, dim3(256), 0, (hipStream_t)stream, a, wc, wcu, wcr, gvp, gvs, grho,
                       invmat1, physics);
}

void launch_residual(const DevAcq &q, int comp, const float *obs, float dt, float w, void *stream)
{
    if (q.ntr <= 0) return;
    dim3 g((q.ntr + RB - 1) / RB);
    hipLaunchKernelGGL(k_residual, g, dim3(RB), 0, (hipStream_t)stream, q, comp, obs, dt, w);
}

void launch_l2(const DevAcq &q, int comp, void *stream)
{
    hipLaunchKernelGGL(k_l2_part, dim3(L2_PARTS), dim3(256), 0, (hipStream_t)stream, q, comp);
    hipLaunchKernelGGL(k_l2_sum, dim3(1), dim3(256), 0, (hipStream_t)stream, q);
}

int l2_scratch_doubles() { return 1 + L2_PARTS; }

}  // namespace dn
