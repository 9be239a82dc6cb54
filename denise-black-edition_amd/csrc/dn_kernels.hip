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
    float *diff = q.adj[comp] + (long)t * ns;
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

}  // namespace

// Tile shapes compiled into the library: X(TBX, BY, TY, name), each in its
// own dn_sweep_<name>.hip translation unit.
// Compiled sweep shapes (tbx, by, ty, vw): one translation unit each
// (dn_sweep_t*.hip scalar columns, dn_sweepv_t*.hip vector columns).
#define DN_TILES(X)                                                                              \
    X(64, 4, 1, 1, t64x4x1) X(128, 2, 1, 1, t128x2x1) X(256, 1, 1, 1, t256x1x1)                   \
    X(64, 4, 2, 1, t64x4x2) X(128, 2, 4, 1, t128x2x4) X(64, 4, 1, 4, t64x4x1v4)                   \
    X(64, 4, 2, 4, t64x4x2v4) X(64, 2, 4, 4, t64x2x4v4) X(64, 4, 1, 2, t64x4x1v2)                 \
    X(128, 2, 2, 2, t128x2x2v2)

#define X(B, Y, T, W, N)                                                                    \
    void launch_vel_##N(const KArgs &, const StepCtl &, const DevAcq &, int, int, void *); \
    void launch_str_##N(const KArgs &, const StepCtl &, const DevAcq &, int, void *);
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

void launch_vel(const KArgs &a, const StepCtl &c0, const DevAcq &q, const Tile &t, int fdoh, int sw, void *stream)
{
    const StepCtl c = with_grid(c0, t, a);
#define X(B, Y, T, W, N) \
    if (DN_SAME(B, Y, T, W)) return launch_vel_##N(a, c, q, fdoh, sw, stream);
    DN_TILES(X)
#undef X
}

void launch_str(const KArgs &a, const StepCtl &c0, const DevAcq &q, const Tile &t, int fdoh, void *stream)
{
    const StepCtl c = with_grid(c0, t, a);
#define X(B, Y, T, W, N) \
    if (DN_SAME(B, Y, T, W)) return launch_str_##N(a, c, q, fdoh, stream);
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
