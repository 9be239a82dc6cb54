// dn_calc_residuals_noalloc: the residual / misfit entry point for callers
// that cannot allocate device memory.  Its device scratch -- the observed
// traces staged for the residual kernel and the misfit with its partial
// sums -- lives in one fixed file-scope device array instead of buffers of
// the context; the call makes no hipMalloc / hipFree and no library call
// that allocates.  Every call on a GPU shares that GPU's copy of the array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "dn_ctx.h"

using namespace dn;

namespace {

// staged samples of one component (ntr * NT) and doubles for the misfit and
// its partial sums
constexpr size_t kObsMax = DN_NOALLOC_MAX_SAMPLES;
constexpr size_t kL2Doubles = 512;

// [0, kL2Doubles): misfit and partial sums; then kObsMax floats
static __device__ __align__(256) double scratch[kL2Doubles + kObsMax / 2];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};  // per device: the address of its copy of scratch

__global__ void k_fill_f(float *p, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) p[k] = 0.0f;
}

__global__ void k_zero_misfit(double *p)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) p[0] = 0.0;
}

// the device address of scratch on the current device (looked up once per
// device: each GPU holds its own copy of the array at its own address)
int scratch_address(int device, void **out)
{
    if (device < 0 || device >= kMaxDevices) return fail(DN_ERR_ARG, "device ordinal beyond the static scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(DN_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return DN_OK;
}

void zero_floats(float *p, size_t n, hipStream_t s)
{
    if (n == 0) return;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_f, dim3(blocks), dim3(256), 0, s, p, n);
}

}  // namespace

extern "C" int dn_calc_residuals_noalloc(dn_ctx *c, const float *obs_vx, const float *obs_vy, const float *obs_p,
                                         int testshot, double *l2)
{
    if (!c) return fail(DN_ERR_ARG, "null context");
    DevAcq q = c->acq;
    const size_t n = (size_t)q.ntr * c->p.nt;
    if (n > kObsMax)
        return fail(DN_ERR_ARG, "dn_calc_residuals_noalloc: ntr * NT = " + std::to_string(n) +
                                    " samples exceed its static scratch of " + std::to_string(kObsMax));
    if ((size_t)l2_scratch_doubles() > kL2Doubles) return fail(DN_ERR_STATE, "misfit partial sums exceed the static scratch");
    HIP_OK(hipSetDevice(c->device));
    void *base = nullptr;
    int rc = scratch_address(c->device, &base);
    if (rc) return rc;
    double *d_l2 = static_cast<double *>(base);
    float *d_obs = reinterpret_cast<float *>(d_l2 + kL2Doubles);
    q.l2 = d_l2;  // the misfit kernels sum into the static scratch
    const int QB = c->p.quelltypb;
    hipLaunchKernelGGL(k_zero_misfit, dim3(1), dim3(64), 0, c->stream, d_l2);
    for (int k = 0; k < 3; ++k)
        if (q.adj[k]) zero_floats(q.adj[k], n, c->stream);
    const float *obs[3] = {obs_vx, obs_vy, obs_p};
    bool use[3];
    float w[3];
    residual_components(c->p.physics, QB, c->res_w, use, w);
    for (int comp = 0; comp < 3; ++comp) {
        if (!use[comp] || !obs[comp] || q.ntr == 0) continue;
        HIP_OK(hipMemcpyAsync(d_obs, obs[comp], n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        launch_residual(q, comp, d_obs, c->p.dt, w[comp], c->stream);
        if (testshot) launch_l2(q, comp, c->stream);
        // the staging area is reused by the next component
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    HIP_OK(hipGetLastError());
    double v = 0.0;
    HIP_OK(d2h(c->stream, &v, d_l2, sizeof(double)));
    if (l2) *l2 = v;
    return DN_OK;
}
