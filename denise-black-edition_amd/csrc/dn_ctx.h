// The per-shot context of libdnhip.so (the object behind dn_ctx*), shared
// by the C-ABI layer (dn_api.hip) and the multi-shot engine (dn_engine.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "dn_arena.h"
#include "dn_internal.h"

namespace dn {
// records the message for dn_last_error() and returns code
int fail(int code, const std::string &msg);

#define HIP_OK(expr)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(DN_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

enum { F_VX = 0, F_VY, F_SXX, F_SYY, F_SXY, F_VXP1, F_VYP1, F_UX, F_UY, F_UXY, F_N };

}  // namespace dn

// dn_ctx is the C ABI's opaque type (global namespace); its members use the
// library's internal types.
using namespace dn;

namespace dn {
// host<->device copies ordered on a stream, returning when complete
inline hipError_t h2d(hipStream_t s, void *dst, const void *src, size_t n)
{
    const hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
}
inline hipError_t d2h(hipStream_t s, void *dst, const void *src, size_t n)
{
    const hipError_t e = hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
}
}  // namespace dn

struct dn_ctx {
    dn_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    Layout lay;
    HostSetup setup;
    int fdoh = 0, nd = 0, ntdtinv = 0;
    bool model_set = false, fwd_complete = false;
    std::vector<std::pair<void *, size_t>> allocs;
    size_t bytes = 0;
    float *fld[F_N] = {};  // plane bases
    float *rip = nullptr, *rjp = nullptr, *muipjp = nullptr, *f = nullptr, *g = nullptr;
    float *vp = nullptr, *vs = nullptr, *rho = nullptr;
    float *rmem = nullptr, *pmem = nullptr, *qmem = nullptr;
    float *vmu = nullptr, *vts = nullptr, *vpi = nullptr, *vtp = nullptr, *vmuij = nullptr, *vtsij = nullptr;
    float *fs_fjm = nullptr, *fs_g = nullptr, *fs_dme = nullptr;
    float *ani[8] = {};  // anisotropic stiffness planes (physics 3 / 4)
    float *psi[9] = {};
    float *pml[P_NVEC] = {};
    float *fwd[5] = {};
    float *wc = nullptr, *wcu = nullptr, *wcr = nullptr, *gvp = nullptr, *gvs = nullptr, *grho = nullptr;
    HostMaterial mat;
    std::vector<float> h_vp, h_vs, h_rho;  // x-major model kept for receiver factors
    DevAcq acq;
    std::vector<int> rec_ix, rec_iy, src_ix, src_iy, src_type;
    int qtype = 0, has_force = 0;
    float res_w[2] = {1.0f, 1.0f};  // COMP_WEIGHT factors of the vx / vy residuals
    int tau_const = 0;
    float ts0 = 0.0f, tp0 = 0.0f, tsij0 = 0.0f;
    KArgs a{};
    // tile shape of the sweeps and the per-tile cell work lists
    // (0 forward velocity, 1 forward stress, 2 adjoint velocity, 3 adjoint stress)
    Tile tile_v, tile_s;  // velocity / stress sweep shapes
    int nt_v = 0, nt_s = 0;  // streaming access to the once-per-step planes of each sweep
    int ntx_v = 0, ntiles_v = 0, ntx_s = 0, ntiles_s = 0;
    int *lst_off[4] = {};   // per list: CSR offsets (ntiles+1), then the lean and the full tile indices
    Ent *lst_ent[4] = {};
    int lst_nlean[4] = {}, lst_nfull[4] = {};
    std::vector<int> lean_h[4], full_h[4];  // host copies of the tile sets
    // timing
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double run_ms = 0.0, v_ms = 0.0, s_ms = 0.0;
    int sample_every = 0, samples = 0;
    // energy preconditioning (dn_set_energy_weights): sum of vx^2+vy^2 of the
    // forward field at its store steps (0) and of the adjoint field at its
    // correlation steps (1), NX x NY in the gradient accumulators' layout
    int eprecond = 0;
    float *ew[2] = {nullptr, nullptr};
    // wavefield snapshots of dn_run_forward (dn_set_snapshots)
    dn_snap_fn snap_fn = nullptr;
    void *snap_user = nullptr;
    int snap_first = 0, snap_inc = 0, snap_last = 0;
    std::vector<hipEvent_t> pool;

    dn_arena *arena = nullptr;  // buffers come from this arena, else hipMalloc
    int dalloc(void **ptr, size_t nbytes)
    {
        if (nbytes == 0) nbytes = 4;
        hipError_t e = hipSuccess;
        if (arena) {
            *ptr = arena->alloc(nbytes);
            if (!*ptr)
                return fail(DN_ERR_OOM, "the arena has no free block of " + std::to_string(nbytes) + " bytes");
        } else if ((e = hipMalloc(ptr, nbytes)) != hipSuccess) {
            *ptr = nullptr;
            return fail(DN_ERR_OOM, "hipMalloc of " + std::to_string(nbytes) + " bytes failed: " +
                                        hipGetErrorString(e));
        }
        allocs.emplace_back(*ptr, nbytes);
        bytes += nbytes;
        // zeroed on the context's own stream: the library makes no
        // null-stream or device-wide calls, so contexts can be prepared on
        // one thread while another thread captures or runs an engine graph
        e = hipMemsetAsync(*ptr, 0, nbytes, stream);
        if (e != hipSuccess) return fail(DN_ERR_HIP, std::string("hipMemsetAsync failed: ") + hipGetErrorString(e));
        return DN_OK;
    }
    void release(void *ptr)
    {
        if (arena)
            (void)arena->release(ptr);
        else
            (void)hipFree(ptr);
    }
    void dfree(void *ptr)
    {
        for (auto it = allocs.begin(); it != allocs.end(); ++it)
            if (it->first == ptr) {
                release(ptr);
                bytes -= it->second;
                allocs.erase(it);
                return;
            }
    }
    ~dn_ctx()
    {
        if (device >= 0) (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &pr : allocs) release(pr.first);
        for (auto ev : pool) (void)hipEventDestroy(ev);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};


namespace dn {
// zero the wavefields, memory variables and CPML state of a shot
int zero_state(dn_ctx *c);
// sweep control with the cell lists lv (velocity) and ls (stress)
StepCtl base_ctl(dn_ctx *c, int lv, int ls);
}  // namespace dn
