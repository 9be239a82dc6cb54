// C ABI of the device-memory arena (the bookkeeping is dn_arena_pool.cpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>

#include "dn_arena.h"
#include "dn_ctx.h"

namespace dn {
// device buffers a context with these parameters allocates (plan order)
std::vector<size_t> plan_sizes(const dn_params &p, int nsrc, int ntr);
}  // namespace dn

int dn_arena_create(int device, size_t bytes, dn_arena **out)
{
    if (!out) return fail(DN_ERR_ARG, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(DN_ERR_ARG, "no such HIP device");
    if (bytes < dn_arena::ALIGN) return fail(DN_ERR_ARG, "arena too small");
    bytes = dn_arena::round(bytes);
    dn_arena *a = new (std::nothrow) dn_arena();
    if (!a) return fail(DN_ERR_OOM, "host allocation failed");
    a->device = device;
    if (hipSetDevice(device) != hipSuccess || hipMalloc((void **)&a->base, bytes) != hipSuccess) {
        delete a;
        return fail(DN_ERR_OOM, "hipMalloc of the arena (" + std::to_string(bytes) + " bytes) failed");
    }
    a->init(bytes);
    *out = a;
    return DN_OK;
}

int dn_arena_destroy(dn_arena *a)
{
    if (!a) return DN_OK;
    {
        std::lock_guard<std::mutex> lk(a->mu);
        if (!a->used.empty()) return fail(DN_ERR_STATE, "contexts still hold arena memory");
    }
    (void)hipSetDevice(a->device);
    (void)hipFree(a->base);
    delete a;
    return DN_OK;
}

int dn_arena_info(dn_arena *a, size_t *free_bytes, size_t *largest_free, size_t *total)
{
    if (!a) return fail(DN_ERR_ARG, "null arena");
    if (largest_free) *largest_free = a->largest();
    std::lock_guard<std::mutex> lk(a->mu);
    if (free_bytes) *free_bytes = a->free_bytes;
    if (total) *total = a->size;
    return DN_OK;
}

int dn_arena_fits(dn_arena *a, const dn_params *p, int nsrc, int ntr, int *fits)
{
    if (!a || !p || !fits || nsrc < 0 || ntr < 0) return fail(DN_ERR_ARG, "bad fit request");
    std::vector<size_t> sizes = plan_sizes(*p, nsrc, ntr);
    // the cell lists of the four sweeps (a few bytes per point and tile)
    for (int k = 0; k < 8; ++k) sizes.push_back(64 * 1024);
    std::sort(sizes.rbegin(), sizes.rend());
    *fits = a->fits(sizes) ? 1 : 0;
    return DN_OK;
}
