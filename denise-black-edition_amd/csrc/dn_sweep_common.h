// Helpers shared by the fused sweep kernels (dn_sweeps.inc,
// dn_sweeps_vec.inc) and the multi-shot kernels (dn_batch.inc).
#pragma once
#include <hip/hip_runtime.h>

#include "dn_internal.h"

namespace dn {
namespace {

constexpr int NXCD = 8;

__device__ __forceinline__ float cpml(float &psi, float b, float a, float K, float d)
{
    psi = b * psi + a * d;
    return d / K + psi;
}

// Block -> tile: workgroups are dispatched to the XCDs round-robin, so the
// blocks of XCD x are x, x+8, x+16, ...; give them the contiguous tile range
// [start_x, start_x + count_x).
__device__ __forceinline__ int xcd_tile(int bid, int ntiles)
{
    const int per = ntiles / NXCD, rem = ntiles % NXCD;
    const int x = bid % NXCD, k = bid / NXCD;
    return x * per + (x < rem ? x : rem) + k;
}

// time derivative of the source signal by central differences, one-sided at
// the ends (src/psource.c:28-31)
__device__ __forceinline__ float src_amp(const float *s, int nt, int NT, float dt)
{
    float amp = 0.0f;
    if (nt == 1) amp = s[1] / dt;
    if (nt > 1 && nt < NT) amp = (s[nt] - s[nt - 2]) / dt;
    if (nt == NT) amp = -s[nt - 2] / dt;
    return amp;
}

template <int VW>
struct VecT;
template <>
struct VecT<2> {
    typedef float __attribute__((ext_vector_type(2))) type;
};
template <>
struct VecT<4> {
    typedef float __attribute__((ext_vector_type(4))) type;
};

// stores of the sweep outputs: plain, or streaming (nontemporal: written
// through without keeping the lines in the caches) when nt is set
template <bool NT>
__device__ __forceinline__ void st1(float *p, float v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

template <int VW>
__device__ __forceinline__ typename VecT<VW>::type ldv(const float *p)
{
    return *reinterpret_cast<const typename VecT<VW>::type *>(p);
}

// loads of planes read once per step (material and coefficient planes):
// streaming (nontemporal) when NT is set, so that they do not displace the
// wavefields the next sweep reads from the caches.  NT is a template
// parameter: with a run-time flag the compiler merges the two loads of the
// same address and drops the nontemporal marking.
template <bool NT>
__device__ __forceinline__ float ld1(const float *p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <int VW, bool NT>
__device__ __forceinline__ typename VecT<VW>::type ldvn(const float *p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(reinterpret_cast<const typename VecT<VW>::type *>(p));
    else
        return ldv<VW>(p);
}

// store the first n (<= VW) elements (streaming when NT is set)
template <int VW, bool NT = false>
__device__ __forceinline__ void stv(float *p, typename VecT<VW>::type v, int n)
{
    if (n >= VW) {
        if constexpr (NT)
            __builtin_nontemporal_store(v, reinterpret_cast<typename VecT<VW>::type *>(p));
        else
            *reinterpret_cast<typename VecT<VW>::type *>(p) = v;
    } else {
#pragma unroll
        for (int e = 0; e < VW; ++e)
            if (e < n) p[e] = v[e];
    }
}

}  // namespace
}  // namespace dn
