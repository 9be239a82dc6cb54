// Multi-shot stress sweep kernels.
#include "dn_batch.inc"

namespace dn {
namespace {

template <int TBX, int BY, int TY, int VW, int F>
void bstr_f(const BatchShot *shots, const BatchBlock *blk, int nb, int nfull, const int *clock, int k, int L,
            hipStream_t s)
{
    switch (L) {
    case 0: hipLaunchKernelGGL((kb_str<F, TY, TBX, BY, 0, VW>), dim3(nb), dim3(TBX, BY), 0, s, shots, blk, nb, nfull, clock, k); break;
    case 1: hipLaunchKernelGGL((kb_str<F, TY, TBX, BY, 1, VW>), dim3(nb), dim3(TBX, BY), 0, s, shots, blk, nb, nfull, clock, k); break;
    case 2: hipLaunchKernelGGL((kb_str<F, TY, TBX, BY, 2, VW>), dim3(nb), dim3(TBX, BY), 0, s, shots, blk, nb, nfull, clock, k); break;
    case 3: hipLaunchKernelGGL((kb_str<F, TY, TBX, BY, 3, VW>), dim3(nb), dim3(TBX, BY), 0, s, shots, blk, nb, nfull, clock, k); break;
    default: hipLaunchKernelGGL((kb_str<F, TY, TBX, BY, -1, VW>), dim3(nb), dim3(TBX, BY), 0, s, shots, blk, nb, nfull, clock, k); break;
    }
}

template <int TBX, int BY, int TY, int VW>
void bstr(const BatchShot *shots, const BatchBlock *blk, int nb, int nfull, const int *clock, int k, int fdoh, int L,
          hipStream_t s)
{
    switch (fdoh) {
    case 1: return bstr_f<TBX, BY, TY, VW, 1>(shots, blk, nb, nfull, clock, k, L, s);
    case 2: return bstr_f<TBX, BY, TY, VW, 2>(shots, blk, nb, nfull, clock, k, L, s);
    case 3: return bstr_f<TBX, BY, TY, VW, 3>(shots, blk, nb, nfull, clock, k, L, s);
    case 4: return bstr_f<TBX, BY, TY, VW, 4>(shots, blk, nb, nfull, clock, k, L, s);
    case 5: return bstr_f<TBX, BY, TY, VW, 5>(shots, blk, nb, nfull, clock, k, L, s);
    default: return bstr_f<TBX, BY, TY, VW, 6>(shots, blk, nb, nfull, clock, k, L, s);
    }
}

// elastic (L = 0) groups only
template <int TBX, int BY, int TY, int VW>
void bstr_el(const BatchShot *shots, const BatchBlock *blk, int nb, int nfull, const int *clock, int k, int fdoh,
             hipStream_t s)
{
#define DN_BS(F)                                                                                          \
    case F:                                                                                               \
        hipLaunchKernelGGL((kb_str<F, TY, TBX, BY, 0, VW>), dim3(nb), dim3(TBX, BY), 0, s, shots, blk, nb, nfull, \
                           clock, k);                                                                     \
        return;
    switch (fdoh) {
        DN_BS(1) DN_BS(2) DN_BS(3) DN_BS(4) DN_BS(5) DN_BS(6)
    }
#undef DN_BS
}

}  // namespace

bool batch_str_supported(const Tile &t, int L)
{
    if (t.tbx == 64 && t.by == 4 && t.ty == 1 && (t.vw == 1 || t.vw == 4)) return true;
    return L == 0 && t.tbx == 64 && t.by == 4 && t.ty == 2 && (t.vw == 1 || t.vw == 2);
}

void launch_batch_str(const BatchShot *shots, const BatchBlock *blk, int nblocks, int nfull, const int *clock,
                      int k, const Tile &t, int fdoh, int L, void *stream)
{
    if (nblocks <= 0) return;
    hipStream_t s = (hipStream_t)stream;
    if (t.tbx == 64 && t.by == 4 && t.ty == 1 && t.vw == 4) return bstr<64, 4, 1, 4>(shots, blk, nblocks, nfull, clock, k, fdoh, L, s);
    if (t.tbx == 64 && t.by == 4 && t.ty == 1 && t.vw == 1) return bstr<64, 4, 1, 1>(shots, blk, nblocks, nfull, clock, k, fdoh, L, s);
    if (L == 0 && t.tbx == 64 && t.by == 4 && t.ty == 2 && t.vw == 1) return bstr_el<64, 4, 2, 1>(shots, blk, nblocks, nfull, clock, k, fdoh, s);
    if (L == 0 && t.tbx == 64 && t.by == 4 && t.ty == 2 && t.vw == 2) return bstr_el<64, 4, 2, 2>(shots, blk, nblocks, nfull, clock, k, fdoh, s);
}

}  // namespace dn
