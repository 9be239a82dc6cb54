// Multi-shot velocity sweep kernels and the clock kernel.
#define DN_BATCH_VEL 1
#include "dn_batch.inc"

namespace dn {
namespace {

// velocity-sweep shapes compiled for multi-shot launches (tbx, by, ty, vw)
#define DN_BATCH_VTILES(X) X(128, 2, 4, 1) X(64, 4, 2, 1)

template <int TBX, int BY, int TY, int VW>
void bvel(const BatchShot *shots, const BatchBlock *blk, int nb, int nfull, const int *clock, int k, int fdoh,
          int sw, hipStream_t s)
{
#define DN_BV(F)                                                                                          \
    case F:                                                                                               \
        if (sw == 0)                                                                                      \
            hipLaunchKernelGGL((kb_vel<F, TY, TBX, BY, 0, VW>), dim3(nb), dim3(TBX, BY), 0, s, shots, blk, nb, nfull, clock, k); \
        else                                                                                              \
            hipLaunchKernelGGL((kb_vel<F, TY, TBX, BY, 1, VW>), dim3(nb), dim3(TBX, BY), 0, s, shots, blk, nb, nfull, clock, k); \
        return;
    switch (fdoh) {
        DN_BV(1) DN_BV(2) DN_BV(3) DN_BV(4) DN_BV(5) DN_BV(6)
    }
#undef DN_BV
}

}  // namespace

bool batch_tile_supported(const Tile &t, int sweep)
{
    if (sweep == 0) {
#define X(B, Y, T, W) \
    if (t.tbx == B && t.by == Y && t.ty == T && t.vw == W) return true;
        DN_BATCH_VTILES(X)
#undef X
        return false;
    }
    // stress-sweep shapes compiled in dn_batch_str.hip (the group's L decides)
    return (t.tbx == 64 && t.by == 4 && t.ty == 1 && (t.vw == 1 || t.vw == 4)) ||
           (t.tbx == 64 && t.by == 4 && t.ty == 2 && (t.vw == 1 || t.vw == 2));
}

void launch_batch_vel(const BatchShot *shots, const BatchBlock *blk, int nblocks, int nfull, const int *clock,
                      int k, const Tile &t, int fdoh, int sw, void *stream)
{
    if (nblocks <= 0) return;
    hipStream_t s = (hipStream_t)stream;
#define X(B, Y, T, W) \
    if (t.tbx == B && t.by == Y && t.ty == T && t.vw == W) return bvel<B, Y, T, W>(shots, blk, nblocks, nfull, clock, k, fdoh, sw, s);
    DN_BATCH_VTILES(X)
#undef X
}

void launch_batch_clock(int *clock, int add, void *stream)
{
    hipLaunchKernelGGL(kb_clock, dim3(1), dim3(64), 0, (hipStream_t)stream, clock, add);
}

}  // namespace dn
