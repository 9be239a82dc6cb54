// Vector-column sweep kernels, tile shape (tbx 64, by 4, ty 1, vw 4).
#define DN_TBX 64
#define DN_BY 4
#define DN_TY 1
#define DN_VW 4
#define DN_CFG_NAME(p) p##t64x4x1v4
#include "dn_sweeps_vec.inc"
