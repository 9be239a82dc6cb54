// Vector-column sweep kernels, tile shape (tbx 64, by 2, ty 4, vw 4).
#define DN_TBX 64
#define DN_BY 2
#define DN_TY 4
#define DN_VW 4
#define DN_CFG_NAME(p) p##t64x2x4v4
#include "dn_sweeps_vec.inc"
