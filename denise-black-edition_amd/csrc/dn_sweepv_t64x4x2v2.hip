// Vector-column sweep kernels, tile shape (tbx 64, by 4, ty 2, vw 2).
#define DN_TBX 64
#define DN_BY 4
#define DN_TY 2
#define DN_VW 2
#define DN_CFG_NAME(p) p##t64x4x2v2
#include "dn_sweeps_vec.inc"
