// Vector-column sweep kernels, tile shape (tbx 128, by 2, ty 2, vw 2).
#define DN_TBX 128
#define DN_BY 2
#define DN_TY 2
#define DN_VW 2
#define DN_CFG_NAME(p) p##t128x2x2v2
#include "dn_sweeps_vec.inc"
