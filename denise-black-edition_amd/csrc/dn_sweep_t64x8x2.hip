// Fused sweeps instantiated for 64x8 threads, 2 row(s) per thread.
#define DN_TBX 64
#define DN_BY 8
#define DN_TY 2
#define DN_CFG_NAME(p) p##t64x8x2
#include "dn_sweeps.inc"
