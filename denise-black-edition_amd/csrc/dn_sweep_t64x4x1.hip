// Fused sweeps instantiated for 64x4 threads, 1 row(s) per thread.
#define DN_TBX 64
#define DN_BY 4
#define DN_TY 1
#define DN_CFG_NAME(p) p##t64x4x1
#include "dn_sweeps.inc"
