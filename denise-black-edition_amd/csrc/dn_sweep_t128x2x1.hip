// Fused sweeps instantiated for 128x2 threads, 1 row(s) per thread.
#define DN_TBX 128
#define DN_BY 2
#define DN_TY 1
#define DN_CFG_NAME(p) p##t128x2x1
#include "dn_sweeps.inc"
