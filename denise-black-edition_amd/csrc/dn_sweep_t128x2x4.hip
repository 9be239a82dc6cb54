// Fused sweeps instantiated for 128x2 threads, 4 row(s) per thread.
#define DN_TBX 128
#define DN_BY 2
#define DN_TY 4
#define DN_CFG_NAME(p) p##t128x2x4
#include "dn_sweeps.inc"
