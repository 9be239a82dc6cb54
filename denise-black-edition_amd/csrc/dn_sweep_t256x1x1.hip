// Fused sweeps instantiated for 256x1 threads, 1 row(s) per thread.
#define DN_TBX 256
#define DN_BY 1
#define DN_TY 1
#define DN_CFG_NAME(p) p##t256x1x1
#include "dn_sweeps.inc"
