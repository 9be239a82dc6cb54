"""CPU restatement of DENISE's wavefield snapshot writer, src/snap_ssg.c:10-184
with src/writedsk.c:15-43, in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
dnio_write_snap (libdnio).  The product never uses it.

snap_values returns, per output file suffix, the float32 values in the
order the reference writes them (i outer, j inner, every IDX-th column and
IDY-th row).  Fields are padded NR arrays (rows -nd+1..NY+nd, columns
-nd+1..NX+nd, x fastest); ppi, pu are (NX, NY) model arrays.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def snap_values(snap, fdorder, dh, vx, vy, sxx, syy, ppi, pu, hc, idx=1, idy=1):
    ny_p, nx_p = vx.shape
    nd = fdorder // 2 + 1
    fdoh = fdorder // 2
    NX, NY = nx_p - 2 * nd, ny_p - 2 * nd

    def at(a, j, i):
        return a[j + nd - 1, i + nd - 1]

    dhi = F32(1.0 / float(F32(dh)))
    pts = [(i, j) for i in range(1, NX + 1, idx) for j in range(1, NY + 1, idy)]
    out = {}
    if snap in (1, 4):
        out["x"] = np.array([at(vx, j, i) for i, j in pts], F32)
        out["y"] = np.array([at(vy, j, i) for i, j in pts], F32)
    if snap == 2:
        out["p"] = np.array([F32(-at(sxx, j, i) - at(syy, j, i)) for i, j in pts], F32)
    if snap in (3, 4):
        rot, div = [], []
        for i, j in pts:
            vyx = F32(0.0)
            vxy = F32(0.0)
            for m in range(1, fdoh + 1):
                vyx = F32(vyx + F32(hc[m] * F32(at(vy, j, i + m) - at(vy, j, i - m + 1))))
                vxy = F32(vxy + F32(hc[m] * F32(at(vx, j + m, i) - at(vx, j - m + 1, i))))
            vyx = F32(vyx * dhi)
            vxy = F32(vxy * dhi)
            rot.append(F32(float(F32(vxy - vyx)) * np.sqrt(float(pu[i - 1, j - 1]))))
        for i, j in pts:
            vxx = F32(0.0)
            vyy = F32(0.0)
            for m in range(1, fdoh + 1):
                vxx = F32(vxx + F32(hc[m] * F32(at(vx, j, i + m - 1) - at(vx, j, i - m))))
                vyy = F32(vyy + F32(hc[m] * F32(at(vy, j + m - 1, i) - at(vy, j - m, i))))
            vxx = F32(vxx * dhi)
            vyy = F32(vyy * dhi)
            div.append(F32(float(F32(vxx + vyy)) * np.sqrt(float(ppi[i - 1, j - 1]))))
        out["rot"] = np.array(rot, F32)
        out["div"] = np.array(div, F32)
    return out
