"""CPU restatement of the reference's time-domain Butterworth filtering
(src/apply_tdfilt.c, src/timedomain_filt.c, and the vendored libcseife:
seife_butterworth cseife.c:139-178, seife_rfk cseife_rfk.c:35-147,
seife_rekfl cseife_rekfl.c:30-78) in plain Python floats (IEEE double,
math.sin / math.atan).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of dnio_tdfilt
(libdnio).  The product never uses it.

The pins of the filter against what the mathematics fixes (the magnitude
response of the bilinear-transformed analog Butterworth filter) are in
tests/test_tdfilt_host.py.
"""
from __future__ import annotations

import math

import numpy as np


def rfk(kind, t0, h):
    """Section coefficients (f0, f1, f2, g1, g2); kind: lp1, hp1, lp2, hp2;
    t0: corner period in samples; h: damping of a 2nd-order section."""
    eps = 8.0 * math.atan(1.0) / t0
    if kind in ("lp1", "hp1"):
        g1 = (2.0 - eps) / (2.0 + eps)
        if kind == "lp1":
            f0 = eps / (2.0 + eps)
            return f0, f0, 0.0, g1, 0.0
        f0 = 2.0 / (2.0 + eps)
        return f0, -f0, 0.0, g1, 0.0
    epsq = eps * eps
    a = 1.0 - eps * h + epsq / 4.0
    b = -2.0 + epsq / 2.0
    c = 1.0 + eps * h + epsq / 4.0
    g1, g2 = -b / c, -a / c
    if kind == "lp2":
        f0 = epsq / 4.0 / c
        return f0, f0 + f0, f0, g1, g2
    f0 = 1.0 / c
    return f0, -f0 - f0, f0, g1, g2


def rekfl(x, coef):
    f0, f1, f2, g1, g2 = coef
    xa = xaa = ya = yaa = 0.0
    y = list(x)
    for j, xn in enumerate(x):
        y[j] = f0 * xn + f1 * xa + f2 * xaa + g1 * ya + g2 * yaa
        xaa, xa, yaa, ya = xa, xn, ya, y[j]
    return y


def butterworth(x, dt, t0, m, highpass):
    """seife_lpb / seife_hpb on a list of floats (double)."""
    x = [float(v) for v in x]
    if highpass:
        x0 = x[0]
        x = [v - x0 for v in x]
    mm = m // 2
    if m > 2 * mm:
        x = rekfl(x, rfk("hp1" if highpass else "lp1", t0 / dt, 0.0))
    pih = 2.0 * math.atan(1.0)
    for j in range(1, mm + 1):
        x = rekfl(x, rfk("hp2" if highpass else "lp2", t0 / dt, math.sin(pih * (2 * j - 1) / m)))
    return x


def apply_tdfilt(section, dt, order, fc2, fc1):
    """section: (ntr, ns) float32; returns the filtered copy."""
    out = np.array(section, np.float32, copy=True)
    dt = float(np.float32(dt))
    for fc, hp in ((fc2, False), (fc1, True)):
        if hp and not fc1 > 0.0:
            continue
        T0 = 1.0 / float(np.float32(fc))
        for i in range(out.shape[0]):
            out[i] = np.array(butterworth(out[i].astype(np.float64), dt, T0, order, hp), np.float32)
    return out
