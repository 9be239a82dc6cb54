"""CPU restatement of the reference's source-time-function inversion
(src/stf.c:13-262, the Wiener deconvolution of the observed vy traces by
the synthetic ones of a spike source) and of its trace normalisations
(src/util.c:489-579), in numpy: transforms by numpy.fft in double
precision, the float32 quantities of the reference (sections, Ebar, the
water level, the output wavelet) rounded where it rounds them.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of dnfwi_stf
(libdnio).  The product never uses it.

The pins against what the mathematics fixes (recovery of a known filter,
the water level of identical traces in closed form, the normalisations)
are in tests/test_stf_host.py.
"""
from __future__ import annotations

import math

import numpy as np

F32 = np.float32


def npad_of(ns):
    """stf.c:62: 2^(ceil(log2 ns) + 2)."""
    return int(math.pow(2.0, math.ceil(math.log(float(ns)) / math.log(2.0)) + 2.0))


def normalize(data):
    """util.c:489-526: each trace divided by its largest |sample| over
    samples 2..ns (1 when 0); sample 1 unchanged.  Returns a new array."""
    d = np.array(data, F32, copy=True)
    for t in d:
        m = F32(max(float(t[1:].max(initial=0.0)), abs(float(t[1:].min(initial=0.0)))))
        if m == 0:
            m = F32(1.0)
        t[1:] = (t[1:] / m).astype(F32)
    return d


def normalize_rel(syn, obs):
    """util.c:528-579: the observed traces scaled by (max|syn| / max|obs|)
    over samples 2..ns.  Returns the new observed array."""
    o = np.array(obs, F32, copy=True)
    for s, t in zip(np.asarray(syn, F32), o):
        m = F32(max(float(s[1:].max(initial=0.0)), abs(float(s[1:].min(initial=0.0)))))
        mo = F32(max(float(t[1:].max(initial=0.0)), abs(float(t[1:].min(initial=0.0)))))
        m = m if m != 0 else F32(1.0)
        mo = mo if mo != 0 else F32(1.0)
        t[1:] = (t[1:] * F32(m / mo)).astype(F32)
    return o


def _plus(x, n):
    """Transform of sign +1 (FFTW_BACKWARD, unnormalised): sum_m x_m e^{+2 pi i k m / n}."""
    return np.fft.ifft(x, n) * n


def _minus(x, n):
    """Transform of sign -1 (FFTW_FORWARD, unnormalised)."""
    return np.fft.fft(x, n)


def stf(obs, syn, mute, norm, eps_stf, spike):
    """The wavelet (float32, ns samples) of stf.c for one shot; obs, syn:
    (ntr, ns) vy sections; mute: bool per trace or None; norm: NORMALIZE;
    spike: the source signal of the synthetic run."""
    obs = np.array(obs, F32, copy=True)
    syn = np.array(syn, F32, copy=True)
    ntr, ns = obs.shape
    n = npad_of(ns)
    if mute is not None:
        obs[np.asarray(mute, bool)] = 0
        syn[np.asarray(mute, bool)] = 0
    if norm == 1:
        syn, obs = normalize(syn), normalize(obs)
    if norm == 2:
        obs = normalize_rel(syn, obs)
    dss = _plus(np.asarray(spike, F32).astype(np.float64), n)
    D = _plus(obs.astype(np.float64), n) if ntr else None
    M = _plus(syn.astype(np.float64), n)
    sumn = np.zeros(n, complex)
    sumd = np.zeros(n, complex)
    ebar = F32(0.0)
    for i in range(ntr):
        sumn = sumn + (ntr * D[i]) * np.conj(M[i])
        sumd = sumd + (ntr * M[i]) * np.conj(M[i])
        # stf.c:235: float Ebar += the running denominator, sample by sample
        for v in sumd.real:
            ebar = F32(float(ebar) + float(v))
    ebar = F32(ebar / F32(n * ntr))
    water = F32(F32(F32(eps_stf) * F32(ntr)) * ebar)
    ds = dss * (sumn / (sumd + float(water)))
    td = _minus(ds, n)
    return (td.real[:ns] / n).astype(F32)
