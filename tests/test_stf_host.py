"""Source-time-function inversion (dnfwi_stf, libdnio; src/stf.c:13-262)
against its numpy restatement (oracle/stf_oracle.py), and both against
what the mathematics fixes:
- observed traces that are the synthetic ones convolved with a known short
  filter h give back the spike wavelet convolved with h (Wiener
  deconvolution with a vanishing water level);
- identical traces give the closed form D_ss D conj(M) / (|M|^2 +
  EPS_STF (ntr+1)/2 sum m^2): the reference's float Ebar adds the running
  denominator after every trace, so its mean energy carries (ntr+1)/2, not
  1 (a plain mean would fail here);
- NORMALIZE 1 scales every trace's samples 2..ns to a largest |value| of 1,
  NORMALIZE 2 scales the observed traces to the synthetic maxima; muted
  traces are zero."""
import numpy as np
import pytest

import dnfwi
import stf_oracle as so
from cases import rel_l2

F32 = np.float32


def _section(rng, ntr, ns, tail=0):
    a = np.zeros((ntr, ns), F32)
    t = np.arange(ns - tail)
    for i in range(ntr):
        f = rng.uniform(0.02, 0.08)
        a[i, :ns - tail] = (np.sin(2 * np.pi * f * t + rng.uniform(0, 6)) * np.exp(-((t - rng.uniform(20, 60)) / 25.0) ** 2)
                            + 0.05 * rng.standard_normal(ns - tail)).astype(F32)
    return a


def _spike(ns, at=5):
    s = np.zeros(ns, F32)
    s[at:at + 4] = F32([0.3, 1.0, -0.5, 0.1])
    return s


@pytest.mark.parametrize("norm", [0, 1, 2])
def test_stf_equals_oracle(norm):
    rng = np.random.default_rng(3 + norm)
    ntr, ns = 7, 150
    syn = _section(rng, ntr, ns)
    obs = (1.7 * syn + 0.1 * _section(rng, ntr, ns)).astype(F32)
    mute = np.zeros(ntr, np.int32)
    mute[[1, 4]] = 1
    got, o2, s2 = dnfwi.stf(obs, syn, mute, norm, 0.05, _spike(ns))
    want = so.stf(obs, syn, mute.astype(bool), norm, 0.05, _spike(ns))
    err = rel_l2(got, want)
    print(f"NORMALIZE {norm}: product vs restatement rel-L2 {err:.2e}")
    assert err < 1e-6
    assert np.abs(got).max() > 0
    assert not o2[[1, 4]].any() and not s2[[1, 4]].any()


@pytest.mark.parametrize("impl", ["product", "oracle"])
def test_stf_recovers_a_known_filter(impl):
    rng = np.random.default_rng(11)
    ntr, ns = 6, 200
    h = np.array([0.0, 0.6, 1.0, -0.4, 0.2, 0.05])
    syn = _section(rng, ntr, ns, tail=len(h))              # zero tail: the convolution stays inside ns
    obs = np.stack([np.convolve(s.astype(np.float64), h)[:ns] for s in syn]).astype(F32)
    spike = _spike(ns)
    want = np.convolve(spike.astype(np.float64), h)[:ns]
    run = (lambda *a: dnfwi.stf(*a)[0]) if impl == "product" else so.stf
    got = run(obs, syn, None, 0, 1e-12, spike)
    err = rel_l2(got, want)
    print(f"{impl}: STF vs spike (*) h rel-L2 {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("impl", ["product", "oracle"])
def test_stf_water_level_of_identical_traces(impl):
    rng = np.random.default_rng(5)
    ns, eps = 120, 0.3
    m = _section(rng, 1, ns)[0]
    d = (m * F32(0.8) + F32(0.2) * np.roll(m, 3)).astype(F32)
    spike = _spike(ns)
    n = so.npad_of(ns)
    M = np.fft.ifft(m.astype(np.float64), n) * n
    D = np.fft.ifft(d.astype(np.float64), n) * n
    S = np.fft.ifft(spike.astype(np.float64), n) * n
    energy = float(np.sum(m.astype(np.float64) ** 2))
    run = (lambda *a: dnfwi.stf(*a)[0]) if impl == "product" else so.stf
    for ntr in (1, 3):
        want = np.fft.fft(S * D * np.conj(M) / (np.abs(M) ** 2 + eps * (ntr + 1) / 2 * energy), n).real[:ns] / n
        got = run(np.tile(d, (ntr, 1)), np.tile(m, (ntr, 1)), None, 0, eps, spike)
        err = rel_l2(got, want)
        print(f"{impl}, {ntr} identical traces: vs closed form rel-L2 {err:.2e}")
        assert err < 1e-5
        # a plain mean energy (factor 1 instead of (ntr+1)/2) differs clearly at ntr = 3
        if ntr == 3:
            plain = np.fft.fft(S * D * np.conj(M) / (np.abs(M) ** 2 + eps * energy), n).real[:ns] / n
            assert rel_l2(got, plain) > 100 * err


def test_stf_normalisations():
    rng = np.random.default_rng(8)
    ntr, ns = 5, 64
    syn = _section(rng, ntr, ns)
    obs = (3.0 * _section(rng, ntr, ns)).astype(F32)
    obs[2] = 0.0                                         # an all-zero trace stays zero (divisor 1)
    _, o1, s1 = dnfwi.stf(obs, syn, None, 1, 0.1, _spike(ns))
    for a, src in ((o1, obs), (s1, syn)):
        for i in range(ntr):
            assert a[i, 0] == src[i, 0]                  # sample 1 untouched
            if np.abs(src[i, 1:]).max() > 0:
                assert np.abs(a[i, 1:]).max() == pytest.approx(1.0, abs=1e-6)
    assert not o1[2].any()
    _, o2, s2 = dnfwi.stf(obs, syn, None, 2, 0.1, _spike(ns))
    assert np.array_equal(s2, syn)
    for i in (0, 1, 3, 4):
        assert np.abs(o2[i, 1:]).max() == pytest.approx(np.abs(syn[i, 1:]).max(), rel=1e-6)
    with pytest.raises(RuntimeError):
        dnfwi.stf(obs, syn, None, 3, 0.1, _spike(ns))
    assert dnfwi.load().dnfwi_stf_npad(150) == 1024 and dnfwi.load().dnfwi_stf_npad(256) == 1024
