"""Butterworth time-domain filter of libdnio (dnio_tdfilt: the reference's
apply_tdfilt over the vendored SEIFE recursive filters) on the CPU: equal
to the plain-Python restatement (oracle/tdfilt_oracle.py) bit for bit for
low-pass and band-pass, odd and even orders, and the band-limited spike of
QUELLART 6; the restatement's double filter is pinned by the magnitude of
the bilinear-transformed analog Butterworth filter,
|H|^2 = 1 / (1 + (W/eps)^(2m)) (low-pass) or 1 / (1 + (eps/W)^(2m))
(high-pass), W = 2 tan(w/2), eps = 2 pi dt fc, on the FFT of the impulse
response."""
import numpy as np
import pytest

import dnio
import tdfilt_oracle as tf

F32 = np.float32


def run(data, dt, order, fc2, fc1):
    return dnio.tdfilt(data, dt, order, fc2, fc1)


@pytest.mark.parametrize("order,fc2,fc1", [(1, 30.0, 0.0), (2, 25.0, 0.0), (4, 20.0, 0.0), (5, 15.0, 3.0),
                                           (6, 40.0, 5.0)])
def test_tdfilt_equals_restatement(order, fc2, fc1):
    rng = np.random.default_rng(order)
    data = rng.standard_normal((3, 700)).astype(F32)
    got = run(data, 1e-3, order, fc2, fc1)
    want = tf.apply_tdfilt(data, 1e-3, order, fc2, fc1)
    assert np.array_equal(got, want)


def test_spike_source_filtered():
    """QUELLART 6: a unit spike at sample 1 + iround(tshift/DT), band-limited
    by ORDER_SPIKE, FC_SPIKE_2, FC_SPIKE_1."""
    sig = np.zeros((2, 900), F32)
    sig[0, 40] = 1.0
    sig[1, 100] = 2.0
    got = run(sig, 5e-4, 4, 25.0, 2.0)
    assert np.array_equal(got, tf.apply_tdfilt(sig, 5e-4, 4, 25.0, 2.0))
    assert np.all(got[0, :40] == 0.0)          # causal: nothing before the spike


@pytest.mark.parametrize("m,highpass", [(4, False), (5, False), (3, True), (6, True)])
def test_filter_response_is_butterworth(m, highpass):
    n, dt, fc = 16384, 1e-3, 25.0
    x = [0.0] * n
    x[8] = 1.0
    h = np.array(tf.butterworth(x, dt, 1.0 / fc, m, highpass))
    H = np.abs(np.fft.rfft(h))
    w = 2 * np.pi * np.arange(H.size) / n
    W = 2 * np.tan(w / 2)
    eps = 2 * np.pi * dt * fc
    with np.errstate(divide="ignore"):
        ratio = (eps / W) if highpass else (W / eps)
        want = 1.0 / np.sqrt(1.0 + ratio ** (2 * m))
    sel = slice(1, H.size - 1)
    assert np.max(np.abs(H[sel] - want[sel])) < 1e-6
    # the corner: -3 dB at fc
    k = int(round(fc * n * dt))
    assert abs(H[k] ** 2 - 0.5) < 0.02
