"""Seeded synthetic inputs for the P/SV workloads: smooth random layered
models, Ricker source signals and surface receiver lines.

Models are returned x-major (shape (NX, NY)) float32, the reference's
model-file order (src/PSV/readmod_elastic_PSV.c:72-76).
"""
from __future__ import annotations

import math

import numpy as np

PI = 3.141592653589793


def gaussian_smooth(field: np.ndarray, sigma: float) -> np.ndarray:
    """Separable Gaussian smoothing with reflecting edges (scipy when present)."""
    try:
        from scipy.ndimage import gaussian_filter
        return gaussian_filter(field, sigma, mode="reflect")
    except ImportError:  # small numpy fallback for the generator only
        r = int(4 * sigma + 0.5)
        x = np.arange(-r, r + 1)
        k = np.exp(-0.5 * (x / sigma) ** 2)
        k /= k.sum()
        out = field
        for ax in range(field.ndim):
            pad = [(0, 0)] * field.ndim
            pad[ax] = (r, r)
            padded = np.pad(out, pad, mode="symmetric")
            out = np.apply_along_axis(lambda v: np.convolve(v, k, mode="valid"), ax, padded)
        return out


def layered_model(nx, ny, dh, seed, vtop=1500.0, vbot=4000.0, pert=0.05, sigma=8.0):
    """vp linear in depth from vtop to vbot, times (1 + pert*p) with p a
    Gaussian-smoothed N(0,1) field normalised to max|p| = 1; vs = vp/sqrt(3);
    rho = 310 vp^0.25.  Returns (vp, vs, rho) x-major float32."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((nx, ny))
    p = gaussian_smooth(noise, sigma)
    p /= max(np.abs(p).max(), 1e-30)
    z = np.arange(ny, dtype=np.float64) * dh
    zmax = max((ny - 1) * dh, dh)
    vp0 = vtop + (vbot - vtop) * z / zmax
    vp = vp0[None, :] * (1.0 + pert * p)
    vs = vp / math.sqrt(3.0)
    rho = 310.0 * vp ** 0.25
    return vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32)


def quick_layered_model(nx, ny, dh, seed, vtop=1500.0, vbot=4000.0, pert=0.05, sigma=8.0, factor=4):
    """layered_model with the random field drawn on a grid `factor` times
    coarser (smoothing sigma/factor there) and interpolated linearly: the
    same kind of seeded smooth-random model at a fraction of the cost, for
    workloads of many large models."""
    from scipy.ndimage import zoom
    rng = np.random.default_rng(seed)
    cx, cy = -(-nx // factor) + 2, -(-ny // factor) + 2
    noise = rng.standard_normal((cx, cy))
    p = gaussian_smooth(noise, sigma / factor)
    p = zoom(p, (nx / cx, ny / cy), order=1)[:nx, :ny]
    p /= max(np.abs(p).max(), 1e-30)
    z = np.arange(ny, dtype=np.float64) * dh
    zmax = max((ny - 1) * dh, dh)
    vp0 = vtop + (vbot - vtop) * z / zmax
    vp = vp0[None, :] * (1.0 + pert * p)
    vs = vp / math.sqrt(3.0)
    rho = 310.0 * vp ** 0.25
    return vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32)


def anomaly(vp, vs, rho, dh, centre_x, centre_z, radius=300.0, amp=0.10):
    """Multiply a model by (1 + amp*G(r; radius)), G a unit Gaussian bump."""
    nx, ny = vp.shape
    x = np.arange(nx)[:, None] * dh
    z = np.arange(ny)[None, :] * dh
    g = np.exp(-((x - centre_x) ** 2 + (z - centre_z) ** 2) / (2.0 * radius ** 2))
    f = (1.0 + amp * g).astype(np.float64)
    return ((vp * f).astype(np.float32), (vs * f).astype(np.float32), (rho * f).astype(np.float32))


def ricker(nt, dt, fc, tshift=0.0, amp=1.0):
    """Ricker wavelet of src/wavelet.c:50-51 (QUELLART=1) evaluated with the
    same float/double rounding; returns signals[1..NT] as float32."""
    out = np.empty(nt, np.float32)
    f32 = np.float32
    ts = f32(1.0 / fc)
    for n in range(1, nt + 1):
        t = f32(f32(n) * f32(dt))
        tau = f32(PI * (float(t) - 1.5 * float(ts) - float(f32(tshift))) / (1.5 * float(ts)))
        a = f32((1.0 - 4.0 * float(tau) * float(tau)) * math.exp(-2.0 * float(tau) * float(tau)))
        out[n - 1] = f32(a * f32(amp))
    return out


def iround(x):
    return int(math.floor(x + 0.5))


def source_table(x, z, dh, fc, amp=1.0, tshift=0.0, stype=1, azimuth=0.0):
    """One source as the 8-row srcpos_loc column (src/splitsrc.c:28-36)."""
    s = np.zeros((8, 1), np.float32)
    s[0, 0] = iround(np.float32(x) / np.float32(dh))
    s[1, 0] = iround(np.float32(z) / np.float32(dh))
    s[2, 0] = 1
    s[3, 0] = tshift
    s[4, 0] = fc
    s[5, 0] = amp
    s[6, 0] = azimuth
    s[7, 0] = stype
    return s


def receiver_line(x0, dx, n, z, dh):
    """recpos (2, n) of receivers at x0 + k*dx, depth z (src/receiver.c:61-63)."""
    xs = [x0 + k * dx for k in range(n)]
    ix = [iround(np.float32(x) / np.float32(dh)) for x in xs]
    iy = [iround(np.float32(z) / np.float32(dh))] * n
    return np.array([ix, iy], np.int32)


def stable_dt(vmax, dh, fdorder=8, maxrelerror=1, safety=0.9):
    """Largest DT under the stability limit DH/(sqrt(2)*sum|hc|*vmax)."""
    hc_sum = {2: 1.0010, 4: 1.1846, 6: 1.2835, 8: 1.3458, 10: 1.3880, 12: 1.4171}
    return safety * dh / (math.sqrt(2.0) * hc_sum[fdorder] * vmax)


def anisotropic_model(nx, ny, dh, seed, eps=0.2, delta=0.1, tilt=None, sigma=8.0):
    """Seeded VTI model from the layered construction: vertical speeds vp0,
    vs0 and rho of layered_model, Thomsen epsilon and delta scaled by a
    smooth random field in [0.5, 1.5]; c33 = rho vp0^2, c44 = rho vs0^2,
    c11 = c33 (1 + 2 eps), c13 = sqrt(2 delta c33 (c33 - c44) + (c33 - c44)^2)
    - c44.  With tilt (degrees), also a smooth random tilt field tilt +- 10.
    Returns (rho, c11, c13, c33, c44, theta or None), x-major float32."""
    vp, vs, rho = layered_model(nx, ny, dh, seed, sigma=sigma)
    rng = np.random.default_rng(seed + 1000)
    f = gaussian_smooth(rng.standard_normal((nx, ny)), sigma)
    f = 1.0 + 0.5 * f / max(np.abs(f).max(), 1e-30)
    r = rho.astype(np.float64)
    c33 = r * vp.astype(np.float64) ** 2
    c44 = r * vs.astype(np.float64) ** 2
    e, d = eps * f, delta * f
    c11 = c33 * (1.0 + 2.0 * e)
    c13 = np.sqrt(2.0 * d * c33 * (c33 - c44) + (c33 - c44) ** 2) - c44
    theta = None
    if tilt is not None:
        g = gaussian_smooth(rng.standard_normal((nx, ny)), sigma)
        theta = (tilt + 10.0 * g / max(np.abs(g).max(), 1e-30)).astype(np.float32)
    return (rho, c11.astype(np.float32), c13.astype(np.float32), c33.astype(np.float32),
            c44.astype(np.float32), theta)
