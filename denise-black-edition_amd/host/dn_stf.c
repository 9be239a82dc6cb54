/*
 * Source-time-function inversion of one shot (include/denise_fwi.h,
 * dnfwi_stf): the Wiener deconvolution of src/stf.c:13-262 with the trace
 * normalisations of src/util.c:489-579.
 *
 * The reference transforms with FFTW (plans of sign +1 for the traces and
 * the spike, sign -1 for the inverse, unnormalised).  FFTW is not part of
 * this build: the transforms here are an iterative radix-2 Cooley-Tukey FFT
 * in double precision with the same signs and lengths (Npad is a power of
 * two), whose twiddle factors are evaluated directly as cos/sin of
 * 2*pi*k/N.  Both are exact transforms up to double rounding, so the
 * wavelet equals the reference's up to that rounding before it is stored
 * as float.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "denise_fwi.h"

int dnio_fail_msg(const char *msg); /* dn_io.c */

/* in-place transform of n = 2^k points: x[k] = sum_m x[m] exp(sign*2*pi*i*k*m/n) */
static void fft_pow2(double complex *x, int n, int sign, const double complex *w)
{
    for (int i = 1, j = 0; i < n; i++) { /* bit-reversed order */
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j |= bit;
        if (i < j) {
            const double complex t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, stride = n / len;
        for (int s = 0; s < n; s += len)
            for (int k = 0; k < half; k++) {
                const double complex tw = sign > 0 ? w[k * stride] : conj(w[k * stride]);
                const double complex u = x[s + k], v = x[s + k + half] * tw;
                x[s + k] = u + v;
                x[s + k + half] = u - v;
            }
    }
}

/* util.c:489-526: each trace divided by its largest |sample| over samples
 * 2..ns (1 when that is 0); sample 1 is left as it is */
static void normalize_traces(float *d, int ntr, int ns)
{
    for (int i = 0; i < ntr; i++) {
        float *t = d + (size_t)i * ns;
        float max = 0.0f, min = 0.0f, tmp = 1.0f;
        for (int j = 1; j < ns; j++) {
            if (t[j] > max) max = t[j];
            if (t[j] < min) min = t[j];
            tmp = max > fabs(min) ? max : (float)fabs(min);
        }
        if (tmp == 0.0f) tmp = 1.0f;
        for (int j = 1; j < ns; j++) t[j] = t[j] / tmp;
    }
}

/* util.c:528-579: observed traces scaled to the synthetic traces' largest
 * |sample| over samples 2..ns */
static void normalize_traces_rel(const float *syn, float *obs, int ntr, int ns)
{
    for (int i = 0; i < ntr; i++) {
        const float *s = syn + (size_t)i * ns;
        float *o = obs + (size_t)i * ns;
        float max = 0.0f, min = 0.0f, max_o = 0.0f, min_o = 0.0f;
        for (int j = 1; j < ns; j++) {
            if (s[j] > max) max = s[j];
            if (s[j] < min) min = s[j];
            if (o[j] > max_o) max_o = o[j];
            if (o[j] < min_o) min_o = o[j];
        }
        float m = max > fabs(min) ? max : (float)fabs(min);
        float mo = max_o > fabs(min_o) ? max_o : (float)fabs(min_o);
        if (m == 0.0f) m = 1.0f;
        if (mo == 0.0f) mo = 1.0f;
        for (int j = 1; j < ns; j++) o[j] = o[j] * (m / mo);
    }
}

int dnfwi_stf_npad(int ns)
{
    /* stf.c:62: 2^(ceil(log2(ns)) + 2) */
    return (int)(pow(2.0, ceil(log((double)ns) / log(2.0)) + 2.0));
}

int dnfwi_stf(float *obs, float *syn, int ntr, int ns, const int *mute, int normalize, float eps_stf,
              const float *spike, float *stf_out)
{
    if (!obs || !syn || !spike || !stf_out || ntr < 1 || ns < 2)
        return dnio_fail_msg("dnfwi_stf: null argument or empty section");
    if (normalize < 0 || normalize > 2) return dnio_fail_msg("dnfwi_stf: NORMALIZE 0, 1 or 2 expected");
    const int npad = dnfwi_stf_npad(ns);
    if (npad < ns || npad > (1 << 26)) return dnio_fail_msg("dnfwi_stf: trace too long");
    double complex *w = malloc(sizeof(double complex) * (size_t)(npad / 2 + 1));
    double complex *sumn = calloc((size_t)npad, sizeof(double complex));
    double complex *sumd = calloc((size_t)npad, sizeof(double complex));
    double complex *dss = calloc((size_t)npad, sizeof(double complex));
    double complex *dd = malloc(sizeof(double complex) * (size_t)npad);
    double complex *dm = malloc(sizeof(double complex) * (size_t)npad);
    if (!w || !sumn || !sumd || !dss || !dd || !dm) {
        free(w), free(sumn), free(sumd), free(dss), free(dd), free(dm);
        return dnio_fail_msg("dnfwi_stf: out of host memory");
    }
    const double tau = 2.0 * 3.14159265358979323846;
    for (int k = 0; k <= npad / 2; k++) w[k] = cos(tau * k / npad) + sin(tau * k / npad) * I;

    /* stf.c:96-179: muted traces (TRKILL, OFFSET_MUTE, OFFSETC_STF) are zero
     * in both sections, then NORMALIZE */
    if (mute)
        for (int i = 0; i < ntr; i++)
            if (mute[i]) {
                memset(syn + (size_t)i * ns, 0, sizeof(float) * ns);
                memset(obs + (size_t)i * ns, 0, sizeof(float) * ns);
            }
    if (normalize == 1) {
        normalize_traces(syn, ntr, ns);
        normalize_traces(obs, ntr, ns);
    }
    if (normalize == 2) normalize_traces_rel(syn, obs, ntr, ns);

    /* stf.c:188-200: transform of the spike wavelet */
    for (int j = 0; j < ns; j++) dss[j] = (double)spike[j];
    fft_pow2(dss, npad, +1, w);

    /* stf.c:203-248: numerator and denominator summed over the traces; the
     * float Ebar takes the real part of the running denominator at every
     * sample of every trace */
    float ebar = 0.0f;
    for (int i = 0; i < ntr; i++) {
        const float *o = obs + (size_t)i * ns, *s = syn + (size_t)i * ns;
        for (int j = 0; j < npad; j++) {
            dd[j] = j < ns ? (double)o[j] : 0.0;
            dm[j] = j < ns ? (double)s[j] : 0.0;
        }
        fft_pow2(dd, npad, +1, w);
        fft_pow2(dm, npad, +1, w);
        for (int j = 0; j < npad; j++) {
            sumn[j] += ntr * dd[j] * conj(dm[j]);
            sumd[j] += ntr * dm[j] * conj(dm[j]);
            ebar = (float)creal(ebar + sumd[j]);
        }
    }
    ebar = ebar / (float)((long long)npad * ntr);

    /* stf.c:253-266: Wiener deconvolution, inverse transform (sign -1) */
    const float water = eps_stf * ntr * ebar;
    for (int j = 0; j < npad; j++) dd[j] = dss[j] * (sumn[j] / (sumd[j] + water));
    fft_pow2(dd, npad, -1, w);
    for (int j = 0; j < ns; j++) stf_out[j] = (float)(creal(dd[j]) / npad);

    free(w), free(sumn), free(sumd), free(dss), free(dd), free(dm);
    return 0;
}
