/*
 * Butterworth time-domain filters of seismograms and source signals
 * (include/denise_io.h: dnio_tdfilt), as the reference applies them:
 * src/apply_tdfilt.c (low-pass at fc2, then high-pass at fc1 when fc1 > 0)
 * over src/timedomain_filt.c (per trace: float -> double, filter, -> float).
 *
 * The filters are E. Wielandt's recursive SEIFE filters (the reference's
 * vendored libcseife, cseife.c:41-52, 139-178; cseife_rfk.c; cseife_rekfl.c):
 * a Butterworth filter of order m is a cascade of bilinear-transformed
 * analog sections -- one first-order section when m is odd, and m/2
 * second-order sections s^2 + 2 h eps s + eps^2 with h = sin(pi (2j-1)/(2m)) --
 * with the corner eps = 2 pi dt / T0 in radians per sample, each run as the
 * recursion y[k] = f0 x[k] + f1 x[k-1] + f2 x[k-2] + g1 y[k-1] + g2 y[k-2]
 * from zero state.  The high-pass first subtracts the first sample from the
 * trace.  Double precision throughout, in the order written here.
 */
#include "denise_io.h"

#include <math.h>
#include <stdlib.h>

int dnio_fail_msg(const char *msg); /* dn_io.c */

typedef struct { double f0, f1, f2, g1, g2; } iir;

/* coefficients of one section; t0 = corner period in samples */
static iir section(int highpass, int second, double t0, double h)
{
    iir c = {0.0, 0.0, 0.0, 0.0, 0.0};
    const double eps = 8.0 * atan(1.0) / t0;
    if (!second) {
        c.g1 = (2.0 - eps) / (2.0 + eps);
        if (highpass) {
            c.f0 = 2.0 / (2.0 + eps);
            c.f1 = -c.f0;
        } else {
            c.f0 = eps / (2.0 + eps);
            c.f1 = c.f0;
        }
        return c;
    }
    const double e2 = eps * eps;
    const double a = 1.0 - eps * h + e2 / 4.0;
    const double b = -2.0 + e2 / 2.0;
    const double cc = 1.0 + eps * h + e2 / 4.0;
    c.g1 = -b / cc;
    c.g2 = -a / cc;
    if (highpass) {
        c.f0 = 1.0 / cc;
        c.f1 = -c.f0 - c.f0;
        c.f2 = c.f0;
    } else {
        c.f0 = e2 / 4.0 / cc;
        c.f1 = c.f0 + c.f0;
        c.f2 = c.f0;
    }
    return c;
}

static void run(double *x, int n, iir c)
{
    double xa = 0.0, xaa = 0.0, ya = 0.0, yaa = 0.0;
    for (int j = 0; j < n; j++) {
        const double xn = x[j];
        x[j] = c.f0 * xn + c.f1 * xa + c.f2 * xaa + c.g1 * ya + c.g2 * yaa;
        xaa = xa;
        xa = xn;
        yaa = ya;
        ya = x[j];
    }
}

static void butterworth(double *x, int n, double dt, int highpass, double t0, int m)
{
    if (highpass) {
        const double x0 = x[0];
        for (int i = 0; i < n; i++) x[i] -= x0;
    }
    const int mm = m / 2;
    if (m > 2 * mm) run(x, n, section(highpass, 0, t0 / dt, 0.0));
    const double pih = 2.0 * atan(1.0);
    for (int j = 1; j <= mm; ++j) run(x, n, section(highpass, 1, t0 / dt, sin(pih * (2 * j - 1) / m)));
}

/* timedomain_filt: one pass (method 1 low-pass, 2 high-pass) over every trace */
static int one_pass(float *data, int ntr, int ns, float dt, float fc, int order, int highpass)
{
    double *tr = malloc(sizeof(double) * (size_t)(ns > 0 ? ns : 1));
    if (!tr) return dnio_fail_msg("dnio_tdfilt: out of host memory");
    const double T0 = 1.0 / fc;
    for (int i = 0; i < ntr; i++) {
        float *d = data + (size_t)i * ns;
        for (int j = 0; j < ns; j++) tr[j] = (double)d[j];
        butterworth(tr, ns, dt, highpass, T0, order);
        for (int j = 0; j < ns; j++) d[j] = (float)tr[j];
    }
    free(tr);
    return 0;
}

int dnio_tdfilt(float *data, int ntr, int ns, float dt, int order, float fc2, float fc1)
{
    if (!data && ntr > 0) return dnio_fail_msg("dnio_tdfilt: null data");
    if (ntr <= 0 || ns <= 0) return 0;
    if (fc2 <= 0.0f) return dnio_fail_msg("dnio_tdfilt: the low-pass corner must be > 0");
    int rc = one_pass(data, ntr, ns, dt, fc2, order, 0);
    if (!rc && fc1 > 0.0f) rc = one_pass(data, ntr, ns, dt, fc1, order, 1);
    return rc;
}
