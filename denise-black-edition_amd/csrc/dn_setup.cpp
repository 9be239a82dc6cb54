// Host-side set-up math of the P/SV path: FD coefficients, CPML profiles,
// stability check and material preparation.  Runs once per model on the host
// (O(NX*NY)); the results are uploaded to the device by dn_api.cpp.
// Every routine keeps the reference's float/double evaluation order so the
// device sees the same coefficients the reference computes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "dn_internal.h"

namespace dn {

static constexpr double kPi = 3.141592653589793;  // include/fd.h:22

// Staggered-grid first-derivative coefficients, Taylor (row 0) and Holberg
// (rows 1-4: maximum group-velocity error 0.1, 0.5, 1, 3 %), per half order.
// Entry 0 is the minimum number of grid points per shortest wavelength.
// Published table as used by src/holbergcoeff.c:29-65.
static const float kHc[5][6][7] = {
    {{23.0f, 1.0f, 0, 0, 0, 0, 0},
     {8.0f, (float)(9.0 / 8.0), (float)(-1.0 / 24.0), 0, 0, 0, 0},
     {6.0f, (float)(75.0 / 64.0), (float)(-25.0 / 384.0), (float)(3.0 / 640.0), 0, 0, 0},
     {5.0f, (float)(1225.0 / 1024.0), (float)(-245.0 / 3072.0), (float)(49.0 / 5120.0),
      (float)(-5.0 / 7168.0), 0, 0},
     {5.0f, (float)(19845.0 / 16384.0), (float)(-735.0 / 8192.0), (float)(567.0 / 40960.0),
      (float)(-405.0 / 229376.0), (float)(35.0 / 294912.0), 0},
     {4.0f, (float)(160083.0 / 131072.0), (float)(-12705.0 / 131072.0),
      (float)(22869.0 / 1310720.0), (float)(-5445.0 / 1835008.0), (float)(847.0 / 2359296.0),
      (float)(-63.0 / 2883584.0)}},
    {{49.7f, 1.0010f, 0, 0, 0, 0, 0},
     {8.32f, 1.1382f, -0.046414f, 0, 0, 0, 0},
     {4.77f, 1.1965f, -0.078804f, 0.0081781f, 0, 0, 0},
     {3.69f, 1.2257f, -0.099537f, 0.018063f, -0.0026274f, 0, 0},
     {3.19f, 1.2415f, -0.11231f, 0.026191f, -0.0064682f, 0.001191f, 0},
     {2.91f, 1.2508f, -0.12034f, 0.032131f, -0.010142f, 0.0029857f, -0.00066667f}},
    {{22.2f, 1.0050f, 0, 0, 0, 0, 0},
     {5.65f, 1.1534f, -0.052806f, 0, 0, 0, 0},
     {3.74f, 1.2111f, -0.088313f, 0.011768f, 0, 0, 0},
     {3.11f, 1.2367f, -0.10815f, 0.023113f, -0.0046905f, 0, 0},
     {2.80f, 1.2496f, -0.11921f, 0.031130f, -0.0093272f, 0.0025161f, 0},
     {2.62f, 1.2568f, -0.12573f, 0.036423f, -0.013132f, 0.0047484f, -0.0015979f}},
    {{15.8f, 1.0100f, 0, 0, 0, 0, 0},
     {4.80f, 1.1640f, -0.057991f, 0, 0, 0, 0},
     {3.39f, 1.2192f, -0.094070f, 0.014608f, 0, 0, 0},
     {2.90f, 1.2422f, -0.11269f, 0.026140f, -0.0064054f, 0, 0},
     {2.65f, 1.2534f, -0.12257f, 0.033755f, -0.011081f, 0.0036784f, 0},
     {2.51f, 1.2596f, -0.12825f, 0.038550f, -0.014763f, 0.0058619f, -0.0024538f}},
    {{9.16f, 1.0300f, 0, 0, 0, 0, 0},
     {3.47f, 1.1876f, -0.072518f, 0, 0, 0, 0},
     {2.91f, 1.2341f, -0.10569f, 0.022589f, 0, 0, 0},
     {2.61f, 1.2516f, -0.12085f, 0.032236f, -0.011459f, 0, 0},
     {2.45f, 1.2596f, -0.12829f, 0.038533f, -0.014681f, 0.0072580f, 0},
     {2.36f, 1.2640f, -0.13239f, 0.042217f, -0.017803f, 0.0081959f, -0.0051848f}},
};

int holberg(int fdorder, int maxrelerror, float hc[7], std::string &err)
{
    if (fdorder < 2 || fdorder > 12 || (fdorder % 2) != 0) {
        err = "Error in selection of FD coefficients: wrong FDORDER!";
        return DN_ERR_ARG;
    }
    if (maxrelerror < 0 || maxrelerror > 4) {
        err = "Error in selection of FD coefficients: wrong choice of maximum relative error!";
        return DN_ERR_ARG;
    }
    for (int k = 0; k < 7; ++k) hc[k] = kHc[maxrelerror][fdorder / 2 - 1][k];
    return DN_OK;
}

namespace {

// The damping gate of the CPML 'a' coefficient: the reference passes the
// float damping to the integer abs(), so 'a' stays zero unless |d| >= 1.
bool pml_gate(float d) { return std::abs((int)d) > 1.0e-6; }

// One CPML point: damping d, stretching K, frequency shift alpha from the
// normalised depth into the layer (Komatitsch & Martin 2007), then b and a.
struct PmlPoint {
    float d = 0.0f, K = 1.0f, al = 0.0f, a = 0.0f, b = 0.0f;
};

// sh: the SH profile of src/SH/PML_pro_SH.c:130-285 adds 0.1 alpha_max to
// alpha inside the layer.
void pml_point(float absc, float thick, float d0, float npower, float kmax, float alpha_max, bool sh,
               PmlPoint &p)
{
    if (absc >= 0.0f) {
        float an = absc / thick;
        p.d = (float)(d0 * std::pow((double)an, (double)npower));
        p.K = (float)(1.0 + (kmax - 1.0) * std::pow((double)an, (double)npower));
        p.al = sh ? (float)(alpha_max * (1.0 - an) + 0.1 * alpha_max) : (float)(alpha_max * (1.0 - an));
    }
}

void pml_finish(PmlPoint &p, float dt)
{
    p.b = (float)std::exp((double)(-(p.d / p.K + p.al) * dt));
    if (pml_gate(p.d)) p.a = (float)((double)p.d * ((double)p.b - 1.0) / (double)(p.K * (p.d + p.K * p.al)));
}

}  // namespace

// src/PML_pro.c:65-407 (SH: src/SH/PML_pro_SH.c:65-407) for a single
// whole-grid domain.
void pml_profiles(const dn_params &p, HostSetup &s)
{
    const bool sh = p.physics == 5;
    const int FW = p.fw;
    for (auto &v : s.pml) v.assign((size_t)2 * FW + 1, 0.0f);
    if (FW <= 0) return;
    const float DH = p.dh, DT = p.dt, npower = p.npower, kmax = p.k_max_pml;
    const float alpha_max = (float)(2.0 * kPi * (p.fpml / 2.0));
    const float thick = FW * DH;
    const float Rcoef = 0.001f;
    const float d0 = (float)((double)((-(npower + 1)) * p.damping) * std::log((double)Rcoef) /
                             (2.0 * (double)thick));
    auto put = [&](int h, const PmlPoint &g, const PmlPoint &hh, bool x) {
        int o = x ? 0 : P_DY;
        s.pml[o + P_DX][h] = g.d;   s.pml[o + P_KX][h] = g.K;   s.pml[o + P_ALX][h] = g.al;
        s.pml[o + P_AX][h] = g.a;   s.pml[o + P_BX][h] = g.b;
        s.pml[o + P_DXH][h] = hh.d; s.pml[o + P_KXH][h] = hh.K; s.pml[o + P_ALXH][h] = hh.al;
        s.pml[o + P_AXH][h] = hh.a; s.pml[o + P_BXH][h] = hh.b;
    };
    // Both axes: the near band (cells 1..FW) measures depth from the layer
    // origin inwards, the far band (N-FW+1..N) from the far origin outwards.
    // Only the x bands clamp a negative alpha to zero (the y loops of the
    // reference do not); the far half point can lie beyond the layer.
    for (int axis = 0; axis < 2; ++axis) {
        const bool x = axis == 0;
        const int N = x ? p.nx : p.ny;
        const float o_near = thick;
        const float o_far = (N - 1) * DH - thick;
        for (int i = 1; i <= FW; ++i) {
            PmlPoint g, hh;
            const float val = DH * (i - 1);
            pml_point(o_near - val, thick, d0, npower, kmax, alpha_max, sh, g);
            pml_point((float)(o_near - (val + DH / 2.0)), thick, d0, npower, kmax, alpha_max, sh, hh);
            if (x) {
                if (g.al < 0.0f) g.al = 0.0f;
                if (hh.al < 0.0f) hh.al = 0.0f;
            }
            pml_finish(g, DT);
            pml_finish(hh, DT);
            put(i, g, hh, x);
        }
        for (int i = N - FW + 1; i <= N; ++i) {
            const int h = i - N + 2 * FW;
            PmlPoint g, hh;
            const float val = DH * (i - 1);
            pml_point(val - o_far, thick, d0, npower, kmax, alpha_max, sh, g);
            pml_point((float)(val + DH / 2.0 - o_far), thick, d0, npower, kmax, alpha_max, sh, hh);
            if (x) {
                if (g.al < 0.0f) g.al = 0.0f;
                if (hh.al < 0.0f) hh.al = 0.0f;
            }
            pml_finish(g, DT);
            pml_finish(hh, DT);
            put(h, g, hh, x);
        }
    }
}

// Stability limit DT <= DH / (sqrt(2) * sum|hc| * vmax):
// src/checkfd_ssg_elastic.c:9-145 and, with the relaxed/unrelaxed phase
// velocities of the standard linear solids, src/checkfd_ssg_visc.c:12-177.
int check_stability(const dn_params &p, const float hc[7], const float *vp, const float *vs,
                    const float *rho, const float *taus, const float *taup, std::string &err)
{
    const size_t n = (size_t)p.nx * p.ny;
    float cmax_s = 0.0f, cmax_p = 0.0f;
    if (p.L == 0) {
        for (size_t k = 0; k < n; ++k) {
            float cs = p.invmat1 == 3 ? (float)std::sqrt(vs[k] / rho[k]) : vs[k];
            float cp = p.invmat1 == 3 ? (float)std::sqrt((vp[k] + 2.0 * vs[k]) / rho[k]) : vp[k];
            if (cmax_s < cs) cmax_s = cs;
            if (cmax_p < cp) cmax_p = cp;
        }
    } else {
        const float TS = p.ts;
        const float ws = (float)(2.0 * kPi / TS);
        float peta[MAXL + 1];
        for (int l = 1; l <= p.L; ++l) {
            float pts = (float)(1.0 / (2.0 * kPi * p.fl[l - 1]));
            peta[l] = p.dt / pts;
        }
        for (size_t k = 0; k < n; ++k) {
            float sumu = 0.0f, sumpi = 0.0f;
            for (int l = 1; l <= p.L; ++l) {
                float ts = p.dt / peta[l];
                sumu = (float)(sumu + ((ws * ws * ts * ts * taus[k]) / (1.0 + ws * ws * ts * ts)));
            }
            for (int l = 1; l <= p.L; ++l) {
                float ts = p.dt / peta[l];
                sumpi = (float)(sumpi + ((ws * ws * ts * ts * taup[k]) / (1.0 + ws * ws * ts * ts)));
            }
            float pu_ref = p.invmat1 == 3 ? vs[k] : rho[k] * vs[k] * vs[k];
            float pp_ref = p.invmat1 == 3 ? vp[k] + 2 * vs[k] : rho[k] * vp[k] * vp[k];
            float cs = (float)std::sqrt(pu_ref * (1.0 + p.L * taus[k]) / (rho[k] * (1.0 + sumu)));
            float cp = (float)std::sqrt(pp_ref * (1.0 + p.L * taup[k]) / (rho[k] * (1.0 + sumpi)));
            if (cmax_s < cs) cmax_s = cs;
            if (cmax_p < cp) cmax_p = cp;
        }
    }
    const float cmax = cmax_s > cmax_p ? cmax_s : cmax_p;
    const float gamma = std::fabs(hc[1]) + std::fabs(hc[2]) + std::fabs(hc[3]) + std::fabs(hc[4]) +
                        std::fabs(hc[5]) + std::fabs(hc[6]);
    const float dtstab = (float)(p.dh / (std::sqrt(2.0) * gamma * cmax));
    if (p.dt > dtstab) {
        char buf[256];
        std::snprintf(buf, sizeof buf,
                      " The simulation will get unstable, choose smaller DT. (DT=%e > %e)",
                      (double)p.dt, (double)dtstab);
        err = buf;
        return DN_ERR_CFL;
    }
    if (p.fw > p.nx || p.fw > p.ny) {
        err = " The width of the absorbing boundary is larger than one computational domain.";
        return DN_ERR_ARG;
    }
    return DN_OK;
}

namespace {

// Host grid with one periodic ghost ring (rows/cols 0 and N+1), reference
// orientation [j][i].  With a single rank the reference's matcopy exchange
// sends each edge to itself: row 0 <- row NY, row NY+1 <- row 1, then
// column 0 <- column NX, column NX+1 <- column 1 (ghost rows included).
struct Ring {
    int nx, ny;
    std::vector<float> v;
    Ring(int nx_, int ny_) : nx(nx_), ny(ny_), v((size_t)(nx_ + 2) * (ny_ + 2), 0.0f) {}
    float &at(int j, int i) { return v[(size_t)j * (nx + 2) + i]; }
    void load(const float *xmajor)
    {
        for (int i = 1; i <= nx; ++i)
            for (int j = 1; j <= ny; ++j) at(j, i) = xmajor[(size_t)(i - 1) * ny + (j - 1)];
        for (int i = 0; i <= nx + 1; ++i) {
            float top = at(1, i), bot = at(ny, i);
            at(ny + 1, i) = top;
            at(0, i) = bot;
        }
        for (int j = 0; j <= ny + 1; ++j) {
            float lef = at(j, 1), rig = at(j, nx);
            at(j, nx + 1) = lef;
            at(j, 0) = rig;
        }
    }
};

}  // namespace

// Material averaging and coefficient preparation:
// src/av_mue.c:5-51, src/av_rho.c:5-28, src/av_tau.c:5-15,
// src/PSV/prepare_update_s_visc_PSV.c:8-94 (peta: src/PSV/readmod_visc_PSV.c:29-32),
// per-column free-surface coefficients of src/PSV/surface_visc_PML_PSV.c:36-169.
void prepare_material(const dn_params &p, const Layout &lay, const float *vp, const float *vs,
                      const float *rho, const float *taus, const float *taup, HostMaterial &m)
{
    const int NX = p.nx, NY = p.ny, L = p.L;
    const float DT = p.dt;
    const size_t plane = lay.plane();
    Ring R(NX, NY), PI(NX, NY), U(NX, NY);
    R.load(rho);
    PI.load(vp);
    U.load(vs);
    auto dev = [&](int j, int i) { return lay.origin() + (size_t)lay.off(j, i); };
    m.rip.assign(plane, 0.0f);
    m.rjp.assign(plane, 0.0f);
    m.muipjp.assign(plane, 0.0f);
    m.f.assign(plane, 0.0f);
    m.g.assign(plane, 0.0f);
    m.vp.assign(plane, 0.0f);
    m.vs.assign(plane, 0.0f);
    m.rho.assign(plane, 0.0f);
    std::vector<float> uipjp((size_t)NX * NY);
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            const size_t o = dev(j, i);
            m.vp[o] = PI.at(j, i);
            m.vs[o] = U.at(j, i);
            m.rho[o] = R.at(j, i);
            // harmonic average of mu over the four cells around (i+1/2, j+1/2)
            float u1, u2, u3, u4;
            if (p.invmat1 == 3) {
                u1 = U.at(j, i); u2 = U.at(j, i + 1); u3 = U.at(j + 1, i); u4 = U.at(j + 1, i + 1);
            } else {
                u1 = R.at(j, i) * U.at(j, i) * U.at(j, i);
                u2 = R.at(j, i + 1) * U.at(j, i + 1) * U.at(j, i + 1);
                u3 = R.at(j + 1, i) * U.at(j + 1, i) * U.at(j + 1, i);
                u4 = R.at(j + 1, i + 1) * U.at(j + 1, i + 1) * U.at(j + 1, i + 1);
            }
            float mu_h = (float)(4.0 / ((1.0 / u1) + (1.0 / u2) + (1.0 / u3) + (1.0 / u4)));
            if (u1 == 0.0f || u2 == 0.0f || u3 == 0.0f || u4 == 0.0f) mu_h = 0.0f;
            uipjp[(size_t)(j - 1) * NX + (i - 1)] = mu_h;
            // reciprocal arithmetic-mean density at (i+1/2, j) and (i, j+1/2)
            float rj = (float)(1.0 / (0.5 * (R.at(j, i) + R.at(j + 1, i))));
            float ri = (float)(1.0 / (0.5 * (R.at(j, i) + R.at(j, i + 1))));
            if (R.at(j, i) < 1e-4 && R.at(j + 1, i) < 1e-4) rj = 0.0f;
            if (R.at(j, i) < 1e-4 && R.at(j, i + 1) < 1e-4) ri = 0.0f;
            m.rjp[o] = rj;
            m.rip[o] = ri;
            if (L == 0) {
                m.muipjp[o] = mu_h;
                if (p.invmat1 == 3) {
                    m.f[o] = U.at(j, i);
                    m.g[o] = PI.at(j, i);
                } else {
                    m.f[o] = R.at(j, i) * U.at(j, i) * U.at(j, i);
                    m.g[o] = R.at(j, i) * ((PI.at(j, i) * PI.at(j, i)) - 2 * U.at(j, i) * U.at(j, i));
                }
            }
        }
    if (L == 0) return;

    Ring TS(NX, NY), TP(NX, NY);
    TS.load(taus);
    TP.load(taup);
    float pts[MAXL + 1], peta[MAXL + 1];
    for (int l = 1; l <= L; ++l) {
        pts[l] = (float)(1.0 / (2.0 * kPi * p.fl[l - 1]));
        peta[l] = DT / pts[l];
    }
    // relaxed moduli are referenced to the first relaxation frequency
    const float ws = (float)(2.0 * kPi * p.fl[0]);
    float sumu = 0.0f, sumpi = 0.0f;
    for (int l = 1; l <= L; ++l) {
        sumu = (float)(sumu + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l])));
        sumpi = (float)(sumpi + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l])));
    }
    for (int l = 1; l <= L; ++l) {
        m.bip[l] = (float)(1.0 / (1.0 + (peta[l] * 0.5)));
        m.bjm[l] = (float)(1.0 / (1.0 + (peta[l] * 0.5)));
        m.cip[l] = (float)(1.0 - (peta[l] * 0.5));
        m.cjm[l] = (float)(1.0 - (peta[l] * 0.5));
        m.fs_bjm[l] = (float)(peta[l] / (1.0 + (peta[l] * 0.5)));
    }
    for (int l = 1; l <= L; ++l) m.peta[l] = peta[l];
    for (auto *v : {&m.vmu, &m.vts, &m.vpi, &m.vtp, &m.vmuij, &m.vtsij}) v->assign(plane, 0.0f);
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            const size_t o = dev(j, i);
            const float ts = TS.at(j, i), tp = TP.at(j, i);
            const float tsij = (float)(0.25 * (TS.at(j, i) + TS.at(j, i + 1) + TS.at(j + 1, i) + TS.at(j + 1, i + 1)));
            float mu, pi;
            if (p.invmat1 == 3) {
                mu = (float)(U.at(j, i) / (1.0 + sumu * ts));
                pi = (float)((PI.at(j, i) + 2 * U.at(j, i)) / (1.0 + sumpi * tp));
            } else {
                mu = (float)((U.at(j, i) * U.at(j, i) * R.at(j, i)) / (1.0 + sumu * ts));
                pi = (float)((PI.at(j, i) * PI.at(j, i) * R.at(j, i)) / (1.0 + sumpi * tp));
            }
            const float muipjp = (float)(uipjp[(size_t)(j - 1) * NX + (i - 1)] / (1.0 + sumu * tsij));
            m.vmu[o] = mu;
            m.vts[o] = ts;
            m.vpi[o] = pi;
            m.vtp[o] = tp;
            m.vmuij[o] = muipjp;
            m.vtsij[o] = tsij;
        }
    // free surface (row 1): relaxed moduli referenced to 1/TS
    const float ws_fs = (float)(2.0 * kPi * (1.0 / p.ts));
    float sumu_fs = 0.0f, sumpi_fs = 0.0f;
    for (int l = 1; l <= L; ++l) {
        sumu_fs = (float)(sumu_fs + ((ws_fs * ws_fs * pts[l] * pts[l]) / (1.0 + ws_fs * ws_fs * pts[l] * pts[l])));
        sumpi_fs = (float)(sumpi_fs + ((ws_fs * ws_fs * pts[l] * pts[l]) / (1.0 + ws_fs * ws_fs * pts[l] * pts[l])));
    }
    m.fs_fjm.assign((size_t)NX + 1, 0.0f);
    m.fs_g.assign((size_t)NX + 1, 0.0f);
    m.fs_dme.assign((size_t)NX + 1, 0.0f);
    for (int i = 1; i <= NX; ++i) {
        const int j = 1;
        const float ts = TS.at(j, i), tp = TP.at(j, i);
        float mu, pi;
        if (p.invmat1 == 3) {
            mu = (float)(U.at(j, i) / (1.0 + sumu_fs * ts));
            pi = (float)((PI.at(j, i) + 2 * U.at(j, i)) / (1.0 + sumpi_fs * tp));
        } else {
            mu = (float)((U.at(j, i) * U.at(j, i) * R.at(j, i)) / (1.0 + sumu_fs * ts));
            pi = (float)((PI.at(j, i) * PI.at(j, i) * R.at(j, i)) / (1.0 + sumpi_fs * tp));
        }
        const float fjm = (float)(mu * 2.0 * (1.0 + L * ts));
        const float g = (float)(pi * (1.0 + L * tp));
        const float djm = (float)(2.0 * mu * ts);
        const float e = pi * tp;
        m.fs_fjm[i] = fjm;
        m.fs_g[i] = g;
        m.fs_dme[i] = djm - e;
    }
}


// Stability limit of the VTI driver, src/VTI/checkfd_ssg_VTI.c:33-125: the
// largest of sqrt(c11/rho), sqrt(c13/rho), sqrt(c33/rho), sqrt(c44/rho)
// (NaN for a negative c13, which no comparison picks) as vmax.
int check_stability_vti(const dn_params &p, const float hc[7], const float *rho, const float *c11,
                        const float *c13, const float *c33, const float *c44, std::string &err)
{
    const size_t n = (size_t)p.nx * p.ny;
    float cmax = 0.0f;
    for (size_t k = 0; k < n; ++k) {
        for (const float *cm : {c11, c13, c33, c44}) {
            const float c = (float)std::sqrt(cm[k] / rho[k]);
            if (cmax < c) cmax = c;
        }
    }
    const float gamma = std::fabs(hc[1]) + std::fabs(hc[2]) + std::fabs(hc[3]) + std::fabs(hc[4]) +
                        std::fabs(hc[5]) + std::fabs(hc[6]);
    const float dtstab = (float)(p.dh / (std::sqrt(2.0) * gamma * cmax));
    if (p.dt > dtstab) {
        char buf[256];
        std::snprintf(buf, sizeof buf, " The simulation will get unstable, choose smaller DT. (DT=%e > %e)",
                      (double)p.dt, (double)dtstab);
        err = buf;
        return DN_ERR_CFL;
    }
    return DN_OK;
}

// Anisotropic material: the one-rank periodic ghosts of matcopy_elastic_VTI
// (src/VTI/matcopy_elastic_VTI.c:32-126), av_rho (src/av_rho.c:5-28) and
// av_harm (src/av_harm.c:13-19) of c44 (VTI, src/VTI/FD_VTI.c:180-197) or of
// the rotated d15, d35, d55 (TTI, src/TTI/FD_TTI.c:180-205); the rotation of
// src/TTI/rot_el_tensor_TTI.c:18-40 with its float locals, theta in degrees
// as read by src/TTI/readmod_elastic_TTI.c:84.
void prepare_material_aniso(const dn_params &p, const Layout &lay, const float *rho, const float *c11,
                            const float *c13, const float *c33, const float *c44, const float *theta,
                            HostMaterial &m)
{
    const int NX = p.nx, NY = p.ny;
    const size_t plane = lay.plane();
    auto dev = [&](int j, int i) { return lay.origin() + (size_t)lay.off(j, i); };
    auto xm = [&](int j, int i) { return (size_t)(i - 1) * NY + (j - 1); };
    Ring R(NX, NY);
    R.load(rho);
    m.rip.assign(plane, 0.0f);
    m.rjp.assign(plane, 0.0f);
    const int np = aniso_planes(p.physics);
    for (int k = 0; k < 8; ++k) m.ani[k].assign(k < np ? plane : 0, 0.0f);
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            const size_t o = dev(j, i);
            float rj = (float)(1.0 / (0.5 * (R.at(j, i) + R.at(j + 1, i))));
            float ri = (float)(1.0 / (0.5 * (R.at(j, i) + R.at(j, i + 1))));
            if (R.at(j, i) < 1e-4 && R.at(j + 1, i) < 1e-4) rj = 0.0f;
            if (R.at(j, i) < 1e-4 && R.at(j, i + 1) < 1e-4) ri = 0.0f;
            m.rjp[o] = rj;
            m.rip[o] = ri;
        }
    auto harm = [&](Ring &M, std::vector<float> &out) {
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                const float a = M.at(j, i), b = M.at(j, i + 1), c = M.at(j + 1, i), d = M.at(j + 1, i + 1);
                float v = (float)(4.0 / ((1.0 / a) + (1.0 / b) + (1.0 / c) + (1.0 / d)));
                if (a == 0.0f || b == 0.0f || c == 0.0f || d == 0.0f) v = 0.0f;
                out[dev(j, i)] = v;
            }
    };
    if (p.physics == 3) {
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                m.ani[0][dev(j, i)] = c11[xm(j, i)];
                m.ani[1][dev(j, i)] = c13[xm(j, i)];
                m.ani[2][dev(j, i)] = c33[xm(j, i)];
            }
        Ring C44(NX, NY);
        C44.load(c44);
        harm(C44, m.ani[3]);
        return;
    }
    // TTI: rotated tensor in x-major order, then the ghosts and averages of d15, d35, d55
    const size_t n = (size_t)NX * NY;
    std::vector<float> d15(n), d35(n), d55(n);
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            const size_t k = xm(j, i);
            const float th = (float)(theta[k] * kPi / 180.0);
            const float cos2 = (float)(std::cos((double)th) * std::cos((double)th));
            const float sin2 = (float)(std::sin((double)th) * std::sin((double)th));
            const float cos4 = cos2 * cos2;
            const float sin4 = sin2 * sin2;
            const float sincos = (float)(std::sin((double)th) * std::cos((double)th));
            const float cos2sin2 = cos2 * sin2;
            const float cos2msin2 = cos2 - sin2;
            const float a11 = c11[k], a13 = c13[k], a33 = c33[k], a44 = c44[k];
            const size_t o = dev(j, i);
            m.ani[0][o] = (float)((double)(cos2 * (a11 * cos2 + a13 * sin2) + sin2 * (a13 * cos2 + a33 * sin2)) +
                                  4.0 * (double)a44 * (double)cos2sin2);
            m.ani[1][o] = (float)((double)(a11 * cos2sin2 + a13 * sin4 + a13 * cos4 + a33 * cos2sin2) -
                                  4.0 * (double)a44 * (double)cos2sin2);
            d15[k] = (float)((double)((a13 - a11) * cos2 * sincos + (a33 - a13) * sin2 * sincos) +
                             2.0 * (double)sincos * (double)a44 * (double)cos2msin2);
            m.ani[4][o] = (float)((double)(sin2 * (a11 * sin2 + a13 * cos2) + cos2 * (a33 * cos2 + a13 * sin2)) +
                                  4.0 * (double)a44 * (double)cos2sin2);
            d35[k] = (float)((double)((a13 - a11) * sin2 * sincos + (a33 - a13) * cos2 * sincos) -
                             2.0 * (double)sincos * (double)a44 * (double)cos2msin2);
            d55[k] = (float)(((double)a44 * (1.0 - 2.0 * (double)sin2) * (1.0 - 2.0 * (double)sin2) +
                              (double)((a33 - a13) * cos2sin2)) -
                             (double)((a13 - a11) * cos2sin2));
            m.ani[2][o] = d15[k];
            m.ani[5][o] = d35[k];
        }
    Ring D15(NX, NY), D35(NX, NY), D55(NX, NY);
    D15.load(d15.data());
    D35.load(d35.data());
    D55.load(d55.data());
    harm(D15, m.ani[3]);
    harm(D35, m.ani[6]);
    harm(D55, m.ani[7]);
}


// Stability limit of the acoustic driver, src/AC/checkfd_acoustic.c:33-68:
// vmax = vp (INVMAT1 1) or sqrt(pi/rho) (INVMAT1 3).
int check_stability_ac(const dn_params &p, const float hc[7], const float *vp, const float *rho, std::string &err)
{
    const size_t n = (size_t)p.nx * p.ny;
    float cmax = 0.0f;
    for (size_t k = 0; k < n; ++k) {
        const float c = p.invmat1 == 3 ? (float)std::sqrt(vp[k] / rho[k]) : vp[k];
        if (cmax < c) cmax = c;
    }
    const float gamma = std::fabs(hc[1]) + std::fabs(hc[2]) + std::fabs(hc[3]) + std::fabs(hc[4]) +
                        std::fabs(hc[5]) + std::fabs(hc[6]);
    const float dtstab = (float)(p.dh / (std::sqrt(2.0) * gamma * cmax));
    if (p.dt > dtstab) {
        char buf[256];
        std::snprintf(buf, sizeof buf, " The simulation will get unstable, choose smaller DT. (DT=%e > %e)",
                      (double)p.dt, (double)dtstab);
        err = buf;
        return DN_ERR_CFL;
    }
    return DN_OK;
}

// checkfd_visc_SH.c:21-95: the unrelaxed shear velocity
// sqrt(mu (1 + L taus) / (rho (1 + sumu))), sumu at the source's centre
// frequency 2 pi / TS, against the Courant limit
int check_stability_sh_visc(const dn_params &p, const float hc[7], const float *vs, const float *rho,
                            const float *taus, std::string &err)
{
    const size_t n = (size_t)p.nx * p.ny;
    const float ws = (float)(2.0 * kPi / p.ts);
    float cmax = 0.0f;
    for (size_t k = 0; k < n; ++k) {
        float sumu = 0.0f;
        for (int l = 1; l <= p.L; ++l) {
            const float ts = (float)(1.0 / (2.0 * kPi * p.fl[l - 1]));
            sumu = (float)(sumu + ((ws * ws * ts * ts * taus[k]) / (1.0 + ws * ws * ts * ts)));
        }
        const float pu_ref = p.invmat1 == 3 ? vs[k] : rho[k] * vs[k] * vs[k];
        const float c = (float)std::sqrt(pu_ref * (1.0 + p.L * taus[k]) / (rho[k] * (1.0 + sumu)));
        if (cmax < c) cmax = c;
    }
    const float gamma = std::fabs(hc[1]) + std::fabs(hc[2]) + std::fabs(hc[3]) + std::fabs(hc[4]) +
                        std::fabs(hc[5]) + std::fabs(hc[6]);
    const float dtstab = (float)(p.dh / (std::sqrt(2.0) * gamma * cmax));
    if (p.dt > dtstab) {
        char buf[256];
        std::snprintf(buf, sizeof buf, " The simulation will get unstable, choose smaller DT. (DT=%e > %e)",
                      (double)p.dt, (double)dtstab);
        err = buf;
        return DN_ERR_CFL;
    }
    return DN_OK;
}

// Acoustic material (src/AC/FD_AC.c:184-207): the one-rank periodic ghosts
// of matcopy_acoustic_AC, av_rho (src/av_rho.c:5-28), and the modulus of
// the pressure update, g = rho vp^2 (INVMAT1 1) or the given modulus (3)
// (update_s_acoustic_PML_AC.c:303-307).
void prepare_material_ac(const dn_params &p, const Layout &lay, const float *vp, const float *rho, HostMaterial &m)
{
    const int NX = p.nx, NY = p.ny;
    const size_t plane = lay.plane();
    auto dev = [&](int j, int i) { return lay.origin() + (size_t)lay.off(j, i); };
    Ring R(NX, NY), PI(NX, NY);
    R.load(rho);
    PI.load(vp);
    m.rip.assign(plane, 0.0f);
    m.rjp.assign(plane, 0.0f);
    for (int k = 0; k < 8; ++k) m.ani[k].assign(k < 1 ? plane : 0, 0.0f);
    // the model itself for the gradient assembly (ass_gradAC.c reads ppi, prho)
    m.vp.assign(p.mode > 0 ? plane : 0, 0.0f);
    m.rho.assign(p.mode > 0 ? plane : 0, 0.0f);
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            const size_t o = dev(j, i);
            if (p.mode > 0) {
                m.vp[o] = PI.at(j, i);
                m.rho[o] = R.at(j, i);
            }
            float rj = (float)(1.0 / (0.5 * (R.at(j, i) + R.at(j + 1, i))));
            float ri = (float)(1.0 / (0.5 * (R.at(j, i) + R.at(j, i + 1))));
            if (R.at(j, i) < 1e-4 && R.at(j + 1, i) < 1e-4) rj = 0.0f;
            if (R.at(j, i) < 1e-4 && R.at(j, i + 1) < 1e-4) ri = 0.0f;
            m.rjp[o] = rj;
            m.rip[o] = ri;
            m.ani[0][o] = p.invmat1 == 3 ? PI.at(j, i) : R.at(j, i) * PI.at(j, i) * PI.at(j, i);
        }
}

// SH material (src/SH/FD_SH.c:175-199): the one-rank periodic ghosts of
// matcopy_elastic_SH, av_mu_SH (src/SH/av_mu_SH.c:14-50: harmonic means of
// the shear modulus at the sxz and syz points, in double; for INVMAT1 1 the
// modulus rho vs^2 in float and 0 next to a fluid cell) and inv_rho_SH
// (src/SH/inv_rho_SH.c:12-24: 1/rho at the vz points, 0 below 1e-4).
void prepare_material_sh(const dn_params &p, const Layout &lay, const float *vs, const float *rho, const float *taus,
                         HostMaterial &m)
{
    const int NX = p.nx, NY = p.ny, L = p.L;
    const size_t plane = lay.plane();
    auto dev = [&](int j, int i) { return lay.origin() + (size_t)lay.off(j, i); };
    Ring R(NX, NY), U(NX, NY), T(NX, NY);
    R.load(rho);
    U.load(vs);
    if (L > 0) T.load(taus);  // matcopy_SH.c: the periodic ghosts of taus too
    m.rip.assign(plane, 0.0f);
    m.rjp.clear();
    // elastic: uip, ujp; viscoelastic: fipjp, f, dip[1..L], d[1..L]
    for (int k = 0; k < 8; ++k) m.ani[k].assign(k < 2 + 2 * L ? plane : 0, 0.0f);
    // prepare_update_s_visc_SH.c:17-42 with readmod_visc_SH.c:24-27 (eta)
    float pts[MAXL + 1] = {0}, sumu = 0.0f;
    const float ws = L > 0 ? (float)(2.0 * kPi * p.fl[0]) : 0.0f;
    for (int l = 1; l <= L; ++l) {
        pts[l] = (float)(1.0 / (2.0 * kPi * p.fl[l - 1]));
        m.peta[l] = p.dt / pts[l];
        sumu = (float)(sumu + ((ws * ws * pts[l] * pts[l]) / (1.0 + ws * ws * pts[l] * pts[l])));
        m.bip[l] = (float)(1.0 / (1.0 + (m.peta[l] * 0.5)));
        m.bjm[l] = (float)(1.0 / (1.0 + (m.peta[l] * 0.5)));
        m.cip[l] = (float)(1.0 - (m.peta[l] * 0.5));
        m.cjm[l] = (float)(1.0 - (m.peta[l] * 0.5));
    }
    // the model itself for the adjoint run (sh.c:193-199 tests mu > 0) and
    // the assembly (ass_gradSH.c reads pu, prho)
    m.vs.assign(p.mode > 0 ? plane : 0, 0.0f);
    m.rho.assign(p.mode > 0 ? plane : 0, 0.0f);
    const bool mod = p.invmat1 == 3;
    auto modulus = [&](int j, int i) { return mod ? U.at(j, i) : R.at(j, i) * U.at(j, i) * U.at(j, i); };
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            const size_t o = dev(j, i);
            const float u1 = modulus(j, i), u2 = modulus(j, i + 1), u3 = modulus(j + 1, i);
            float uip = (float)(2.0 / ((1.0 / u1) + (1.0 / u2)));
            float ujp = (float)(2.0 / ((1.0 / u1) + (1.0 / u3)));
            if (!mod && (u1 == 0.0f || u2 == 0.0f)) uip = 0.0f;
            if (!mod && (u1 == 0.0f || u3 == 0.0f)) ujp = 0.0f;
            m.ani[0][o] = uip;
            m.ani[1][o] = ujp;
            if (L > 0) {
                // av_tau.c:5-15 and prepare_update_s_visc_SH.c:48-58
                const float ts = T.at(j, i);
                const float tsij =
                    (float)(0.25 * (T.at(j, i) + T.at(j, i + 1) + T.at(j + 1, i) + T.at(j + 1, i + 1)));
                const float mujp = (float)(ujp / (1.0 + sumu * ts));
                const float muip = (float)(uip / (1.0 + sumu * tsij));
                m.ani[0][o] = (float)(muip * p.dt * (1.0 + L * tsij));
                m.ani[1][o] = (float)(mujp * p.dt * (1.0 + L * ts));
                for (int l = 1; l <= L; ++l) {
                    m.ani[1 + l][o] = muip * m.peta[l] * tsij;
                    m.ani[1 + L + l][o] = mujp * m.peta[l] * ts;
                }
            }
            float ri = (float)(1.0 / R.at(j, i));
            if (R.at(j, i) < 1e-4) ri = 0.0f;
            m.rip[o] = ri;
            if (p.mode > 0) {
                m.vs[o] = U.at(j, i);
                m.rho[o] = R.at(j, i);
            }
        }
}
}  // namespace dn
