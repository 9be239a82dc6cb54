// C ABI of libdnhip.so: context, device memory, the per-shot time loop.
// The loop enqueues every time step of a shot on the context's stream and
// synchronises once at the end (the reference crosses its process boundary
// twice per step for the halo exchange; here nothing leaves the device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "dn_ctx.h"

using namespace dn;

namespace dn {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

}  // namespace dn

namespace {

int upload(dn_ctx *c, float *dst, const float *src, size_t n)
{
    HIP_OK(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    return DN_OK;
}

void fill_args(dn_ctx *c)
{
    KArgs &a = c->a;
    const dn_params &p = c->p;
    const size_t org = c->lay.origin();
    a.nx = p.nx;
    a.ny = p.ny;
    a.pitch = c->lay.pitch;
    a.fw = p.fw;
    a.fw2 = 2 * p.fw;
    a.free_surf = p.free_surf;
    a.physics = p.physics == 0 ? 1 : p.physics;
    // the P/SV and acoustic loops mirror the ghost rows above the free
    // surface (surface_*_PML_PSV.c, surface_acoustic_PML_AC.c); VTI/TTI do not
    a.mirror = p.free_surf && (a.physics == 1 || a.physics == 2);
    a.boundary = p.boundary;
    a.grad_form = p.grad_form;
    a.L = p.L;
    a.nt_total = p.nt;
    a.dt = p.dt;
    a.dh = p.dh;
    a.dhi_el = p.dt / p.dh;
    a.dhi_vis = (float)(1.0 / p.dh);
    a.dthalbe = (float)(p.dt / 2.0);
    for (int k = 0; k < 7; ++k) a.hc[k] = c->setup.hc[k];
    a.hc[7] = 0.0f;
    a.plane = (long)c->lay.plane();
    auto at = [&](float *base) { return base ? base + org : nullptr; };
    a.vx = at(c->fld[F_VX]);
    a.vy = at(c->fld[F_VY]);
    a.sxx = at(c->fld[F_SXX]);
    a.syy = at(c->fld[F_SYY]);
    a.sxy = at(c->fld[F_SXY]);
    a.vxp1 = at(c->fld[F_VXP1]);
    a.vyp1 = at(c->fld[F_VYP1]);
    a.ux = at(c->fld[F_UX]);
    a.uy = at(c->fld[F_UY]);
    a.uxy = at(c->fld[F_UXY]);
    a.rip = at(c->rip);
    a.rjp = at(c->rjp);
    a.muipjp = at(c->muipjp);
    a.f = at(c->f);
    a.g = at(c->g);
    a.vp = at(c->vp);
    a.vs = at(c->vs);
    a.rho = at(c->rho);
    for (int k = 0; k < 8; ++k) a.ani[k] = at(c->ani[k]);
    // acoustic: the pressure lives in the sxx plane and is both normal
    // stress of the velocity sweep; sxy stays 0
    if (a.physics == 2) a.syy = a.sxx;
    a.rmem = at(c->rmem);
    a.pmem = at(c->pmem);
    a.qmem = at(c->qmem);
    a.vmu = at(c->vmu);
    a.vts = at(c->vts);
    a.vpi = at(c->vpi);
    a.vtp = at(c->vtp);
    a.vmuij = at(c->vmuij);
    a.vtsij = at(c->vtsij);
    a.tau_const = c->tau_const;
    a.ts0 = c->ts0;
    a.tp0 = c->tp0;
    a.tsij0 = c->tsij0;
    for (int l = 0; l <= MAXL; ++l) {
        a.peta[l] = c->mat.peta[l];
        a.bip[l] = c->mat.bip[l];
        a.cip[l] = c->mat.cip[l];
        a.bjm[l] = c->mat.bjm[l];
        a.cjm[l] = c->mat.cjm[l];
        a.fs_bjm[l] = c->mat.fs_bjm[l];
    }
    a.fs_fjm = c->fs_fjm;
    a.fs_g = c->fs_g;
    a.fs_dme = c->fs_dme;
    a.psi_sxx_x = c->psi[0];
    a.psi_sxy_x = c->psi[1];
    a.psi_vxx = c->psi[2];
    a.psi_vyx = c->psi[3];
    a.psi_vxxs = c->psi[4];
    a.psi_syy_y = c->psi[5];
    a.psi_sxy_y = c->psi[6];
    a.psi_vyy = c->psi[7];
    a.psi_vxy = c->psi[8];
    a.K_x = c->pml[P_KX];
    a.a_x = c->pml[P_AX];
    a.b_x = c->pml[P_BX];
    a.K_xh = c->pml[P_KXH];
    a.a_xh = c->pml[P_AXH];
    a.b_xh = c->pml[P_BXH];
    a.K_y = c->pml[P_KY];
    a.a_y = c->pml[P_AY];
    a.b_y = c->pml[P_BY];
    a.K_yh = c->pml[P_KYH];
    a.a_yh = c->pml[P_AYH];
    a.b_yh = c->pml[P_BYH];
}

// Tile shape used when the caller does not pick one: the widest tile that
// still gives every CU several workgroups.
Tile default_tile(int nx, int ny)
{
    (void)nx;
    (void)ny;
    Tile t;
    t.tbx = 64;
    t.by = 4;
    t.ty = 1;
    return t;
}

// Per-tile cell work lists of the two sweeps (see Ent in dn_internal.h).
int build_lists(dn_ctx *c)
{
    const int NX = c->p.nx, NY = c->p.ny;
    c->ntx_v = c->tile_v.ntx(NX);
    c->ntiles_v = c->ntx_v * c->tile_v.nty(NY);
    c->ntx_s = c->tile_s.ntx(NX);
    c->ntiles_s = c->ntx_s * c->tile_s.nty(NY);
    struct Item {
        int tile;
        Ent e;
    };
    std::vector<Item> lists[4];
    // lists 0 and 2 belong to the velocity sweep, 1 and 3 to the stress sweep
    auto add = [&](int which, int i, int j, int kind, int idx) {
        const Tile &t = (which == 0 || which == 2) ? c->tile_v : c->tile_s;
        const int ntx = (which == 0 || which == 2) ? c->ntx_v : c->ntx_s;
        const int tx = (i - 1) / t.width(), ty = (j - 1) / t.rows();
        Item it;
        it.tile = ty * ntx + tx;
        it.e.lane = (i - 1) - tx * t.width();
        it.e.row = j;
        it.e.kind = kind;
        it.e.idx = idx;
        lists[which].push_back(it);
    };
    const int ns = (int)c->src_ix.size(), nr = (int)c->rec_ix.size();
    const int seismo = c->p.seismo;
    // SH: QUELLTYP 1 is the force on vz (update_v_PML_SH.c:560-568); the SH
    // loop has no explosive source
    const bool sh = c->p.physics == 5;
    for (int l = 0; l < ns; ++l)
        if (sh ? c->src_type[l] == 1 : (c->src_type[l] >= 2 && c->src_type[l] <= 4))
            add(0, c->src_ix[l], c->src_iy[l], E_FORCE, l);
    if (seismo == 1 || seismo == 4)
        for (int r = 0; r < nr; ++r) add(0, c->rec_ix[r], c->rec_iy[r], E_REC_V, r);
    if (c->qtype == 1 && !sh)
        for (int l = 0; l < ns; ++l) add(1, c->src_ix[l], c->src_iy[l], E_EXPLOSIVE, l);
    if (seismo >= 2)
        for (int r = 0; r < nr; ++r) add(1, c->rec_ix[r], c->rec_iy[r], E_REC_S, r);
    for (int r = 0; r < nr; ++r) add(2, c->rec_ix[r], c->rec_iy[r], E_ADJ_FORCE, r);
    if (c->p.quelltypb >= 4)
        for (int r = 0; r < nr; ++r) add(3, c->rec_ix[r], c->rec_iy[r], E_ADJ_P, r);
    for (int w = 0; w < 4; ++w) {
        // by tile, then lane; entries of one cell keep their order
        std::stable_sort(lists[w].begin(), lists[w].end(), [](const Item &x, const Item &y) {
            return x.tile < y.tile || (x.tile == y.tile && x.e.lane < y.e.lane);
        });
        const int ntiles = (w == 0 || w == 2) ? c->ntiles_v : c->ntiles_s;
        std::vector<int> off((size_t)ntiles + 1, 0);
        std::vector<Ent> ent(lists[w].size() + 1);
        for (auto &it : lists[w]) off[(size_t)it.tile + 1]++;
        for (int k = 0; k < ntiles; ++k) off[(size_t)k + 1] += off[k];
        for (size_t k = 0; k < lists[w].size(); ++k) ent[k] = lists[w][k].e;
        // lean tiles: none of the sweep body's boundary or point paths can
        // trigger in them (the conditions below are the body's own tests,
        // taken over the tile's column and row range)
        {
            const bool vel = w == 0 || w == 2;
            const Tile &t = vel ? c->tile_v : c->tile_s;
            const int ntx = vel ? c->ntx_v : c->ntx_s;
            const dn_params &p = c->p;
            c->lean_h[w].clear();
            c->full_h[w].clear();
            for (int k = 0; k < ntiles; ++k) {
                const int tx = k % ntx, ty = k / ntx;
                const int ilo = tx * t.width() + 1, ihi = ilo + t.width() - 1;
                const int jt = ty * t.rows() + 1, jb = jt + t.rows() - 1;
                bool lean = true;
                if (!p.boundary && (ilo <= p.fw || ihi >= p.nx - p.fw + 1)) lean = false;  // x CPML
                if (!p.free_surf && jt <= p.fw) lean = false;                              // top CPML
                if (jb >= p.ny - p.fw + 1) lean = false;                                   // bottom CPML
                if (p.free_surf && jt <= c->fdoh) lean = false;                            // surface row / ghost rows
                if (off[(size_t)k + 1] > off[k]) lean = false;                             // sources, receivers
                if (p.nx / 2 >= ilo && p.nx / 2 <= ihi && p.ny / 2 >= jt && p.ny / 2 <= jb) lean = false;  // NaN probe
                (lean ? c->lean_h[w] : c->full_h[w]).push_back(k);
            }
            c->lst_nlean[w] = (int)c->lean_h[w].size();
            c->lst_nfull[w] = (int)c->full_h[w].size();
            off.insert(off.end(), c->lean_h[w].begin(), c->lean_h[w].end());
            off.insert(off.end(), c->full_h[w].begin(), c->full_h[w].end());
            // per-lane entry ranges of the tiles that have entries
            const int width = t.width();
            std::vector<int> slot((size_t)ntiles, -1), lo;
            int ns = 0;
            for (int k = 0; k < ntiles; ++k) {
                const int b0 = off[k], b1 = off[(size_t)k + 1];
                if (b1 == b0) continue;
                slot[k] = ns++;
                int e = b0;
                for (int l = 0; l <= width; ++l) {
                    while (e < b1 && lists[w][(size_t)e].e.lane < l) ++e;
                    lo.push_back(e);
                }
            }
            off.insert(off.end(), slot.begin(), slot.end());
            off.insert(off.end(), lo.begin(), lo.end());
        }
        if (c->lst_off[w]) c->dfree(c->lst_off[w]);
        if (c->lst_ent[w]) c->dfree(c->lst_ent[w]);
        c->lst_off[w] = nullptr;
        c->lst_ent[w] = nullptr;
        int rc;
        if ((rc = c->dalloc((void **)&c->lst_off[w], off.size() * sizeof(int))) ||
            (rc = c->dalloc((void **)&c->lst_ent[w], ent.size() * sizeof(Ent))))
            return rc;
        HIP_OK(h2d(c->stream, c->lst_off[w], off.data(), off.size() * sizeof(int)));
        HIP_OK(h2d(c->stream, c->lst_ent[w], ent.data(), ent.size() * sizeof(Ent)));
    }
    return DN_OK;
}

}  // namespace

StepCtl dn::base_ctl(dn_ctx *c, int lv, int ls)
{
    StepCtl s{};
    s.vlist.off = c->lst_off[lv];
    s.vlist.ent = c->lst_ent[lv];
    s.slist.off = c->lst_off[ls];
    s.slist.ent = c->lst_ent[ls];
    s.vlist.nlean = c->lst_nlean[lv];
    s.vlist.nfull = c->lst_nfull[lv];
    s.vlist.lean = c->lst_off[lv] + (c->ntiles_v + 1);
    s.vlist.full = s.vlist.lean + s.vlist.nlean;
    s.vlist.slot = s.vlist.full + s.vlist.nfull;
    s.vlist.lane_off = s.vlist.slot + c->ntiles_v;
    s.slist.nlean = c->lst_nlean[ls];
    s.slist.nfull = c->lst_nfull[ls];
    s.slist.lean = c->lst_off[ls] + (c->ntiles_s + 1);
    s.slist.full = s.slist.lean + s.slist.nlean;
    s.slist.slot = s.slist.full + s.slist.nfull;
    s.slist.lane_off = s.slist.slot + c->ntiles_s;
    for (int k = 0; k < 5; ++k) s.fwd[k] = c->fwd[k];
    s.wc = c->wc;
    s.wcu = c->wcu;
    s.wcr = c->wcr;
    s.quelltypb = c->p.quelltypb;
    s.invmat1 = c->p.invmat1;
    s.store_k = -1;
    s.corr_k = -1;
    return s;
}

namespace {

int validate(const dn_params *p)
{
    if (!p) return fail(DN_ERR_ARG, "null params");
    if (p->nx < 1 || p->ny < 1) return fail(DN_ERR_ARG, "NX and NY must be positive");
    if (p->nt < 2) return fail(DN_ERR_ARG, "NT must be at least 2");
    if (p->fdorder < 2 || p->fdorder > 12 || p->fdorder % 2)
        return fail(DN_ERR_ARG, "Error in selection of FD coefficients: wrong FDORDER!");
    if (p->L < 0 || p->L > MAXL) return fail(DN_ERR_ARG, "L must be within 0..8");
    if (p->boundary != 0) return fail(DN_ERR_ARG, "BOUNDARY=1 (periodic x) is not supported");
    if (p->fw < 0 || p->fw > p->nx || p->fw > p->ny)
        return fail(DN_ERR_ARG, " The width of the absorbing boundary is larger than one computational domain.");
    if (p->seismo < 0 || p->seismo > 4) return fail(DN_ERR_ARG, "SEISMO must be within 0..4");
    if (p->invmat1 != 1 && p->invmat1 != 3) return fail(DN_ERR_ARG, "INVMAT1 must be 1 or 3");
    if (p->mode > 0 && p->dtinv < 1) return fail(DN_ERR_ARG, "DTINV must be >= 1");
    if (p->mode > 0 && p->grad_form != 1 && p->grad_form != 2) return fail(DN_ERR_ARG, "GRAD_FORM must be 1 or 2");
    if (!(p->dh > 0.0f) || !(p->dt > 0.0f)) return fail(DN_ERR_ARG, "DH and DT must be positive");
    if (p->physics < 0 || p->physics > 5)
        return fail(DN_ERR_ARG, "PHYSICS must be 1 (P/SV), 2 (acoustic), 3 (VTI), 4 (TTI) or 5 (SH)");
    if (p->physics >= 2 && p->physics <= 4 && p->L != 0)
        return fail(DN_ERR_ARG, "acoustic, VTI and TTI runs are elastic: L = 0");
    // FD_SH.c:171-206: viscoelastic SH forward modelling; its coefficients
    // share the eight material planes (2 + 2 L)
    if (p->physics == 5 && p->L > 3) return fail(DN_ERR_ARG, "viscoelastic SH runs take L <= 3");
    if (p->physics == 5 && p->L > 0 && p->mode != 0)
        return fail(DN_ERR_ARG, "viscoelastic SH runs are forward modelling (MODE 0; FWI_SH_visc is not built)");
    // the acoustic and VTI / TTI loops also run the RTM passes (physics_AC.c,
    // physics_VTI.c:27-29, physics_TTI.c); the SH loop its FWI gradient
    // (sh.c mode 1, FWI_SH.c)
    // TTI.c stores ux / uy / uxy for GRAD_FORM 2, which its stress update never writes
    if (p->physics == 4 && p->mode > 0 && p->grad_form == 2)
        return fail(DN_ERR_ARG, "TTI adjoint runs need GRAD_FORM 1 (the TTI update writes no ux/uy/uxy)");
    // sh.c:133-135 samples vz into every section; outseis_SHfor.c writes SEISMO 1 only
    if (p->physics == 5 && p->seismo > 1) return fail(DN_ERR_ARG, "SH runs record vz (SEISMO 0 or 1)");
    if (p->L > 0) {
        for (int l = 0; l < p->L; ++l)
            if (!(p->fl[l] > 0.0f)) return fail(DN_ERR_ARG, "relaxation frequencies FL must be positive");
        if (p->free_surf && !(p->ts > 0.0f)) return fail(DN_ERR_ARG, "TS must be positive for the viscoelastic free surface");
    }
    return DN_OK;
}

}  // namespace

int dn::zero_state(dn_ctx *c)
{
    const size_t pb = c->lay.plane() * sizeof(float);
    for (int k = 0; k < F_N; ++k)
        if (c->fld[k]) HIP_OK(hipMemsetAsync(c->fld[k], 0, pb, c->stream));
    if (c->p.L > 0) {
        HIP_OK(hipMemsetAsync(c->rmem, 0, pb * c->p.L, c->stream));
        HIP_OK(hipMemsetAsync(c->pmem, 0, pb * c->p.L, c->stream));
        HIP_OK(hipMemsetAsync(c->qmem, 0, pb * c->p.L, c->stream));
    }
    const size_t nxb = (size_t)c->p.ny * 2 * c->p.fw, nyb = (size_t)2 * c->p.fw * c->p.nx;
    for (int k = 0; k < 9; ++k) HIP_OK(hipMemsetAsync(c->psi[k], 0, (k < 5 ? nxb : nyb) * sizeof(float) + 4, c->stream));
    HIP_OK(hipMemsetAsync(c->acq.nan_step, 0, sizeof(int), c->stream));
    return DN_OK;
}

namespace {

hipEvent_t pool_event(dn_ctx *c, size_t k)
{
    while (c->pool.size() <= k) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->pool.push_back(e);
    }
    return c->pool[k];
}

int finish_run(dn_ctx *c, size_t nsample_events)
{
    HIP_OK(hipEventRecord(c->ev1, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->run_ms = ms;
    c->v_ms = c->s_ms = 0.0;
    c->samples = 0;
    for (size_t k = 0; k + 3 < nsample_events; k += 4) {
        float tv = 0.0f, ts = 0.0f;
        HIP_OK(hipEventElapsedTime(&tv, c->pool[k], c->pool[k + 1]));
        HIP_OK(hipEventElapsedTime(&ts, c->pool[k + 2], c->pool[k + 3]));
        c->v_ms += tv;
        c->s_ms += ts;
        c->samples++;
    }
    if (c->samples) {
        c->v_ms /= c->samples;
        c->s_ms /= c->samples;
    }
    int nan_step = 0;
    HIP_OK(d2h(c->stream, &nan_step, c->acq.nan_step, sizeof(int)));
    if (nan_step)
        return fail(DN_ERR_UNSTABLE, " Simulation is unstable ! (NaN at time step " + std::to_string(nan_step) + ")");
    return DN_OK;
}

int receiver_factors(dn_ctx *c)
{
    const int ntr = c->acq.ntr;
    if (ntr == 0 || c->h_vp.empty()) return DN_OK;
    std::vector<double> sl(ntr), sm(ntr);
    for (int t = 0; t < ntr; ++t) {
        const size_t k = (size_t)(c->rec_ix[t] - 1) * c->p.ny + (c->rec_iy[t] - 1);
        const float pi = c->h_vp[k], u = c->h_vs[k], r = c->h_rho[k];
        if (c->p.invmat1 == 3) {
            sl[t] = std::sqrt((double)pi);
            sm[t] = std::sqrt((double)u);
        } else {
            const float muss = u * u * r;
            const float lamss = pi * pi * r;
            sl[t] = std::sqrt((double)lamss);
            sm[t] = std::sqrt((double)muss);
        }
    }
    HIP_OK(h2d(c->stream, c->acq.rec_sqlam, sl.data(), sizeof(double) * ntr));
    HIP_OK(h2d(c->stream, c->acq.rec_sqmu, sm.data(), sizeof(double) * ntr));
    return DN_OK;
}

}  // namespace

namespace dn {
// the device buffers, one entry per allocation, that dn_ctx_create,
// dn_set_sources and dn_set_receivers make (cell lists aside)
std::vector<size_t> plan_sizes(const dn_params &p, int nsrc, int ntr)
{
    Layout lay;
    lay.init(p.nx, p.ny);
    const size_t pb = lay.plane() * sizeof(float);
    const bool gf2 = p.mode > 0 && p.grad_form == 2;
    std::vector<size_t> v;
    auto add = [&](size_t n, size_t count) {
        for (size_t k = 0; k < count; ++k) v.push_back(n == 0 ? 4 : n);
    };
    add(pb, F_N - (gf2 ? 0 : 3));
    add(pb, 2);                                   // rip, rjp
    const int nani = aniso_planes(p.physics, p.L);
    if (nani) add(pb, nani);                      // stiffness planes
    else if (p.L == 0) add(pb, 3);                // muipjp, f, g
    if (p.mode > 0) add(pb, 3);                   // vp, vs, rho
    if (p.L > 0) {
        add(pb * p.L, 3);
        add(pb, 6);
        add(sizeof(float) * (p.nx + 1), 3);
    }
    add((size_t)p.ny * 2 * p.fw * sizeof(float) + 4, 5);
    add((size_t)2 * p.fw * p.nx * sizeof(float) + 4, 4);
    add(sizeof(float) * (2 * p.fw + 1), P_NVEC);
    if (p.mode > 0) {
        const size_t ntd = (size_t)std::ceil((float)p.nt / (float)p.dtinv);
        add((size_t)p.nx * p.ny * ntd * sizeof(float), (size_t)snap_planes(p.physics));
        add((size_t)p.nx * p.ny * sizeof(float), 6);
    }
    add(sizeof(int), 1);
    add(sizeof(double) * l2_scratch_doubles(), 1);
    if (nsrc > 0) {
        add(sizeof(int) * nsrc, 3);
        add(sizeof(double) * nsrc, 2);
        add(sizeof(float) * (size_t)nsrc * p.nt, 1);
    }
    if (ntr > 0) {
        add(sizeof(int) * ntr, 2);
        add(sizeof(double) * ntr, 2);
        add((size_t)ntr * p.nt * sizeof(float), 9);
    }
    return v;
}
}  // namespace dn

extern "C" {

int dn_abi_version(void) { return DN_ABI_VERSION; }

const char *dn_last_error(void) { return g_err.c_str(); }

int dn_device_count(int *count)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    if (count) *count = n;
    return e == hipSuccess ? DN_OK : fail(DN_ERR_HIP, hipGetErrorString(e));
}

// Device bytes a context with these parameters, nsrc sources and ntr
// receivers holds (cell lists aside, which are a few bytes per point).
static size_t plan_bytes(const dn_params &p, int nsrc, int ntr)
{
    size_t b = 0;
    for (size_t n : plan_sizes(p, nsrc, ntr)) b += n;
    return b;
}

int dn_device_memory(int device, size_t *free_bytes, size_t *total_bytes)
{
    if (!free_bytes || !total_bytes) return fail(DN_ERR_ARG, "null output pointer");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(DN_ERR_ARG, "no such HIP device");
    HIP_OK(hipSetDevice(device));
    HIP_OK(hipMemGetInfo(free_bytes, total_bytes));
    return DN_OK;
}

int dn_estimate_footprint(const dn_params *p, int nsrc, int ntr, size_t *bytes)
{
    if (!p || !bytes || nsrc < 0 || ntr < 0) return fail(DN_ERR_ARG, "bad footprint request");
    int rc = validate(p);
    if (rc) return rc;
    *bytes = plan_bytes(*p, nsrc, ntr);
    return DN_OK;
}

static int ctx_create(const dn_params *p, int device, dn_arena *arena, dn_ctx **out);

int dn_ctx_create(const dn_params *p, int device, dn_ctx **out) { return ctx_create(p, device, nullptr, out); }

int dn_ctx_create_in(const dn_params *p, dn_arena *arena, dn_ctx **out)
{
    if (!arena) return fail(DN_ERR_ARG, "null arena");
    return ctx_create(p, arena->device, arena, out);
}

static int ctx_create(const dn_params *p, int device, dn_arena *arena, dn_ctx **out)
{
    if (!out) return fail(DN_ERR_ARG, "null output pointer");
    *out = nullptr;
    int rc = validate(p);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(DN_ERR_HIP, "no HIP device available: libdnhip needs an AMD GPU");
    if (device < 0 || device >= ndev) return fail(DN_ERR_ARG, "device ordinal out of range");
    dn_ctx *c = new (std::nothrow) dn_ctx();
    if (!c) return fail(DN_ERR_OOM, "host allocation failed");
    c->p = *p;
    c->device = device;
    c->arena = arena;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return fail(DN_ERR_HIP, "hipSetDevice failed");
    }
    std::string err;
    rc = holberg(p->fdorder, p->maxrelerror, c->setup.hc, err);
    if (rc) {
        delete c;
        return fail(rc, err);
    }
    c->fdoh = p->fdorder / 2;
    c->nd = c->fdoh + 1;
    c->ntdtinv = p->mode > 0 ? (int)std::ceil((float)p->nt / (float)p->dtinv) : 0;
    c->lay.init(p->nx, p->ny);
    pml_profiles(*p, c->setup);
    auto bail = [&](int code) {
        std::string m = g_err;
        delete c;
        g_err = m;
        return code;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(fail(DN_ERR_HIP, "stream creation failed"));
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return bail(fail(DN_ERR_HIP, "event creation failed"));
    const size_t pb = c->lay.plane() * sizeof(float);
    const bool gf2 = p->mode > 0 && p->grad_form == 2;
    for (int k = 0; k < F_N; ++k) {
        if ((k == F_UX || k == F_UY || k == F_UXY) && !gf2) continue;
        if ((rc = c->dalloc((void **)&c->fld[k], pb))) return bail(rc);
    }
    if ((rc = c->dalloc((void **)&c->rip, pb)) || (rc = c->dalloc((void **)&c->rjp, pb))) return bail(rc);
    if (aniso_planes(p->physics, p->L)) {
        for (int k = 0; k < aniso_planes(p->physics, p->L); ++k)
            if ((rc = c->dalloc((void **)&c->ani[k], pb))) return bail(rc);
    } else if (p->L == 0) {
        float **mats[] = {&c->muipjp, &c->f, &c->g};
        for (auto m : mats)
            if ((rc = c->dalloc((void **)m, pb))) return bail(rc);
    }
    if (p->mode > 0) {
        if ((rc = c->dalloc((void **)&c->vp, pb)) || (rc = c->dalloc((void **)&c->vs, pb)) ||
            (rc = c->dalloc((void **)&c->rho, pb)))
            return bail(rc);
    }
    if (p->L > 0) {
        float **mem[] = {&c->rmem, &c->pmem, &c->qmem};
        for (auto m : mem)
            if ((rc = c->dalloc((void **)m, pb * p->L))) return bail(rc);
        float **vis[] = {&c->vmu, &c->vts, &c->vpi, &c->vtp, &c->vmuij, &c->vtsij};
        for (auto m : vis)
            if ((rc = c->dalloc((void **)m, pb))) return bail(rc);
        if ((rc = c->dalloc((void **)&c->fs_fjm, sizeof(float) * (p->nx + 1))) ||
            (rc = c->dalloc((void **)&c->fs_g, sizeof(float) * (p->nx + 1))) ||
            (rc = c->dalloc((void **)&c->fs_dme, sizeof(float) * (p->nx + 1))))
            return bail(rc);
    }
    const size_t nxb = (size_t)p->ny * 2 * p->fw, nyb = (size_t)2 * p->fw * p->nx;
    for (int k = 0; k < 9; ++k)
        if ((rc = c->dalloc((void **)&c->psi[k], (k < 5 ? nxb : nyb) * sizeof(float) + 4))) return bail(rc);
    for (int k = 0; k < P_NVEC; ++k) {
        if ((rc = c->dalloc((void **)&c->pml[k], sizeof(float) * (2 * p->fw + 1)))) return bail(rc);
        if (h2d(c->stream, c->pml[k], c->setup.pml[k].data(), sizeof(float) * (2 * p->fw + 1)) != hipSuccess)
            return bail(fail(DN_ERR_HIP, "upload of CPML profiles failed"));
    }
    if (p->mode > 0) {
        const size_t snap = (size_t)p->nx * p->ny * c->ntdtinv * sizeof(float);
        for (int k = 0; k < 5; ++k)
            if (snap_plane_used(p->physics, k) && (rc = c->dalloc((void **)&c->fwd[k], snap))) return bail(rc);
        const size_t g = (size_t)p->nx * p->ny * sizeof(float);
        float **grads[] = {&c->wc, &c->wcu, &c->wcr, &c->gvp, &c->gvs, &c->grho};
        for (auto m : grads)
            if ((rc = c->dalloc((void **)m, g))) return bail(rc);
    }
    if ((rc = c->dalloc((void **)&c->acq.nan_step, sizeof(int))) || (rc = c->dalloc((void **)&c->acq.l2, sizeof(double) * l2_scratch_doubles())))
        return bail(rc);
    c->acq.nt = p->nt;
    c->tile_v = c->tile_s = default_tile(p->nx, p->ny);
    if ((rc = build_lists(c))) return bail(rc);
    fill_args(c);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(fail(DN_ERR_HIP, "stream synchronisation failed"));
    *out = c;
    return DN_OK;
}

int dn_ctx_destroy(dn_ctx *ctx)
{
    delete ctx;
    return DN_OK;
}

// acoustic model (params.physics 2): readmod_AC's vp (INVMAT1 1) or
// modulus (3) and rho, x-major; vs unused
static int set_model_ac(dn_ctx *c, const float *vp, const float *rho)
{
    if (!vp || !rho) return fail(DN_ERR_ARG, "null argument");
    HIP_OK(hipSetDevice(c->device));
    std::string err;
    int rc = check_stability_ac(c->p, c->setup.hc, vp, rho, err);
    if (rc) return fail(rc, err);
    prepare_material_ac(c->p, c->lay, vp, rho, c->mat);
    const size_t pl = c->lay.plane();
    if ((rc = upload(c, c->rip, c->mat.rip.data(), pl)) || (rc = upload(c, c->rjp, c->mat.rjp.data(), pl)) ||
        (rc = upload(c, c->ani[0], c->mat.ani[0].data(), pl)))
        return rc;
    if (c->vp && ((rc = upload(c, c->vp, c->mat.vp.data(), pl)) || (rc = upload(c, c->rho, c->mat.rho.data(), pl))))
        return rc;
    HIP_OK(hipStreamSynchronize(c->stream));
    for (auto *v : {&c->mat.rip, &c->mat.rjp, &c->mat.ani[0], &c->mat.vp, &c->mat.rho}) {
        v->clear();
        v->shrink_to_fit();
    }
    // seismo_AC's div and curl factors both come from vp (src/AC/ac.c:199-201)
    const size_t n = (size_t)c->p.nx * c->p.ny;
    c->h_vp.assign(vp, vp + n);
    c->h_vs.assign(vp, vp + n);
    c->h_rho.assign(rho, rho + n);
    fill_args(c);
    c->model_set = true;
    c->fwd_complete = false;
    return receiver_factors(c);
}

// SH model (params.physics 5): readmod_elastic_SH's vs (INVMAT1 1) or mu
// (INVMAT1 3) and rho; checkfd_elast_SH's DT limit is the acoustic one with
// the shear velocity (checkfd_elast_SH.c:30-75), then matcopy_elastic_SH,
// av_mu_SH and inv_rho_SH (src/SH/FD_SH.c:175-199).
static int set_model_sh(dn_ctx *c, const float *vs, const float *rho, const float *taus)
{
    if (!vs || !rho) return fail(DN_ERR_ARG, "null argument");
    if (c->p.L > 0 && !taus) return fail(DN_ERR_ARG, "a viscoelastic SH model needs taus");
    HIP_OK(hipSetDevice(c->device));
    std::string err;
    int rc = c->p.L > 0 ? check_stability_sh_visc(c->p, c->setup.hc, vs, rho, taus, err)
                        : check_stability_ac(c->p, c->setup.hc, vs, rho, err);
    if (rc) return fail(rc, err);
    prepare_material_sh(c->p, c->lay, vs, rho, taus, c->mat);
    const size_t pl = c->lay.plane();
    if ((rc = upload(c, c->rip, c->mat.rip.data(), pl))) return rc;
    for (int k = 0; k < aniso_planes(5, c->p.L); ++k)
        if ((rc = upload(c, c->ani[k], c->mat.ani[k].data(), pl))) return rc;
    if (c->vs && ((rc = upload(c, c->vs, c->mat.vs.data(), pl)) || (rc = upload(c, c->rho, c->mat.rho.data(), pl))))
        return rc;
    HIP_OK(hipStreamSynchronize(c->stream));
    for (int k = 2; k < 8; ++k) {
        c->mat.ani[k].clear();
        c->mat.ani[k].shrink_to_fit();
    }
    for (auto *v : {&c->mat.rip, &c->mat.ani[0], &c->mat.ani[1], &c->mat.vs, &c->mat.rho}) {
        v->clear();
        v->shrink_to_fit();
    }
    fill_args(c);
    c->model_set = true;
    c->fwd_complete = false;
    return DN_OK;
}

int dn_set_model(dn_ctx *c, const float *vp, const float *vs, const float *rho, const float *taus,
                 const float *taup)
{
    if (c && c->p.physics == 2) return set_model_ac(c, vp, rho);
    if (c && c->p.physics == 5) return set_model_sh(c, vs, rho, taus);
    if (!c || !vp || !vs || !rho) return fail(DN_ERR_ARG, "null argument");
    if (aniso_planes(c->p.physics)) return fail(DN_ERR_ARG, "VTI/TTI models are set with dn_set_model_aniso");
    if (c->p.L > 0 && (!taus || !taup)) return fail(DN_ERR_ARG, "viscoelastic model needs taus and taup");
    HIP_OK(hipSetDevice(c->device));
    std::string err;
    int rc = check_stability(c->p, c->setup.hc, vp, vs, rho, taus, taup, err);
    if (rc) return fail(rc, err);
    prepare_material(c->p, c->lay, vp, vs, rho, taus, taup, c->mat);
    const size_t pl = c->lay.plane();
    if ((rc = upload(c, c->rip, c->mat.rip.data(), pl)) || (rc = upload(c, c->rjp, c->mat.rjp.data(), pl)))
        return rc;
    if (c->p.L == 0) {
        if ((rc = upload(c, c->muipjp, c->mat.muipjp.data(), pl)) || (rc = upload(c, c->f, c->mat.f.data(), pl)) ||
            (rc = upload(c, c->g, c->mat.g.data(), pl)))
            return rc;
    }
    if (c->vp) {
        if ((rc = upload(c, c->vp, c->mat.vp.data(), pl)) || (rc = upload(c, c->vs, c->mat.vs.data(), pl)) ||
            (rc = upload(c, c->rho, c->mat.rho.data(), pl)))
            return rc;
    }
    if (c->p.L > 0) {
        if ((rc = upload(c, c->vmu, c->mat.vmu.data(), pl)) || (rc = upload(c, c->vts, c->mat.vts.data(), pl)) ||
            (rc = upload(c, c->vpi, c->mat.vpi.data(), pl)) || (rc = upload(c, c->vtp, c->mat.vtp.data(), pl)) ||
            (rc = upload(c, c->vmuij, c->mat.vmuij.data(), pl)) || (rc = upload(c, c->vtsij, c->mat.vtsij.data(), pl)) ||
            (rc = upload(c, c->fs_fjm, c->mat.fs_fjm.data(), c->p.nx + 1)) ||
            (rc = upload(c, c->fs_g, c->mat.fs_g.data(), c->p.nx + 1)) ||
            (rc = upload(c, c->fs_dme, c->mat.fs_dme.data(), c->p.nx + 1)))
            return rc;
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    // uniform relaxation times over the physical grid: pass them as scalars
    c->tau_const = 0;
    if (c->p.L > 0) {
        auto uniform = [&](const std::vector<float> &v, float &val) {
            val = v[c->lay.origin()];
            for (int j = 1; j <= c->p.ny; ++j) {
                const float *row = v.data() + c->lay.origin() + (size_t)(j - 1) * c->lay.pitch;
                for (int i = 0; i < c->p.nx; ++i)
                    if (std::memcmp(&row[i], &val, sizeof(float)) != 0) return false;
            }
            return true;
        };
        c->tau_const = uniform(c->mat.vts, c->ts0) && uniform(c->mat.vtp, c->tp0) && uniform(c->mat.vtsij, c->tsij0);
    }
    const size_t n = (size_t)c->p.nx * c->p.ny;
    c->h_vp.assign(vp, vp + n);
    c->h_vs.assign(vs, vs + n);
    c->h_rho.assign(rho, rho + n);
    // host staging of the device-layout material is no longer needed
    c->mat.rip.clear(); c->mat.rip.shrink_to_fit();
    c->mat.rjp.clear(); c->mat.rjp.shrink_to_fit();
    c->mat.muipjp.clear(); c->mat.muipjp.shrink_to_fit();
    c->mat.f.clear(); c->mat.f.shrink_to_fit();
    c->mat.g.clear(); c->mat.g.shrink_to_fit();
    c->mat.vp.clear(); c->mat.vp.shrink_to_fit();
    c->mat.vs.clear(); c->mat.vs.shrink_to_fit();
    c->mat.rho.clear(); c->mat.rho.shrink_to_fit();
    for (auto *v : {&c->mat.vmu, &c->mat.vts, &c->mat.vpi, &c->mat.vtp, &c->mat.vmuij, &c->mat.vtsij}) {
        v->clear();
        v->shrink_to_fit();
    }
    fill_args(c);
    c->model_set = true;
    c->fwd_complete = false;
    return receiver_factors(c);
}

int dn_set_model_aniso(dn_ctx *c, const float *rho, const float *c11, const float *c13, const float *c33,
                       const float *c44, const float *theta)
{
    if (!c || !rho || !c11 || !c13 || !c33 || !c44) return fail(DN_ERR_ARG, "null argument");
    const int np = aniso_planes(c->p.physics);
    if (np < 4) return fail(DN_ERR_ARG, "dn_set_model_aniso needs params.physics 3 (VTI) or 4 (TTI)");
    if (c->p.physics == 4 && !theta) return fail(DN_ERR_ARG, "a TTI model needs theta");
    HIP_OK(hipSetDevice(c->device));
    std::string err;
    int rc;
    if (c->p.physics == 3 && (rc = check_stability_vti(c->p, c->setup.hc, rho, c11, c13, c33, c44, err)))
        return fail(rc, err);
    prepare_material_aniso(c->p, c->lay, rho, c11, c13, c33, c44, theta, c->mat);
    const size_t pl = c->lay.plane();
    if ((rc = upload(c, c->rip, c->mat.rip.data(), pl)) || (rc = upload(c, c->rjp, c->mat.rjp.data(), pl)))
        return rc;
    for (int k = 0; k < np; ++k)
        if ((rc = upload(c, c->ani[k], c->mat.ani[k].data(), pl))) return rc;
    HIP_OK(hipStreamSynchronize(c->stream));
    c->mat.rip.clear(); c->mat.rip.shrink_to_fit();
    c->mat.rjp.clear(); c->mat.rjp.shrink_to_fit();
    for (auto &v : c->mat.ani) {
        v.clear();
        v.shrink_to_fit();
    }
    fill_args(c);
    c->model_set = true;
    c->fwd_complete = false;
    return DN_OK;
}

int dn_set_sources(dn_ctx *c, int nsrc, const float *srcpos, const float *signals)
{
    if (!c || nsrc < 0 || (nsrc > 0 && (!srcpos || !signals))) return fail(DN_ERR_ARG, "bad source arguments");
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipStreamSynchronize(c->stream));
    DevAcq &q = c->acq;
    for (void *ptr : {(void *)q.src_ix, (void *)q.src_iy, (void *)q.src_type, (void *)q.src_sin, (void *)q.src_cos, (void *)q.signals})
        if (ptr) c->dfree(ptr);
    q.src_ix = q.src_iy = q.src_type = nullptr;
    q.src_sin = q.src_cos = nullptr;
    q.signals = nullptr;
    q.nsrc = 0;
    c->qtype = 0;
    c->has_force = 0;
    c->src_ix.clear();
    c->src_iy.clear();
    c->src_type.clear();
    if (nsrc == 0) return build_lists(c);
    std::vector<int> ix(nsrc), iy(nsrc), type(nsrc);
    std::vector<double> sn(nsrc), cs(nsrc);
    for (int l = 0; l < nsrc; ++l) {
        ix[l] = (int)srcpos[0 * nsrc + l];
        iy[l] = (int)srcpos[1 * nsrc + l];
        type[l] = (int)srcpos[7 * nsrc + l];
        // azimuth in radians as the reference's float: src/PSV/update_v_PML_PSV.c:743
        const float azi = (float)(srcpos[6 * nsrc + l] * 3.141592653589793 / 180);
        sn[l] = std::sin((double)azi);
        cs[l] = std::cos((double)azi);
        if (ix[l] < 1 || ix[l] > c->p.nx || iy[l] < 1 || iy[l] > c->p.ny)
            return fail(DN_ERR_ARG, "source " + std::to_string(l + 1) + " lies outside the grid");
        if (type[l] == 5 && (ix[l] < 2 || iy[l] < 1)) return fail(DN_ERR_ARG, "moment-tensor source on the left edge");
        if (type[l] >= 2 && type[l] <= 4) c->has_force = 1;
    }
    c->qtype = type[nsrc - 1];
    int rc;
    if ((rc = c->dalloc((void **)&q.src_ix, sizeof(int) * nsrc)) || (rc = c->dalloc((void **)&q.src_iy, sizeof(int) * nsrc)) ||
        (rc = c->dalloc((void **)&q.src_type, sizeof(int) * nsrc)) ||
        (rc = c->dalloc((void **)&q.src_sin, sizeof(double) * nsrc)) ||
        (rc = c->dalloc((void **)&q.src_cos, sizeof(double) * nsrc)) ||
        (rc = c->dalloc((void **)&q.signals, sizeof(float) * (size_t)nsrc * c->p.nt)))
        return rc;
    HIP_OK(h2d(c->stream, q.src_ix, ix.data(), sizeof(int) * nsrc));
    HIP_OK(h2d(c->stream, q.src_iy, iy.data(), sizeof(int) * nsrc));
    HIP_OK(h2d(c->stream, q.src_type, type.data(), sizeof(int) * nsrc));
    HIP_OK(h2d(c->stream, q.src_sin, sn.data(), sizeof(double) * nsrc));
    HIP_OK(h2d(c->stream, q.src_cos, cs.data(), sizeof(double) * nsrc));
    HIP_OK(h2d(c->stream, q.signals, signals, sizeof(float) * (size_t)nsrc * c->p.nt));
    q.nsrc = nsrc;
    c->src_ix = ix;
    c->src_iy = iy;
    c->src_type = type;
    return build_lists(c);
}

int dn_set_receivers(dn_ctx *c, int ntr, const int *recpos)
{
    if (!c || ntr < 0 || (ntr > 0 && !recpos)) return fail(DN_ERR_ARG, "bad receiver arguments");
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipStreamSynchronize(c->stream));
    DevAcq &q = c->acq;
    std::set<std::pair<int, int>> seen;
    for (int t = 0; t < ntr; ++t) {
        const int ix = recpos[t], iy = recpos[ntr + t];
        if (ix < 1 || ix > c->p.nx || iy < 1 || iy > c->p.ny)
            return fail(DN_ERR_ARG, "receiver " + std::to_string(t + 1) + " lies outside the grid");
        if (!seen.insert({ix, iy}).second)
            return fail(DN_ERR_ARG, "receiver " + std::to_string(t + 1) + " repeats a grid position");
    }
    std::vector<void *> old = {q.rec_ix, q.rec_iy, q.rec_sqlam, q.rec_sqmu, q.obs};
    for (int k = 0; k < 5; ++k) old.push_back(q.sec[k]);
    for (int k = 0; k < 3; ++k) old.push_back(q.adj[k]);
    for (void *ptr : old)
        if (ptr) c->dfree(ptr);
    q.rec_ix = q.rec_iy = nullptr;
    q.rec_sqlam = q.rec_sqmu = nullptr;
    q.obs = nullptr;
    for (auto &s : q.sec) s = nullptr;
    for (auto &s : q.adj) s = nullptr;
    q.ntr = 0;
    c->rec_ix.assign(recpos, recpos + ntr);
    c->rec_iy.assign(recpos + ntr, recpos + 2 * ntr);
    if (ntr == 0) return build_lists(c);
    const size_t tr = (size_t)ntr * c->p.nt * sizeof(float);
    int rc;
    if ((rc = c->dalloc((void **)&q.rec_ix, sizeof(int) * ntr)) || (rc = c->dalloc((void **)&q.rec_iy, sizeof(int) * ntr)) ||
        (rc = c->dalloc((void **)&q.rec_sqlam, sizeof(double) * ntr)) ||
        (rc = c->dalloc((void **)&q.rec_sqmu, sizeof(double) * ntr)) || (rc = c->dalloc((void **)&q.obs, tr)))
        return rc;
    for (int k = 0; k < 5; ++k)
        if ((rc = c->dalloc((void **)&q.sec[k], tr))) return rc;
    for (int k = 0; k < 3; ++k)
        if ((rc = c->dalloc((void **)&q.adj[k], tr))) return rc;
    HIP_OK(h2d(c->stream, q.rec_ix, c->rec_ix.data(), sizeof(int) * ntr));
    HIP_OK(h2d(c->stream, q.rec_iy, c->rec_iy.data(), sizeof(int) * ntr));
    q.ntr = ntr;
    int rc2 = build_lists(c);
    if (rc2) return rc2;
    return receiver_factors(c);
}

int dn_set_tile(dn_ctx *c, int sweep, int tbx, int by, int ty, int vw)
{
    if (!c) return fail(DN_ERR_ARG, "null context");
    if (sweep < -1 || sweep > 1) return fail(DN_ERR_ARG, "sweep must be -1 (both), 0 (velocity) or 1 (stress)");
    Tile t;
    t.tbx = tbx;
    t.by = by;
    t.ty = ty;
    t.vw = vw;
    if (!tile_supported(t)) return fail(DN_ERR_ARG, "tile shape not compiled into the library");
    if (vw != 1 && scalar_only(c->p.physics))
        return fail(DN_ERR_ARG, "the acoustic and SH sweeps are compiled for scalar columns (vw = 1) only");
    if (sweep != 0 && vw != 1 && aniso_planes(c->p.physics))
        return fail(DN_ERR_ARG, "the VTI/TTI stress sweep is compiled for scalar columns (vw = 1) only");
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipStreamSynchronize(c->stream));
    if (sweep != 1) c->tile_v = t;
    if (sweep != 0) c->tile_s = t;
    return build_lists(c);
}

int dn_get_tile(dn_ctx *c, int sweep, int *tbx, int *by, int *ty, int *vw)
{
    if (!c || sweep < 0 || sweep > 1) return fail(DN_ERR_ARG, "null context or sweep not 0/1");
    const Tile &t = sweep == 0 ? c->tile_v : c->tile_s;
    if (tbx) *tbx = t.tbx;
    if (by) *by = t.by;
    if (ty) *ty = t.ty;
    if (vw) *vw = t.vw;
    return DN_OK;
}

int dn_autotune_tile(dn_ctx *c, int nsteps, int *vel_shape, int *str_shape)
{
    if (!c || nsteps < 1) return fail(DN_ERR_ARG, "autotune needs a context and nsteps >= 1");
    if (!c->model_set) return fail(DN_ERR_STATE, "dn_set_model must precede dn_autotune_tile");
    Tile shapes[32];
    const int n = tile_list(shapes, 32);
    const int keep_sampling = c->sample_every;
    c->sample_every = 1;  // time every launch of both sweeps
    Tile best_v = c->tile_v, best_s = c->tile_s;
    double bv = -1.0, bs = -1.0;
    int rc = DN_OK;
    // the VTI/TTI stress sweep has scalar columns only: with a vector shape
    // the velocity sweep is timed beside the default stress shape
    const bool scalar_s = aniso_planes(c->p.physics) > 0;
    for (int k = 0; k < n && k < 32 && !rc; ++k) {
        if (scalar_only(c->p.physics) && shapes[k].vw != 1) continue;  // acoustic / SH: scalar columns only
        const bool both = !(scalar_s && shapes[k].vw != 1);
        if (both)
            rc = dn_set_tile(c, -1, shapes[k].tbx, shapes[k].by, shapes[k].ty, shapes[k].vw);
        else if (!(rc = dn_set_tile(c, 0, shapes[k].tbx, shapes[k].by, shapes[k].ty, shapes[k].vw)))
            rc = dn_set_tile(c, 1, 64, 4, 1, 1);
        if (!rc) rc = dn_run_forward(c, nsteps);  // warm-up (code load, caches)
        if (!rc) rc = dn_run_forward(c, nsteps);
        if (rc) break;
        if (bv < 0.0 || c->v_ms < bv) {
            bv = c->v_ms;
            best_v = shapes[k];
        }
        if (both && (bs < 0.0 || c->s_ms < bs)) {
            bs = c->s_ms;
            best_s = shapes[k];
        }
    }
    if (!rc) {
        // second pass: plain or streaming access to the once-per-step
        // planes, per sweep, at the best shapes.  The two sweeps interact
        // through the caches, so the four pairs are timed by the whole step.
        c->tile_v = best_v;
        c->tile_s = best_s;
        rc = build_lists(c);
        double best = -1.0;
        int hv = 0, hs = 0;
        for (int h = 0; h < 4 && !rc; ++h) {
            c->nt_v = h & 1;
            c->nt_s = h >> 1;
            rc = dn_run_forward(c, nsteps);
            if (!rc) rc = dn_run_forward(c, nsteps);
            const double t = c->v_ms + c->s_ms;
            if (!rc && (best < 0.0 || t < best)) {
                best = t;
                hv = c->nt_v;
                hs = c->nt_s;
            }
        }
        c->nt_v = hv;
        c->nt_s = hs;
    }
    c->sample_every = keep_sampling;
    if (rc) return rc;
    c->tile_v = best_v;
    c->tile_s = best_s;
    if ((rc = build_lists(c))) return rc;
    if ((rc = zero_state(c))) return rc;
    c->fwd_complete = false;
    const Tile *bt[2] = {&best_v, &best_s};
    int *out[2] = {vel_shape, str_shape};
    for (int w = 0; w < 2; ++w)
        if (out[w]) {
            out[w][0] = bt[w]->tbx;
            out[w][1] = bt[w]->by;
            out[w][2] = bt[w]->ty;
            out[w][3] = bt[w]->vw;
        }
    return DN_OK;
}

int dn_set_store_hint(dn_ctx *c, int sweep, int streaming)
{
    if (!c || sweep < -1 || sweep > 1 || streaming < 0 || streaming > 1)
        return fail(DN_ERR_ARG, "bad store hint");
    if (sweep != 1) c->nt_v = streaming;
    if (sweep != 0) c->nt_s = streaming;
    return DN_OK;
}

int dn_get_store_hint(dn_ctx *c, int sweep, int *streaming)
{
    if (!c || !streaming || sweep < 0 || sweep > 1) return fail(DN_ERR_ARG, "bad store hint request");
    *streaming = sweep == 0 ? c->nt_v : c->nt_s;
    return DN_OK;
}

int dn_set_energy_weights(dn_ctx *c, int eprecond)
{
    if (!c || (eprecond != 0 && eprecond != 1 && eprecond != 3))
        return fail(DN_ERR_ARG, "EPRECOND 0, 1 or 3 expected");
    if (eprecond && c->p.mode <= 0) return fail(DN_ERR_STATE, "energy weights need params.mode > 0");
    HIP_OK(hipSetDevice(c->device));
    const size_t n = (size_t)c->p.nx * c->p.ny * sizeof(float);
    for (int w = 0; w < 2 && eprecond; ++w)
        if (!c->ew[w]) {
            int rc = c->dalloc((void **)&c->ew[w], n);
            if (rc) return rc;
            HIP_OK(hipMemsetAsync(c->ew[w], 0, n, c->stream));
        }
    c->eprecond = eprecond;
    return DN_OK;
}

int dn_get_energy(dn_ctx *c, int which, float *out)
{
    if (!c || !out || which < 0 || which > 1) return fail(DN_ERR_ARG, "bad energy request");
    if (!c->ew[which]) return fail(DN_ERR_STATE, "energy weights are off (dn_set_energy_weights)");
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipStreamSynchronize(c->stream));
    const int NX = c->p.nx, NY = c->p.ny;
    std::vector<float> tmp((size_t)NX * NY);
    HIP_OK(d2h(c->stream, tmp.data(), c->ew[which], tmp.size() * sizeof(float)));
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j) out[(size_t)i * NY + j] = tmp[(size_t)j * NX + i];
    return DN_OK;
}

int dn_set_snapshots(dn_ctx *c, int first, int inc, int last, dn_snap_fn fn, void *user)
{
    if (!c) return fail(DN_ERR_ARG, "null context");
    c->snap_fn = fn;
    c->snap_user = user;
    c->snap_first = first;
    c->snap_inc = inc;
    c->snap_last = last;
    return DN_OK;
}

int dn_set_kernel_sampling(dn_ctx *c, int every)
{
    if (!c || every < 0) return fail(DN_ERR_ARG, "bad sampling interval");
    c->sample_every = every;
    return DN_OK;
}

int dn_run_forward(dn_ctx *c, int nsteps)
{
    if (!c) return fail(DN_ERR_ARG, "null context");
    if (!c->model_set) return fail(DN_ERR_STATE, "dn_set_model must precede dn_run_forward");
    const int NT = c->p.nt;
    if (nsteps <= 0 || nsteps > NT) nsteps = NT;
    HIP_OK(hipSetDevice(c->device));
    int rc = zero_state(c);
    if (rc) return rc;
    c->fwd_complete = false;
    KArgs av = c->a, as = c->a;
    const int fwdstore = c->p.mode > 0;
    const int vp1_form = c->p.grad_form == 1 || c->p.grad_form == 2;
    StepCtl ctl = base_ctl(c, 0, 1);
    ctl.mode0_gf2 = fwdstore && c->p.grad_form == 2;
    ctl.seismo = c->p.seismo;
    ctl.probe = 1;
    size_t nev = 0;
    if (c->eprecond)  // grad_obj_psv.c:71-76: the energies start at zero for each shot
        for (int w = 0; w < 2; ++w) HIP_OK(hipMemsetAsync(c->ew[w], 0, (size_t)c->p.nx * c->p.ny * sizeof(float), c->stream));
    HIP_OK(hipEventRecord(c->ev0, c->stream));
    int hin = 1, hin1 = 1, lsnap = c->snap_first, nsnap = 0;
    std::vector<float> snapbuf;
    for (int nt = 1; nt <= nsteps; ++nt) {
        const bool store = fwdstore && nt == hin1;
        const bool sample = c->sample_every > 0 && (nt % c->sample_every) == 0;
        ctl.nt = nt;
        ctl.store_k = (store && c->qtype != 5) ? hin - 1 : -1;
        ctl.write_vp1 = vp1_form && (nt == nsteps || (store && c->qtype == 5));
        hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
        if (sample) {
            e0 = pool_event(c, nev); e1 = pool_event(c, nev + 1); e2 = pool_event(c, nev + 2);
            if (!e0 || !e1 || !e2 || !pool_event(c, nev + 3)) return fail(DN_ERR_HIP, "event creation failed");
            HIP_OK(hipEventRecord(e0, c->stream));
        }
        launch_vel(av, ctl, c->acq, c->tile_v, c->fdoh, 0, c->nt_v, c->stream);
        if (sample) {
            HIP_OK(hipEventRecord(e1, c->stream));
            HIP_OK(hipEventRecord(c->pool[nev + 2], c->stream));
        }
        launch_str(as, ctl, c->acq, c->tile_s, c->fdoh, c->nt_s, c->stream);
        if (sample) {
            HIP_OK(hipEventRecord(c->pool[nev + 3], c->stream));
            nev += 4;
        }
        // the acoustic and SH loops have no moment-tensor source (src/AC/ac.c:155-170, src/SH/sh.c:80-130)
        if (c->qtype == 5 && c->acq.nsrc > 0 && c->a.physics != 2 && c->a.physics != 5)
            launch_msource(c->a, c->acq, nt, c->stream);
        if (store && c->qtype == 5) launch_store(c->a, c->fwd, hin - 1, c->stream);
        if (store && c->eprecond) launch_energy(c->a, c->ew[0], c->stream);  // psv.c:300-303
        if (store) {
            ++hin;
            hin1 += c->p.dtinv;
        }
        if (c->snap_fn && nt == lsnap && nt <= c->snap_last) {
            // psv.c:225-231: the step's wavefields, after sources and the free surface
            const size_t np = (size_t)(c->p.nx + 2 * c->nd) * (c->p.ny + 2 * c->nd);
            snapbuf.resize(4 * np);
            const int ids[4] = {F_VX, F_VY, F_SXX, F_SYY};
            for (int k = 0; k < 4; ++k)
                if ((rc = dn_get_field(c, ids[k], snapbuf.data() + k * np))) return rc;
            c->snap_fn(c->snap_user, nt, ++nsnap, snapbuf.data(), snapbuf.data() + np, snapbuf.data() + 2 * np,
                       snapbuf.data() + 3 * np);
            lsnap += c->snap_inc > 0 ? c->snap_inc : nsteps + 1;
        }
    }
    HIP_OK(hipGetLastError());
    rc = finish_run(c, nev);
    if (rc) return rc;
    c->fwd_complete = fwdstore && nsteps == NT;
    return DN_OK;
}

int dn_get_seismogram(dn_ctx *c, int comp, float *out)
{
    if (!c || !out || comp < 0 || comp > 4) return fail(DN_ERR_ARG, "bad seismogram request");
    if (c->acq.ntr == 0) return DN_OK;
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipStreamSynchronize(c->stream));
    HIP_OK(d2h(c->stream, out, c->acq.sec[comp], (size_t)c->acq.ntr * c->p.nt * sizeof(float)));
    return DN_OK;
}

int dn_calc_residuals(dn_ctx *c, const float *obs_vx, const float *obs_vy, const float *obs_p, int testshot,
                      double *l2)
{
    if (!c) return fail(DN_ERR_ARG, "null context");
    HIP_OK(hipSetDevice(c->device));
    DevAcq &q = c->acq;
    const int QB = c->p.quelltypb;
    const size_t n = (size_t)q.ntr * c->p.nt;
    HIP_OK(hipMemsetAsync(q.l2, 0, sizeof(double), c->stream));
    for (int k = 0; k < 3; ++k)
        if (q.adj[k]) HIP_OK(hipMemsetAsync(q.adj[k], 0, n * sizeof(float), c->stream));
    const float *obs[3] = {obs_vx, obs_vy, obs_p};
    bool use[3];
    float w[3];
    residual_components(c->p.physics, QB, c->res_w, use, w);
    for (int comp = 0; comp < 3; ++comp) {
        if (!use[comp] || !obs[comp] || q.ntr == 0) continue;
        HIP_OK(hipMemcpyAsync(q.obs, obs[comp], n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        launch_residual(q, comp, q.obs, c->p.dt, w[comp], c->stream);
        if (testshot) launch_l2(q, comp, c->stream);
        HIP_OK(hipStreamSynchronize(c->stream));
    }
    HIP_OK(hipGetLastError());
    double v = 0.0;
    HIP_OK(d2h(c->stream, &v, q.l2, sizeof(double)));
    if (l2) *l2 = v;
    return DN_OK;
}

int dn_set_residual_weights(dn_ctx *c, float w_vx, float w_vy)
{
    if (!c) return fail(DN_ERR_ARG, "null context");
    if (!(std::isfinite(w_vx) && std::isfinite(w_vy))) return fail(DN_ERR_ARG, "residual weights must be finite");
    c->res_w[0] = w_vx;
    c->res_w[1] = w_vy;
    return DN_OK;
}

int dn_set_adjoint_sources(dn_ctx *c, const float *dvx, const float *dvy, const float *dp)
{
    if (!c) return fail(DN_ERR_ARG, "null context");
    HIP_OK(hipSetDevice(c->device));
    const size_t n = (size_t)c->acq.ntr * c->p.nt;
    const float *src[3] = {dvx, dvy, dp};
    for (int k = 0; k < 3; ++k) {
        if (!c->acq.adj[k]) continue;
        if (src[k])
            HIP_OK(h2d(c->stream, c->acq.adj[k], src[k], n * sizeof(float)));
        else
            HIP_OK(hipMemsetAsync(c->acq.adj[k], 0, n * sizeof(float), c->stream));
    }
    return DN_OK;
}

int dn_get_adjoint_sources(dn_ctx *c, int comp, float *out)
{
    if (!c || !out || comp < 0 || comp > 2) return fail(DN_ERR_ARG, "bad adjoint-source request");
    if (c->acq.ntr == 0) return DN_OK;
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipStreamSynchronize(c->stream));
    HIP_OK(d2h(c->stream, out, c->acq.adj[comp], (size_t)c->acq.ntr * c->p.nt * sizeof(float)));
    return DN_OK;
}

int dn_run_adjoint(dn_ctx *c, int nsteps)
{
    if (!c) return fail(DN_ERR_ARG, "null context");
    if (c->p.mode <= 0) return fail(DN_ERR_STATE, "adjoint run needs params.mode > 0");
    if (!c->fwd_complete) return fail(DN_ERR_STATE, "adjoint run needs a complete forward run with snapshots");
    const int NT = c->p.nt;
    if (nsteps <= 0 || nsteps > NT) nsteps = NT;
    HIP_OK(hipSetDevice(c->device));
    int rc = zero_state(c);
    if (rc) return rc;
    const size_t g = (size_t)c->p.nx * c->p.ny * sizeof(float);
    HIP_OK(hipMemsetAsync(c->wc, 0, g, c->stream));
    HIP_OK(hipMemsetAsync(c->wcu, 0, g, c->stream));
    HIP_OK(hipMemsetAsync(c->wcr, 0, g, c->stream));
    KArgs av = c->a, as = c->a;
    StepCtl ctl = base_ctl(c, 2, 3);
    ctl.probe = 1;
    size_t nev = 0;
    HIP_OK(hipEventRecord(c->ev0, c->stream));
    int hin = 1;
    for (int nt = 1; nt <= nsteps; ++nt) {
        const bool sample = c->sample_every > 0 && (nt % c->sample_every) == 0;
        const int nfwd = NT - nt + 1;  // forward step this adjoint step meets
        const bool corr = (nfwd - 1) % c->p.dtinv == 0;
        ctl.nt = nt;
        ctl.corr_k = corr ? c->ntdtinv - hin : -1;
        if (sample) {
            for (int k = 0; k < 4; ++k)
                if (!pool_event(c, nev + k)) return fail(DN_ERR_HIP, "event creation failed");
            HIP_OK(hipEventRecord(c->pool[nev], c->stream));
        }
        launch_vel(av, ctl, c->acq, c->tile_v, c->fdoh, 1, c->nt_v, c->stream);
        if (sample) {
            HIP_OK(hipEventRecord(c->pool[nev + 1], c->stream));
            HIP_OK(hipEventRecord(c->pool[nev + 2], c->stream));
        }
        launch_str(as, ctl, c->acq, c->tile_s, c->fdoh, c->nt_s, c->stream);
        if (sample) {
            HIP_OK(hipEventRecord(c->pool[nev + 3], c->stream));
            nev += 4;
        }
        if (corr && c->eprecond == 1) launch_energy(c->a, c->ew[1], c->stream);  // psv.c:380-383
        if (corr) ++hin;
    }
    HIP_OK(hipGetLastError());
    return finish_run(c, nev);
}

int dn_assemble_gradient(dn_ctx *c)
{
    if (!c) return fail(DN_ERR_ARG, "null context");
    if (c->p.mode <= 0) return fail(DN_ERR_STATE, "gradient assembly needs params.mode > 0");
    if (c->p.physics == 3 || c->p.physics == 4)
        return fail(DN_ERR_STATE, "VTI/TTI images are the raw correlation sums (ass_gradVTI assembles nothing): "
                                  "read them with dn_get_gradient 0..2");
    HIP_OK(hipSetDevice(c->device));
    launch_assemble(c->a, c->wc, c->wcu, c->wcr, c->gvp, c->gvs, c->grho, c->p.invmat1, c->a.physics, c->stream);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
    return DN_OK;
}

int dn_get_gradient(dn_ctx *c, int which, float *out)
{
    if (!c || !out || which < 0 || which > 5) return fail(DN_ERR_ARG, "bad gradient request");
    if (c->p.mode <= 0) return fail(DN_ERR_STATE, "no gradient without params.mode > 0");
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipStreamSynchronize(c->stream));
    const int NX = c->p.nx, NY = c->p.ny;
    const size_t n = (size_t)NX * NY;
    if (which >= 3) {
        float *src = which == 3 ? c->gvp : which == 4 ? c->gvs : c->grho;
        HIP_OK(d2h(c->stream, out, src, n * sizeof(float)));
        return DN_OK;
    }
    std::vector<float> tmp(n);
    float *src = which == 0 ? c->wc : which == 1 ? c->wcu : c->wcr;
    HIP_OK(d2h(c->stream, tmp.data(), src, n * sizeof(float)));
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j) out[(size_t)i * NY + j] = tmp[(size_t)j * NX + i];
    return DN_OK;
}

int dn_get_dims(dn_ctx *c, int *nd, int *ntdtinv)
{
    if (!c) return fail(DN_ERR_ARG, "null context");
    if (nd) *nd = c->nd;
    if (ntdtinv) *ntdtinv = c->ntdtinv;
    return DN_OK;
}

static int get_plane_nr(dn_ctx *c, const float *plane, float *out)
{
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipStreamSynchronize(c->stream));
    std::vector<float> h(c->lay.plane());
    HIP_OK(d2h(c->stream, h.data(), plane, h.size() * sizeof(float)));
    const int nd = c->nd, NX = c->p.nx, NY = c->p.ny;
    const int ncol = NX + 2 * nd;
    size_t k = 0;
    for (int j = -nd + 1; j <= NY + nd; ++j)
        for (int i = -nd + 1; i <= NX + nd; ++i, ++k) out[k] = h[c->lay.origin() + c->lay.off(j, i)];
    (void)ncol;
    return DN_OK;
}

int dn_get_field(dn_ctx *c, int id, float *out)
{
    if (!c || !out || id < 0 || id >= F_N) return fail(DN_ERR_ARG, "bad field request");
    if (!c->fld[id]) return fail(DN_ERR_STATE, "field not allocated in this configuration");
    // acoustic contexts keep the pressure in the sxx plane, read as syy too
    int rc = get_plane_nr(c, (id == F_SYY && c->a.physics == 2) ? c->fld[F_SXX] : c->fld[id], out);
    if (rc || !c->a.mirror || (id != F_SYY && id != F_SXY && !(id == F_SXX && c->a.physics == 2))) return rc;
    if (id == F_SXX) id = F_SYY;  // the acoustic pressure: mirrored like syy (surface_acoustic_PML_AC.c:20-25)
    // With a free surface the sweeps mirror the stress ghost rows on the fly;
    // materialise them as the reference's surface routine stores them:
    // syy[1-m] = -syy[1+m], sxy[1-m] = -sxy[m], m = 1..FDORDER/2.
    const int nd = c->nd, ncol = c->p.nx + 2 * nd;
    auto row = [&](int j) { return out + (size_t)(j + nd - 1) * ncol; };
    for (int m = 1; m <= c->fdoh; ++m) {
        float *dst = row(1 - m);
        const float *src = id == F_SYY ? row(1 + m) : row(m);
        for (int k = 0; k < ncol; ++k) dst[k] = -src[k];
    }
    return DN_OK;
}

int dn_get_psi(dn_ctx *c, int id, float *out)
{
    if (!c || !out || id < 0 || id > 8) return fail(DN_ERR_ARG, "bad CPML memory request");
    HIP_OK(hipSetDevice(c->device));
    HIP_OK(hipStreamSynchronize(c->stream));
    const size_t n = id < 5 ? (size_t)c->p.ny * 2 * c->p.fw : (size_t)2 * c->p.fw * c->p.nx;
    HIP_OK(d2h(c->stream, out, c->psi[id], n * sizeof(float)));
    return DN_OK;
}

int dn_get_memvar(dn_ctx *c, int which, int l, float *out)
{
    if (!c || !out || which < 0 || which > 2 || l < 1 || l > c->p.L) return fail(DN_ERR_ARG, "bad memory-variable request");
    const float *base = which == 0 ? c->rmem : which == 1 ? c->pmem : c->qmem;
    return get_plane_nr(c, base + (size_t)(l - 1) * c->lay.plane(), out);
}

int dn_get_pml(dn_ctx *c, int which, float *out)
{
    if (!c || !out || which < 0 || which >= P_NVEC) return fail(DN_ERR_ARG, "bad CPML profile request");
    std::memcpy(out, c->setup.pml[which].data() + 1, sizeof(float) * 2 * c->p.fw);
    return DN_OK;
}

int dn_get_hc(dn_ctx *c, float *out7)
{
    if (!c || !out7) return fail(DN_ERR_ARG, "null argument");
    std::memcpy(out7, c->setup.hc, sizeof(float) * 7);
    return DN_OK;
}

int dn_get_timing(dn_ctx *c, double *run_ms, double *v_kernel_ms, double *s_kernel_ms, int *samples)
{
    if (!c) return fail(DN_ERR_ARG, "null context");
    if (run_ms) *run_ms = c->run_ms;
    if (v_kernel_ms) *v_kernel_ms = c->v_ms;
    if (s_kernel_ms) *s_kernel_ms = c->s_ms;
    if (samples) *samples = c->samples;
    return DN_OK;
}

int dn_get_footprint(dn_ctx *c, size_t *bytes)
{
    if (!c || !bytes) return fail(DN_ERR_ARG, "null argument");
    *bytes = c->bytes;
    return DN_OK;
}

}  // extern "C"
