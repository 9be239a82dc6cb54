// Internal declarations shared by the host set-up, the HIP kernels and the
// C-ABI layer of libdnhip.so.  Not part of the public interface.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "denise_hip.h"

namespace dn {

// Device layout of every 2D field: row-major, x fastest, with HX halo columns
// on the left/right and HY halo rows on top/bottom.  Element (j,i) in the
// reference's 1-based indices lives at (j-1+HY)*pitch + (i-1+HX).  HX = 32
// puts column i=1 on a 128-byte boundary; the pitch is a multiple of 64
// floats so every row starts 256-byte aligned.  HY = 8 covers the widest
// stencil reach (FDORDER 12: +-6) and the free-surface ghost rows.
constexpr int HX = 32;
constexpr int HY = 8;
constexpr int MAXL = 8;

struct Layout {
    int nx = 0, ny = 0, pitch = 0, rows = 0;
    void init(int nx_, int ny_) {
        nx = nx_;
        ny = ny_;
        pitch = ((nx + 2 * HX) + 63) / 64 * 64;
        rows = ny + 2 * HY;
    }
    size_t plane() const { return (size_t)pitch * rows; }
    size_t origin() const { return (size_t)HY * pitch + HX; }
    // offset of (j,i) relative to the origin (j=1,i=1)
    long off(int j, int i) const { return (long)(j - 1) * pitch + (i - 1); }
};

// CPML profile vectors in the argument order of src/PML_pro.c:65-68.
enum PmlVec {
    P_DX = 0, P_KX, P_ALX, P_AX, P_BX, P_DXH, P_KXH, P_ALXH, P_AXH, P_BXH,
    P_DY, P_KY, P_ALY, P_AY, P_BY, P_DYH, P_KYH, P_ALYH, P_AYH, P_BYH, P_NVEC
};

// Host-side results of the set-up math (all float, reference operation order).
struct HostSetup {
    float hc[7] = {0, 0, 0, 0, 0, 0, 0};
    std::vector<float> pml[P_NVEC];    // index 0..2FW (0 unused)
};

// Material fields prepared on the host in the device layout.
struct HostMaterial {
    std::vector<float> rip, rjp;           // 1/rho at i+1/2 and j+1/2
    std::vector<float> muipjp, f, g;       // elastic: uipjp, rho vs^2, rho(vp^2-2vs^2)
    // viscoelastic: relaxed moduli and relaxation times; the update rebuilds
    // fipjp, f, g (with DT) and the per-mechanism d, e, dip from them
    std::vector<float> vmu, vts, vpi, vtp, vmuij, vtsij;
    float peta[MAXL + 1] = {0};
    float bip[MAXL + 1] = {0}, cip[MAXL + 1] = {0}, bjm[MAXL + 1] = {0}, cjm[MAXL + 1] = {0};
    std::vector<float> fs_fjm, fs_g, fs_dme; // visco free-surface per column
    float fs_bjm[MAXL + 1] = {0};
    std::vector<float> vp, vs, rho;        // raw model in device layout
    // anisotropic stiffness planes: VTI c11, c13, c33, c44h; TTI d11, d13,
    // d15, d15h, d33, d35, d35h, d55h
    std::vector<float> ani[8];
};

// material planes of the acoustic (physics 2: the modulus g), the
// anisotropic (3 VTI, 4 TTI: the stiffness planes) and the SH runs (5: the
// shear moduli uip, ujp); 0 for isotropic P/SV
inline int aniso_planes(int physics, int L = 0)
{
    // SH: uip, ujp, or with L > 0 fipjp, f, dip[1..L], d[1..L]
    return physics == 2 ? 1 : (physics == 3 ? 4 : (physics == 4 ? 8 : (physics == 5 ? 2 + 2 * L : 0)));
}

// the physics whose sweeps have scalar columns only in both sweeps
inline bool scalar_only(int physics) { return physics == 2 || physics == 5; }

// forward snapshot planes a physics stores: vxp1, vyp1 and three stress (or
// ux, uy, uxy) planes for P/SV, VTI and TTI; vxp1, vyp1 and p (or ux) for
// the acoustic loop (ac.c:222-247); vzp1 and sxz, syz (or uzx, uz) for SH
// (sh.c:150-172)
inline bool snap_plane_used(int physics, int k)
{
    if (physics == 2) return k <= 2;
    if (physics == 5) return k == 0 || k == 2 || k == 3;
    return true;
}
inline int snap_planes(int physics)
{
    int n = 0;
    for (int k = 0; k < 5; ++k) n += snap_plane_used(physics, k);
    return n;
}

// the residual components of calc_res_PSV.c (vx, vy, p by QUELLTYPB) with
// their COMP_WEIGHT factors; calc_res_SH.c:20-44 takes the vz residual
// (component 1) for any QUELLTYPB and no weights
inline void residual_components(int physics, int qb, const float res_w[2], bool use[3], float w[3])
{
    use[0] = qb == 1 || qb == 3 || qb == 5 || qb == 7;
    use[1] = qb == 1 || qb == 2 || qb == 6 || qb == 7;
    use[2] = qb >= 4;
    w[0] = res_w[0];
    w[1] = res_w[1];
    w[2] = 1.0f;
    if (physics == 5) {
        use[0] = use[2] = false;
        use[1] = qb != 0;
        w[1] = 1.0f;
    }
}

// set-up math (dn_setup.cpp)
int holberg(int fdorder, int maxrelerror, float hc[7], std::string &err);
void pml_profiles(const dn_params &p, HostSetup &s);
int check_stability(const dn_params &p, const float hc[7], const float *vp, const float *vs,
                    const float *rho, const float *taus, const float *taup, std::string &err);
void prepare_material(const dn_params &p, const Layout &lay, const float *vp, const float *vs,
                      const float *rho, const float *taus, const float *taup, HostMaterial &m);
int check_stability_ac(const dn_params &p, const float hc[7], const float *vp, const float *rho, std::string &err);
void prepare_material_ac(const dn_params &p, const Layout &lay, const float *vp, const float *rho, HostMaterial &m);
void prepare_material_sh(const dn_params &p, const Layout &lay, const float *vs, const float *rho, const float *taus,
                         HostMaterial &m);
int check_stability_sh_visc(const dn_params &p, const float hc[7], const float *vs, const float *rho,
                            const float *taus, std::string &err);
int check_stability_vti(const dn_params &p, const float hc[7], const float *rho, const float *c11,
                        const float *c13, const float *c33, const float *c44, std::string &err);
void prepare_material_aniso(const dn_params &p, const Layout &lay, const float *rho, const float *c11,
                            const float *c13, const float *c33, const float *c44, const float *theta,
                            HostMaterial &m);

// Kernel argument block (passed by value).
struct KArgs {
    int nx, ny, pitch, fw, fw2;
    int free_surf, boundary, grad_form, L, nt_total;
    // mirror: the velocity sweep forms the free-surface ghost rows of syy and
    // sxy (isotropic P/SV with FREE_SURF); the VTI/TTI loops leave them 0
    int mirror, physics;
    float dt, dh;
    float dhi_el;   // DT/DH  (update_s_elastic, seismo_ssg)
    float dhi_vis;  // 1/DH   (update_s_visc, surfaces)
    float dthalbe;  // DT/2
    float hc[8];
    long plane;
    // wavefields at (1,1)
    float *vx, *vy, *sxx, *syy, *sxy, *vxp1, *vyp1, *ux, *uy, *uxy;
    // velocity material
    const float *rip, *rjp;
    // elastic stress material
    const float *muipjp, *f, *g;
    // viscoelastic material and memory variables (L planes each)
    const float *vmu, *vts, *vpi, *vtp, *vmuij, *vtsij;
    float *rmem, *pmem, *qmem;
    // relaxation times uniform over the grid (the reference's model reader
    // sets Q = 30 everywhere): the update takes them from ts0/tp0/tsij0
    // instead of reading the vts/vtp/vtsij planes
    int tau_const;
    float ts0, tp0, tsij0;
    float peta[MAXL + 1];
    float bip[MAXL + 1], cip[MAXL + 1], bjm[MAXL + 1], cjm[MAXL + 1];
    const float *fs_fjm, *fs_g, *fs_dme;
    float fs_bjm[MAXL + 1];
    // raw model (FWI)
    const float *vp, *vs, *rho;
    // anisotropic stiffness planes (HostMaterial::ani order)
    const float *ani[8];
    // CPML memory and profiles (profiles indexed 1..2FW)
    float *psi_sxx_x, *psi_sxy_x, *psi_vxx, *psi_vyx, *psi_vxxs;
    float *psi_syy_y, *psi_sxy_y, *psi_vyy, *psi_vxy;
    const float *K_x, *a_x, *b_x, *K_xh, *a_xh, *b_xh;
    const float *K_y, *a_y, *b_y, *K_yh, *a_yh, *b_yh;
};

// Per-shot acquisition on the device.
struct DevAcq {
    int nsrc = 0, ntr = 0, nt = 0;
    int *src_ix = nullptr, *src_iy = nullptr, *src_type = nullptr;
    double *src_sin = nullptr, *src_cos = nullptr;
    float *signals = nullptr;             // nsrc x NT
    int *rec_ix = nullptr, *rec_iy = nullptr;
    double *rec_sqlam = nullptr, *rec_sqmu = nullptr;   // seismo_ssg SEISMO 3/4
    float *sec[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};   // ntr x NT
    float *adj[3] = {nullptr, nullptr, nullptr};                     // ntr x NT
    float *obs = nullptr;                 // scratch for observed traces
    double *l2 = nullptr;                 // misfit, then the partial sums of its reduction
    int *nan_step = nullptr;              // device scalar
};

// Work items a sweep applies to individual cells after its own update:
// sources, adjoint sources and receivers, grouped by the tile that owns the
// cell (CSR: off[tile]..off[tile+1] into ent, in source / trace order).
// ent = {lane within tile, row j, kind, index}.
enum { E_FORCE = 0, E_ADJ_FORCE = 1, E_REC_V = 2, E_EXPLOSIVE = 3, E_ADJ_P = 4, E_REC_S = 5 };
struct Ent {
    int lane, row, kind, idx;
};
struct CellList {
    const int *off = nullptr;
    const Ent *ent = nullptr;
    // tiles of the sweep in two disjoint sets, together all tiles: lean
    // (interior: no point, no boundary work) and full (the rest)
    const int *lean = nullptr, *full = nullptr;
    int nlean = 0, nfull = 0;
    // entries of a tile are sorted by lane; for a tile with entries,
    // lane_off[slot[tile] * (width + 1) + l] is its first entry of lane >= l
    // (absolute index), so a thread visits only its own columns' entries
    const int *slot = nullptr, *lane_off = nullptr;
};

// Tile shape of the two sweeps: blocks of tbx x by threads, each thread
// marching ty consecutive rows of vw adjacent columns; a tile is
// (tbx*vw) x (by*ty) cells.
struct Tile {
    int tbx = 64, by = 4, ty = 1, vw = 1;
    int width() const { return tbx * vw; }
    int rows() const { return by * ty; }
    int ntx(int nx) const { return (nx + width() - 1) / width(); }
    int nty(int ny) const { return (ny + rows() - 1) / rows(); }
};

// The per-step part of the sweep control (time step, snapshot slots,
// vxp1/vyp1 output), passed separately so that a kernel can derive it from a
// device-side step counter.
struct StepDyn {
    int nt, store_k, corr_k, write_vp1;
};

// Per-step control of the fused sweeps.
struct StepCtl {
    int nt;          // time step (1..NT)
    int store_k;     // snapshot slot written this step by the sweeps, -1 none
    int corr_k;      // snapshot slot correlated this step, -1 none
    int write_vp1;   // forward: also write the vxp1/vyp1 planes
    int seismo;      // receiver sampling mode this step (0 = none)
    int mode0_gf2;   // forward with GRAD_FORM 2: write ux/uy/uxy
    int probe;       // NaN probe of vy[NY/2][NX/2] at the start of the step
    int quelltypb;
    int invmat1;
    int ntx, ntiles; // tile grid
    CellList vlist, slist;
    float *fwd[5];   // snapshots (compact NX*NY per slot)
    float *wc, *wcu, *wcr;   // correlation accumulators (compact, [j][i])
};

// One shot of a multi-shot launch: its sweep arguments and the static sweep
// control of both sweeps (cell lists, tile grids); the per-step fields are
// derived in the kernel from the engine clock.
struct BatchShot {
    KArgs a;
    StepCtl cv, cs;
    DevAcq q;
    int nt_total, dtinv, fwdstore, vp1_form, sw;
    // energy weights (EPRECOND, src/eprecond.c:9-18): vx^2+vy^2 summed at
    // the store steps (forward, Ws) or correlation steps (adjoint with
    // EPRECOND 1, Wr) into this NX x NY plane, or nullptr
    float *ew;
};

// One block of a multi-shot launch: its shot, the tile it sweeps, and what
// decides whether the shot runs at this step (time step = clock + k - base,
// running while 1 <= step <= nt_total), so that a block can leave before
// it touches the shot's arguments.
struct alignas(16) BatchBlock {
    int shot, tile, base, nt_total;
};

// multi-shot launchers (dn_batch_*.hip).  clock: the engine's device clock;
// k: offset of this step from it (runs of steps are enqueued, or captured
// in a graph, with k = 0, 1, ... and the clock moved on once at the end)
bool batch_tile_supported(const Tile &t, int sweep);
// stress-sweep shapes of the multi-shot kernels for a group with L relaxation mechanisms
bool batch_str_supported(const Tile &t, int L);
// the first nfull blocks of blk are full tiles, the rest lean tiles
void launch_batch_vel(const BatchShot *shots, const BatchBlock *blk, int nblocks, int nfull, const int *clock,
                      int k, const Tile &t, int fdoh, int sw, void *stream);
void launch_batch_str(const BatchShot *shots, const BatchBlock *blk, int nblocks, int nfull, const int *clock,
                      int k, const Tile &t, int fdoh, int L, void *stream);
void launch_batch_clock(int *clock, int add, void *stream);

// kernel launchers (dn_kernels.hip); all enqueue on `stream`
// ntl: streaming (nontemporal) access to the planes a step touches once
// (material / coefficient loads, memory-variable stores)
void launch_vel(const KArgs &a, const StepCtl &c, const DevAcq &q, const Tile &t, int fdoh, int sw, int ntl,
                void *stream);
void launch_str(const KArgs &a, const StepCtl &c, const DevAcq &q, const Tile &t, int fdoh, int ntl, void *stream);
void launch_msource(const KArgs &a, const DevAcq &q, int nt, void *stream);
void launch_energy(const KArgs &a, float *W, void *stream);
void launch_store(const KArgs &a, float *const fwd[5], long k, void *stream);
void launch_assemble(const KArgs &a, const float *wc, const float *wcu, const float *wcr, float *gvp,
                     float *gvs, float *grho, int invmat1, int physics, void *stream);
void launch_residual(const DevAcq &q, int comp, const float *obs, float dt, float w, void *stream);
void launch_l2(const DevAcq &q, int comp, void *stream);
int l2_scratch_doubles();  // doubles behind DevAcq::l2
bool tile_supported(const Tile &t);
// compiled tile shapes (returns their count; fills at most max entries)
int tile_list(Tile *out, int max);

}  // namespace dn
