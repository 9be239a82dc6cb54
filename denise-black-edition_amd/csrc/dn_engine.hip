// Multi-shot engine: many independent shots (contexts) advanced together.
//
// Every time step is one velocity launch and one stress launch per group of
// shots that share FD order, relaxation count, direction and tile shapes.
// The kernels read the time step from one device clock plus the offset k of
// the step in the current run of steps (shot step = clock + k - base, base
// fixed when the shot is attached); a run of steps ends with one launch that
// moves the clock on.  So the same launch sequence serves every run: the
// engine can capture a run of steps in a hipGraph once and replay it.  Shots of different grid size and length share the
// launches; a shot that has reached its NT leaves its blocks idle until the
// host detaches it.  Shots can be attached and detached between calls
// (continuous batching); their state stays in their contexts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "dn_ctx.h"

using namespace dn;

namespace {

struct Job {
    dn_ctx *ctx;
    int phase;  // 0 forward, 1 adjoint
    long base;  // engine clock before its first step
};

struct Group {
    int fdoh, L, sw;
    Tile tv, ts;
    int off_v = 0, nblk_v = 0, off_s = 0, nblk_s = 0;
    int nfull_v = 0, nfull_s = 0;  // full-tile blocks at the head of each table
};


}  // namespace

struct dn_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<Job> jobs;              // running shots
    std::map<dn_ctx *, int> status;     // last known state per shot
    int *d_clock = nullptr;             // device clock (steps run)
    long clock = 0;                     // its value once the queued work is done
    bool inflight = false;              // a launched run not yet waited for
    bool dirty = true;
    std::vector<Group> groups;
    BatchShot *d_shots = nullptr;
    int n_shots = 0, cap_shots = 0;
    BatchBlock *d_blk = nullptr;
    size_t cap_blk = 0;
    int graph_steps = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;

    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }
    ~dn_engine()
    {
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graph();
        if (d_clock) (void)hipFree(d_clock);
        if (d_shots) (void)hipFree(d_shots);
        if (d_blk) (void)hipFree(d_blk);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// group every running shot and upload the launch tables
int rebuild(dn_engine *e)
{
    HIP_OK(hipStreamSynchronize(e->stream));
    e->drop_graph();
    std::map<std::tuple<int, int, int, int, int, int, int, int, int, int, int>, std::vector<int>> by;
    for (int k = 0; k < (int)e->jobs.size(); ++k) {
        const dn_ctx *c = e->jobs[k].ctx;
        const Tile &v = c->tile_v, &s = c->tile_s;
        by[std::make_tuple(c->fdoh, c->p.L, e->jobs[k].phase, v.tbx, v.by, v.ty, v.vw, s.tbx, s.by, s.ty, s.vw)]
            .push_back(k);
    }
    std::vector<BatchShot> shots;
    std::vector<BatchBlock> blk;
    e->groups.clear();
    for (auto &kv : by) {
        Group g;
        const dn_ctx *c0 = e->jobs[kv.second[0]].ctx;
        g.fdoh = c0->fdoh;
        g.L = c0->p.L;
        g.sw = e->jobs[kv.second[0]].phase;
        g.tv = c0->tile_v;
        g.ts = c0->tile_s;
        std::vector<BatchBlock> bv, bs, lv, ls;  // full and lean blocks of each sweep
        for (int k : kv.second) {
            const Job &j = e->jobs[k];
            dn_ctx *c = j.ctx;
            BatchShot b{};
            b.a = c->a;
            const StepCtl base = j.phase == 0 ? base_ctl(c, 0, 1) : base_ctl(c, 2, 3);
            b.cv = base;
            b.cv.seismo = j.phase == 0 ? c->p.seismo : 0;
            b.cv.probe = 1;
            b.cv.mode0_gf2 = j.phase == 0 && c->p.mode > 0 && c->p.grad_form == 2;
            b.cs = b.cv;
            b.cv.ntx = c->tile_v.ntx(c->p.nx);
            b.cv.ntiles = b.cv.ntx * c->tile_v.nty(c->p.ny);
            b.cs.ntx = c->tile_s.ntx(c->p.nx);
            b.cs.ntiles = b.cs.ntx * c->tile_s.nty(c->p.ny);
            b.q = c->acq;
            b.nt_total = c->p.nt;
            b.dtinv = c->p.dtinv;
            b.fwdstore = c->p.mode > 0;
            b.vp1_form = c->p.grad_form == 1 || c->p.grad_form == 2;
            b.sw = j.phase;
            b.ew = j.phase == 0 ? (c->eprecond ? c->ew[0] : nullptr) : (c->eprecond == 1 ? c->ew[1] : nullptr);
            const int idx = (int)shots.size();
            shots.push_back(b);
            const int b0 = (int)j.base;
            const int wv = j.phase == 0 ? 0 : 2, ws = wv + 1;
            for (int t : c->full_h[wv]) bv.push_back(BatchBlock{idx, t, b0, b.nt_total});
            for (int t : c->lean_h[wv]) lv.push_back(BatchBlock{idx, t, b0, b.nt_total});
            for (int t : c->full_h[ws]) bs.push_back(BatchBlock{idx, t, b0, b.nt_total});
            for (int t : c->lean_h[ws]) ls.push_back(BatchBlock{idx, t, b0, b.nt_total});
        }
        g.off_v = (int)blk.size();
        g.nfull_v = (int)bv.size();
        g.nblk_v = (int)(bv.size() + lv.size());
        blk.insert(blk.end(), bv.begin(), bv.end());
        blk.insert(blk.end(), lv.begin(), lv.end());
        g.off_s = (int)blk.size();
        g.nfull_s = (int)bs.size();
        g.nblk_s = (int)(bs.size() + ls.size());
        blk.insert(blk.end(), bs.begin(), bs.end());
        blk.insert(blk.end(), ls.begin(), ls.end());
        e->groups.push_back(g);
    }
    if ((int)shots.size() > e->cap_shots) {
        if (e->d_shots) HIP_OK(hipFree(e->d_shots));
        e->d_shots = nullptr;
        e->cap_shots = std::max<int>((int)shots.size(), 2 * e->cap_shots);
        HIP_OK(hipMalloc((void **)&e->d_shots, sizeof(BatchShot) * e->cap_shots));
    }
    if (blk.size() > e->cap_blk) {
        if (e->d_blk) HIP_OK(hipFree(e->d_blk));
        e->d_blk = nullptr;
        e->cap_blk = std::max(blk.size(), 2 * e->cap_blk);
        HIP_OK(hipMalloc((void **)&e->d_blk, sizeof(BatchBlock) * e->cap_blk));
    }
    if (!shots.empty()) HIP_OK(h2d(e->stream, e->d_shots, shots.data(), sizeof(BatchShot) * shots.size()));
    if (!blk.empty()) HIP_OK(h2d(e->stream, e->d_blk, blk.data(), sizeof(BatchBlock) * blk.size()));
    e->n_shots = (int)shots.size();
    e->dirty = false;
    return DN_OK;
}

// step k of a run of steps
void enqueue_step(dn_engine *e, int k)
{
    for (const Group &g : e->groups)
        launch_batch_vel(e->d_shots, e->d_blk + g.off_v, g.nblk_v, g.nfull_v, e->d_clock, k, g.tv, g.fdoh, g.sw, e->stream);
    for (const Group &g : e->groups)
        launch_batch_str(e->d_shots, e->d_blk + g.off_s, g.nblk_s, g.nfull_s, e->d_clock, k, g.ts, g.fdoh, g.L, e->stream);
}

// a run of n steps, then the clock moves on by n
void enqueue_run(dn_engine *e, int n)
{
    for (int k = 0; k < n; ++k) enqueue_step(e, k);
    launch_batch_clock(e->d_clock, n, e->stream);
}

int capture(dn_engine *e)
{
    HIP_OK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeRelaxed));
    enqueue_run(e, e->graph_steps);
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamEndCapture(e->stream, &g));
    e->graph = g;
    HIP_OK(hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0));
    return DN_OK;
}

// shots past their NT leave the engine
int collect(dn_engine *e)
{
    std::vector<Job> still;
    for (const Job &j : e->jobs) {
        if (e->clock - j.base <= j.ctx->p.nt) {
            still.push_back(j);
            continue;
        }
        int nan_step = 0;
        HIP_OK(d2h(e->stream, &nan_step, j.ctx->acq.nan_step, sizeof(int)));
        if (j.phase == 0) j.ctx->fwd_complete = j.ctx->p.mode > 0;
        e->status[j.ctx] = nan_step ? DN_ERR_UNSTABLE : 0;
        e->dirty = true;
    }
    e->jobs.swap(still);
    return DN_OK;
}

}  // namespace

int dn_engine_create(int device, dn_engine **out)
{
    if (!out) return fail(DN_ERR_ARG, "null output pointer");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(DN_ERR_ARG, "no such HIP device");
    dn_engine *e = new (std::nothrow) dn_engine();
    if (!e) return fail(DN_ERR_OOM, "host allocation failed");
    e->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return fail(DN_ERR_HIP, "stream creation failed");
    }
    const int zero = 0;
    if (hipMalloc((void **)&e->d_clock, sizeof(int)) != hipSuccess ||
        h2d(e->stream, e->d_clock, &zero, sizeof(int)) != hipSuccess) {
        delete e;
        return fail(DN_ERR_OOM, "engine clock");
    }
    *out = e;
    return DN_OK;
}

int dn_engine_destroy(dn_engine *e)
{
    delete e;
    return DN_OK;
}

int dn_engine_set_graph(dn_engine *e, int steps)
{
    if (!e || steps < 0) return fail(DN_ERR_ARG, "bad graph length");
    HIP_OK(hipSetDevice(e->device));
    HIP_OK(hipStreamSynchronize(e->stream));
    e->drop_graph();
    e->graph_steps = steps;
    return DN_OK;
}

int dn_engine_attach(dn_engine *e, dn_ctx *c, int phase)
{
    if (!e || !c || (phase != 0 && phase != 1)) return fail(DN_ERR_ARG, "bad attach request");
    if (c->device != e->device) return fail(DN_ERR_ARG, "context and engine are on different devices");
    if (!c->model_set) return fail(DN_ERR_STATE, "dn_set_model must precede dn_engine_attach");
    if (c->qtype == 5) return fail(DN_ERR_ARG, "moment-tensor shots run through dn_run_forward, not the engine");
    if (c->a.physics != 1) return fail(DN_ERR_ARG, "acoustic, SH and VTI/TTI shots run through dn_run_forward, not the engine");
    for (const Job &j : e->jobs)
        if (j.ctx == c) return fail(DN_ERR_STATE, "context already attached");
    if (phase == 1 && (c->p.mode <= 0 || !c->fwd_complete))
        return fail(DN_ERR_STATE, "adjoint phase needs a finished forward phase with snapshots");
    if (e->clock > (1L << 30)) return fail(DN_ERR_STATE, "engine clock exhausted; create a new engine");
    HIP_OK(hipSetDevice(e->device));
    // tile shapes the multi-shot kernels are compiled for
    int rc;
    if (!batch_tile_supported(c->tile_v, 0) && (rc = dn_set_tile(c, 0, 128, 2, 4, 1))) return rc;
    // (elastic groups default to 64,4,2,2: on the 64-shot mix 50.7 Gcell/s
    // against 49.2 with 64,4,1,4 and 42.2 with 64,4,2,1)
    if ((!batch_str_supported(c->tile_s, c->p.L) || (c->p.L == 0 && c->tile_s.ty == 1 && c->tile_s.vw == 1)) &&
        (rc = c->p.L == 0 ? dn_set_tile(c, 1, 64, 4, 2, 2) : dn_set_tile(c, 1, 64, 4, 1, 1)))
        return rc;
    if ((rc = zero_state(c))) return rc;
    if (phase == 1) {
        const size_t g = (size_t)c->p.nx * c->p.ny * sizeof(float);
        HIP_OK(hipMemsetAsync(c->wc, 0, g, c->stream));
        HIP_OK(hipMemsetAsync(c->wcu, 0, g, c->stream));
        HIP_OK(hipMemsetAsync(c->wcr, 0, g, c->stream));
    } else {
        c->fwd_complete = false;
        // grad_obj_psv.c:71-76: the energies start at zero for each shot
        if (c->eprecond)
            for (int w = 0; w < 2; ++w)
                HIP_OK(hipMemsetAsync(c->ew[w], 0, (size_t)c->p.nx * c->p.ny * sizeof(float), c->stream));
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    e->jobs.push_back(Job{c, phase, e->clock - 1});
    e->status[c] = 1;
    e->dirty = true;
    return DN_OK;
}

int dn_engine_launch(dn_engine *e, int nsteps)
{
    if (!e || nsteps < 0) return fail(DN_ERR_ARG, "bad step request");
    if (e->inflight) return fail(DN_ERR_STATE, "a launched run is still in flight (dn_engine_wait first)");
    HIP_OK(hipSetDevice(e->device));
    int rc;
    const bool changed = e->dirty;
    if (changed && (rc = rebuild(e))) return rc;
    if (e->jobs.empty() || nsteps == 0) return DN_OK;
    // no more steps than the longest remaining shot needs
    long need = 0;
    for (const Job &j : e->jobs) need = std::max<long>(need, j.ctx->p.nt - (e->clock - j.base) + 1);
    int left = (int)std::min<long>(nsteps, need);
    e->clock += left;
    e->inflight = true;
    // a graph is captured only for a set of shots that has already run
    // unchanged once: capture and instantiation cost milliseconds of host
    // time, which a set that changes at every call (a request stream) would
    // pay each time while plain launches overlap the GPU work
    if (e->graph_steps > 0 && left >= e->graph_steps && (e->gexec || !changed)) {
        if (!e->gexec && (rc = capture(e))) return rc;
        for (; left >= e->graph_steps; left -= e->graph_steps) HIP_OK(hipGraphLaunch(e->gexec, e->stream));
    }
    if (left > 0) enqueue_run(e, left);
    HIP_OK(hipGetLastError());
    return DN_OK;
}

int dn_engine_wait(dn_engine *e, int *running)
{
    if (!e) return fail(DN_ERR_ARG, "null engine");
    if (e->inflight) {
        HIP_OK(hipSetDevice(e->device));
        e->inflight = false;
        HIP_OK(hipStreamSynchronize(e->stream));
        int rc;
        if ((rc = collect(e))) return rc;
    }
    if (running) *running = (int)e->jobs.size();
    return DN_OK;
}

int dn_engine_step(dn_engine *e, int nsteps, int *running)
{
    const int rc = dn_engine_launch(e, nsteps);
    return rc ? rc : dn_engine_wait(e, running);
}

int dn_engine_status(dn_engine *e, dn_ctx *c, int *status)
{
    if (!e || !c || !status) return fail(DN_ERR_ARG, "bad status request");
    auto it = e->status.find(c);
    if (it == e->status.end()) return fail(DN_ERR_ARG, "context was never attached to this engine");
    *status = it->second;
    return DN_OK;
}

int dn_engine_forget(dn_engine *e, dn_ctx *c)
{
    if (!e || !c) return fail(DN_ERR_ARG, "bad request");
    for (const Job &j : e->jobs)
        if (j.ctx == c) return fail(DN_ERR_STATE, "context still running");
    e->status.erase(c);
    return DN_OK;
}
