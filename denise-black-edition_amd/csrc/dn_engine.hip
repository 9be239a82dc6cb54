// Multi-shot engine: many independent shots (contexts) advanced together.
//
// Every time step is one velocity launch and one stress launch per group of
// shots that share FD order, relaxation count, direction and tile shapes,
// plus one launch that moves every shot's device step counter on.  Because
// the kernels read the step from the counter, the same launch sequence
// serves every step: the engine can capture a run of steps in a hipGraph
// once and replay it.  Shots of different grid size and length share the
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
    int slot;   // index of its step counter
};

struct Group {
    int fdoh, L, sw;
    Tile tv, ts;
    int off_v = 0, nblk_v = 0, off_s = 0, nblk_s = 0;
};

}  // namespace

struct dn_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<Job> jobs;              // running shots
    std::map<dn_ctx *, int> status;     // last known state per shot
    int *d_steps = nullptr;             // step counters, one per slot
    int cap = 0;
    std::vector<int> free_slots;
    bool dirty = true;
    std::vector<Group> groups;
    BatchShot *d_shots = nullptr;
    int n_shots = 0, cap_shots = 0;
    int *d_blk = nullptr;
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
        if (d_steps) (void)hipFree(d_steps);
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
    std::vector<int> blk;
    e->groups.clear();
    for (auto &kv : by) {
        Group g;
        const dn_ctx *c0 = e->jobs[kv.second[0]].ctx;
        g.fdoh = c0->fdoh;
        g.L = c0->p.L;
        g.sw = e->jobs[kv.second[0]].phase;
        g.tv = c0->tile_v;
        g.ts = c0->tile_s;
        std::vector<int> bv, bs;
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
            b.step = e->d_steps + j.slot;
            b.blk0_v = (int)bv.size();
            b.blk0_s = (int)bs.size();
            b.nt_total = c->p.nt;
            b.dtinv = c->p.dtinv;
            b.fwdstore = c->p.mode > 0;
            b.vp1_form = c->p.grad_form == 1 || c->p.grad_form == 2;
            b.sw = j.phase;
            b.nt_v = c->nt_v;
            b.nt_s = c->nt_s;
            const int idx = (int)shots.size();
            shots.push_back(b);
            bv.insert(bv.end(), (size_t)b.cv.ntiles, idx);
            bs.insert(bs.end(), (size_t)b.cs.ntiles, idx);
        }
        g.off_v = (int)blk.size();
        g.nblk_v = (int)bv.size();
        blk.insert(blk.end(), bv.begin(), bv.end());
        g.off_s = (int)blk.size();
        g.nblk_s = (int)bs.size();
        blk.insert(blk.end(), bs.begin(), bs.end());
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
        HIP_OK(hipMalloc((void **)&e->d_blk, sizeof(int) * e->cap_blk));
    }
    if (!shots.empty()) HIP_OK(hipMemcpy(e->d_shots, shots.data(), sizeof(BatchShot) * shots.size(), hipMemcpyHostToDevice));
    if (!blk.empty()) HIP_OK(hipMemcpy(e->d_blk, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice));
    e->n_shots = (int)shots.size();
    e->dirty = false;
    return DN_OK;
}

void enqueue_step(dn_engine *e)
{
    for (const Group &g : e->groups)
        launch_batch_vel(e->d_shots, e->d_blk + g.off_v, g.nblk_v, g.tv, g.fdoh, g.sw, e->stream);
    for (const Group &g : e->groups)
        launch_batch_str(e->d_shots, e->d_blk + g.off_s, g.nblk_s, g.ts, g.fdoh, g.L, e->stream);
    launch_batch_advance(e->d_shots, e->n_shots, e->stream);
}

int capture(dn_engine *e)
{
    HIP_OK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < e->graph_steps; ++k) enqueue_step(e);
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamEndCapture(e->stream, &g));
    e->graph = g;
    HIP_OK(hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0));
    return DN_OK;
}

// shots whose counters passed their NT leave the engine
int collect(dn_engine *e)
{
    if (e->jobs.empty()) return DN_OK;
    std::vector<int> steps((size_t)e->cap);
    HIP_OK(hipMemcpy(steps.data(), e->d_steps, sizeof(int) * e->cap, hipMemcpyDeviceToHost));
    std::vector<Job> still;
    for (const Job &j : e->jobs) {
        if (steps[(size_t)j.slot] <= j.ctx->p.nt) {
            still.push_back(j);
            continue;
        }
        int nan_step = 0;
        HIP_OK(hipMemcpy(&nan_step, j.ctx->acq.nan_step, sizeof(int), hipMemcpyDeviceToHost));
        if (j.phase == 0) j.ctx->fwd_complete = j.ctx->p.mode > 0;
        e->status[j.ctx] = nan_step ? DN_ERR_UNSTABLE : 0;
        e->free_slots.push_back(j.slot);
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
    e->cap = 4096;
    if (hipMalloc((void **)&e->d_steps, sizeof(int) * e->cap) != hipSuccess) {
        delete e;
        return fail(DN_ERR_OOM, "step counters");
    }
    for (int k = e->cap - 1; k >= 0; --k) e->free_slots.push_back(k);
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
    for (const Job &j : e->jobs)
        if (j.ctx == c) return fail(DN_ERR_STATE, "context already attached");
    if (phase == 1 && (c->p.mode <= 0 || !c->fwd_complete))
        return fail(DN_ERR_STATE, "adjoint phase needs a finished forward phase with snapshots");
    if (e->free_slots.empty()) return fail(DN_ERR_ARG, "engine full");
    HIP_OK(hipSetDevice(e->device));
    // tile shapes the multi-shot kernels are compiled for
    int rc;
    if (!batch_tile_supported(c->tile_v, 0) && (rc = dn_set_tile(c, 0, 128, 2, 4, 1))) return rc;
    if (!batch_tile_supported(c->tile_s, 1) && (rc = dn_set_tile(c, 1, 64, 4, 1, c->p.L == 0 ? 4 : 1))) return rc;
    if ((rc = zero_state(c))) return rc;
    if (phase == 1) {
        const size_t g = (size_t)c->p.nx * c->p.ny * sizeof(float);
        HIP_OK(hipMemsetAsync(c->wc, 0, g, c->stream));
        HIP_OK(hipMemsetAsync(c->wcu, 0, g, c->stream));
        HIP_OK(hipMemsetAsync(c->wcr, 0, g, c->stream));
    } else {
        c->fwd_complete = false;
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    const int slot = e->free_slots.back();
    e->free_slots.pop_back();
    const int one = 1;
    HIP_OK(hipMemcpy(e->d_steps + slot, &one, sizeof(int), hipMemcpyHostToDevice));
    e->jobs.push_back(Job{c, phase, slot});
    e->status[c] = 1;
    e->dirty = true;
    return DN_OK;
}

int dn_engine_step(dn_engine *e, int nsteps, int *running)
{
    if (!e || nsteps < 0) return fail(DN_ERR_ARG, "bad step request");
    HIP_OK(hipSetDevice(e->device));
    int rc;
    if (e->dirty && (rc = rebuild(e))) return rc;
    if (!e->jobs.empty() && nsteps > 0) {
        int left = nsteps;
        if (e->graph_steps > 0 && nsteps >= e->graph_steps) {
            if (!e->gexec && (rc = capture(e))) return rc;
            for (; left >= e->graph_steps; left -= e->graph_steps) HIP_OK(hipGraphLaunch(e->gexec, e->stream));
        }
        for (; left > 0; --left) enqueue_step(e);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(e->stream));
        if ((rc = collect(e))) return rc;
    }
    if (running) *running = (int)e->jobs.size();
    return DN_OK;
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
