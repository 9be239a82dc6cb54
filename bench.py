#!/usr/bin/env python3
"""Benchmark of the P/SV hot path on one MI355X.

A "step" here is one whole shot: NT time steps of the viscoelastic P/SV
update over the 4000 x 1200 grid (the single-shot workload the metric is
quoted on: DH=5 m, FD8 Holberg, DT=0.4 ms, NT=6000, L=2 relaxation
mechanisms with Q=50, CPML 20 on three sides plus free surface, Ricker
15 Hz, 800 receivers), inputs resident in HBM before timing starts.

value  = cell-updates per second over the K timed shots (Gcell-updates/s).
roofline = algorithmic HBM bytes of the dominant kernel (the viscoelastic
           stress/memory-variable update: 4*(14+6L) = 104 B per cell at L=2,
           92 B with the uniform relaxation times of this workload)
           divided by its average launch time from HIP events on the
           context stream, against the 8 TB/s MI355X HBM peak.
cpu_baseline = the FP32 CPU oracle (the reference algorithm restated in C,
           OpenMP over the host cores) on the first steps of the same shot.
The elastic FD8 4000x1200 shot (whose step should reach 60% of the HBM
roofline) is reported alongside in "elastic"; 64 uneven shots through the
multi-shot engine in "batched"; a request stream through the server in
"serving"; the elastic FWI gradient job of 2000x600, NT=4000 (forward with
the snapshot store, adjoint with the correlation) in "gradient"; the
drop-in executable over the shipped Marmousi example's 100 shots, with the
multi-shot engine and shot by shot, in "executable"; a VTI and a TTI shot of
the same 4000 x 1200 grid in "detail.aniso".
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "denise-black-edition_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "Gcell-updates/s and %HBM roofline per shot; batched shots/s + p50/p99"


def bytes_per_cell(L, tau_uniform=False):
    """Algorithmic HBM bytes per cell-update of each sweep, the per-unit
    figures of the path's byte accounting: velocity 36 B; stress elastic
    44 B, viscoelastic 4*(14+6L) B (104 B at L=2).  Third value: the bytes
    the stress kernel moves -- with uniform relaxation times it does not read
    the three relaxation-time planes, 12 B fewer."""
    vel = 36
    stress = 44 if L == 0 else 4 * (14 + 6 * L)
    moved = stress - (12 if L > 0 and tau_uniform else 0)
    return vel, stress, moved


def workload(L, nt=6000, nrec=800, seed=11):
    import dnhip
    import synth
    nx, ny, dh, dt, fc = 4000, 1200, 5.0, 4.0e-4, 15.0
    vp, vs, rho = synth.layered_model(nx, ny, dh, seed=seed)
    taus = taup = None
    if L:
        taus = np.full((nx, ny), 2.0 / 50.0, np.float32)
        taup = taus.copy()
    p = dnhip.Params(nx=nx, ny=ny, nt=nt, dh=dh, dt=dt, fdorder=8, maxrelerror=1, fw=20,
                     free_surf=1, L=L, seismo=1, ts=float(np.float32(1.0 / fc)),
                     damping=1500.0, fpml=fc, fl=[3.2, 32.0][:L] if L else [])
    src = synth.source_table(nx * dh / 2, 2 * dh, dh, fc=fc)
    sig = synth.ricker(nt, dt, fc)[None, :]
    rec = synth.receiver_line(2 * dh, 25.0, nrec, 2 * dh, dh)
    return p, dict(vp=vp, vs=vs, rho=rho, taus=taus, taup=taup), src, sig, rec


def run_gpu(L, steps, warmup, sample_every=20, nt=6000, tune_steps=40, tiles=None, policy=None):
    import dnhip
    p, m, src, sig, rec = workload(L, nt=nt)
    ctx = dnhip.Context(p)
    ctx.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    ctx.set_sources(src, sig)
    ctx.set_receivers(rec)
    if tiles:
        ctx.set_tile(*tiles[0], sweep=0)
        ctx.set_tile(*tiles[1], sweep=1)
    elif tune_steps > 0:
        ctx.autotune(tune_steps)
    if policy is not None:
        ctx.set_store_hint(policy[0], sweep=0)
        ctx.set_store_hint(policy[1], sweep=1)
    tile = [list(ctx.tile(0)), list(ctx.tile(1))]
    cache_policy = [ctx.store_hint(0), ctx.store_hint(1)]
    for _ in range(warmup):
        ctx.forward()
    ctx.set_kernel_sampling(sample_every)
    walls, v_ms, s_ms, dev_ms = [], [], [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        ctx.forward()
        walls.append(time.perf_counter() - t0)
        tm = ctx.timing()
        v_ms.append(tm["v_kernel_ms"])
        s_ms.append(tm["s_kernel_ms"])
        dev_ms.append(tm["run_ms"])
    seis = ctx.seismogram(0)
    assert np.isfinite(seis).all()
    cells = p.nx * p.ny
    total = float(sum(walls))
    tau_uniform = L > 0 and all(np.all(m[k] == m[k].flat[0]) for k in ("taus", "taup"))
    vel_b, str_b, str_moved = bytes_per_cell(L, tau_uniform)
    s_avg = float(np.mean(s_ms))
    v_avg = float(np.mean(v_ms))
    out = {
        "cells_per_step": cells, "nt": p.nt,
        "gcells_per_s": cells * p.nt * steps / total / 1e9,
        "ms_per_shot": total / steps * 1e3,
        "device_ms_per_shot": float(np.mean(dev_ms)),
        "v_kernel_ms": v_avg, "s_kernel_ms": s_avg,
        "v_kernel_GBs": vel_b * cells / (v_avg * 1e-3) / 1e9 if v_avg > 0 else None,
        "s_kernel_GBs": str_b * cells / (s_avg * 1e-3) / 1e9 if s_avg > 0 else None,
        "step_GBs": (vel_b + str_b) * cells * p.nt * steps / total / 1e9,
        "stress_bytes_per_cell": str_b, "vel_bytes_per_cell": vel_b, "tau_uniform": bool(tau_uniform),
        "stress_moved_bytes_per_cell": str_moved, "cache_policy": cache_policy,
        "s_kernel_moved_GBs": str_moved * cells / (s_avg * 1e-3) / 1e9 if s_avg > 0 else None,
        "footprint_GB": ctx.footprint() / 1e9,
        "tile": {"velocity": tile[0], "stress": tile[1]},
    }
    ctx.close()
    return out


def run_batched(nshots, graph_steps=32):
    """shots/s of the multi-shot engine vs the same shots one after another
    with the plain per-shot solver (dn_run_forward), on prepared contexts
    (models, sources, receivers resident in HBM before timing)."""
    import dnhip
    from workloads import batch_workload
    shots = batch_workload(nshots)
    cells = sum(p.nx * p.ny * p.nt for p, *_ in shots)

    def prepare():
        ctxs = []
        for p, (vp, vs, rho), src, sig, rec in shots:
            g = dnhip.Context(p)
            g.set_model(vp, vs, rho)
            g.set_sources(src, sig)
            g.set_receivers(rec)
            ctxs.append(g)
        return ctxs

    ctxs = prepare()
    ctxs[0].forward(50)                   # load the kernels
    t0 = time.perf_counter()
    for g in ctxs:
        g.forward()
    t_seq = time.perf_counter() - t0
    ref = [g.seismogram(0) for g in ctxs[:4]]
    for g in ctxs:
        g.close()
    ctxs = prepare()
    eng = dnhip.Engine()
    eng.set_graph(graph_steps)
    t0 = time.perf_counter()
    for g in ctxs:
        eng.attach(g)
    eng.run(chunk=256)
    t_bat = time.perf_counter() - t0
    same = all(np.array_equal(r, g.seismogram(0)) for r, g in zip(ref, ctxs[:4]))
    for g in ctxs:
        g.close()
    eng.close()
    return {
        "shots": nshots, "cell_updates": cells,
        "sequential_s": t_seq, "batched_s": t_bat,
        "sequential_shots_per_s": nshots / t_seq, "batched_shots_per_s": nshots / t_bat,
        "speedup": t_seq / t_bat,
        "batched_gcells_per_s": cells / t_bat / 1e9, "sequential_gcells_per_s": cells / t_seq / 1e9,
        "graph_steps": graph_steps, "results_equal_single_shot": bool(same),
    }


def run_serving(nreq, qps, graph_steps=32, saturate_qps=0.0):
    """p50/p99 latency and throughput of the request server, and the same
    requests one at a time with the plain solver (no overlap).  With
    saturate_qps, the server also runs the same requests arriving that fast
    (its throughput when requests queue up; the plain solver is already
    saturated at the base rate)."""
    import serve
    from workloads import serving_workload
    reqs = serving_workload(nreq, qps)
    srv = serve.Server(graph_steps=graph_steps)
    stats = srv.serve(reqs)
    srv.close()
    if saturate_qps > 0:
        burst = serving_workload(nreq, saturate_qps)
        srv = serve.Server(graph_steps=graph_steps)
        sat = srv.serve(burst)
        srv.close()
        stats["saturated"] = {k: sat[k] for k in ("shots_per_s", "gcells_per_s", "latency_p50_s", "latency_p99_s",
                                                  "wall_s", "completed")}
        stats["saturated"]["qps"] = saturate_qps
    # the plain one-at-a-time solver on the same stream: each request starts
    # when it has arrived and the previous one is done
    t_free, lat, t0 = 0.0, [], time.perf_counter()
    done = 0
    for r in sorted(reqs, key=lambda q: q.arrival):
        if r.error:
            continue
        now = time.perf_counter() - t0
        if now < r.arrival:
            time.sleep(r.arrival - now)
        serve.solve_alone(r)
        lat.append(time.perf_counter() - t0 - r.arrival)
        done += 1
    wall = time.perf_counter() - t0
    stats["plain"] = {"shots_per_s": done / wall, "latency_p50_s": float(np.percentile(lat, 50)),
                      "latency_p99_s": float(np.percentile(lat, 99)), "wall_s": wall}
    stats["qps"] = qps
    return stats


def gradient_workload(nt=4000):
    """The FWI gradient job of one shot: elastic P/SV 2000 x 600, DH = 10 m,
    DT = 1 ms, NT = 4000, FD8, CPML 20 + free surface, forward wavefield
    stored every 2nd step (5 fields x 2000 snapshots x 1.2 M cells = 48 GB
    in HBM), Ricker 10 Hz at the centre 20 m deep, 999 receivers every 20 m
    recording vx and vy.  Observed data: the same shot on the true model
    (smooth background + 10% Gaussian anomaly of radius 300 m); the
    gradient is taken at the background."""
    import dnhip
    import synth
    nx, ny, dh, dt, fc = 2000, 600, 10.0, 1e-3, 10.0
    bg = synth.layered_model(nx, ny, dh, seed=3)
    true = synth.anomaly(*bg, dh, 10000.0, 3000.0, radius=300.0, amp=0.10)
    common = dict(nx=nx, ny=ny, nt=nt, dh=dh, dt=dt, fdorder=8, fw=20, free_surf=1, seismo=1, dtinv=2,
                  ts=float(np.float32(1.0 / fc)))
    src = synth.source_table(nx * dh / 2, 20.0, dh, fc=fc)
    sig = synth.ricker(nt, dt, fc)[None, :]
    rec = synth.receiver_line(20.0, 20.0, 999, 20.0, dh)
    return common, bg, true, src, sig, rec


def run_gradient_other(physics, nt=4000, jobs=2, tune_steps=40):
    """The gradient job of the acoustic (PHYSICS 2, FWI_AC.c) or SH (5,
    FWI_SH.c) inversion on the elastic job's grid and shot (2000 x 600,
    NT 4000, snapshots every 2nd step: 3 planes per stored step), vp/rho or
    vs/rho of the same background and true models: forward with the store,
    residuals (vx/vy, or vz), adjoint with the correlations, assembly;
    whole-job Gcell-updates/s (2 NX NY NT per job) over `jobs` timed jobs."""
    import dnhip
    common, bg, true, src, sig, rec = gradient_workload(nt)
    pick = (lambda m: (m[0], None, m[2])) if physics == 2 else (lambda m: (None, m[1], m[2]))
    fwd = dnhip.Context(dnhip.Params(mode=0, physics=physics, **common))
    fwd.set_model(*pick(true))
    fwd.set_sources(src, sig)
    fwd.set_receivers(rec)
    fwd.forward()
    obs = (fwd.seismogram(0), fwd.seismogram(1))
    fwd.close()
    g = dnhip.Context(dnhip.Params(mode=1, physics=physics, quelltypb=1, **common))
    g.set_model(*pick(bg))
    g.set_sources(src, sig)
    g.set_receivers(rec)
    if tune_steps > 0:
        g.autotune(tune_steps)

    def job():
        g.forward()
        l2 = g.residuals(obs[0] if physics == 2 else None, obs[1], None, 1)
        g.adjoint()
        g.assemble()
        return l2

    job()
    t0 = time.perf_counter()
    for _ in range(jobs):
        l2 = job()
    wall = (time.perf_counter() - t0) / jobs
    k = 4 if physics == 5 else 3
    assert np.isfinite(l2) and np.abs(g.gradient(k)).max() > 0
    out = {"workload": f"{'acoustic' if physics == 2 else 'SH'} gradient job 2000x600, FD8, NT={nt}, DTINV 2",
           "s_per_job": wall, "gcells_per_s": 2.0 * common["nx"] * common["ny"] * nt / wall / 1e9,
           "tile": {"velocity": list(g.tile(0)), "stress": list(g.tile(1))}, "misfit": l2}
    g.close()
    return out


def run_fwi_iterations(g, model, obs, iters, nlbfgs=20):
    """FWI iterations of the gradient job's shot (the only, and test, shot)
    through the product path: gradient on the device with energy
    preconditioning (EPRECOND 3: device energy sums, host weights), the
    horizontal gradient taper (rows 3..10 ramp, (j*DH)^1 below), L-BFGS
    direction and model update on the host (libdnio), step-length search
    with device forward runs of the test shot; EPS_SCALE 0.01, SCALEFAC 2,
    STEPMAX 6, no limits, MIN_ITER above the iteration count (a search that
    finds no lower misfit still applies its smallest trial step,
    FWI_PSV.c:698-708).  Seconds per part and iteration, and the misfit."""
    import dnfwi
    vp, vs, rho = (a.copy() for a in model)
    nx, ny = vp.shape
    lb = dnfwi.LBFGS(nx, ny, nlbfgs)
    par = dnfwi.update_par(vpupperlim=1e9, vplowerlim=0.0, vsupperlim=1e9, vslowerlim=0.0, rhoupperlim=1e9,
                           rholowerlim=0.0, scalerho=0.5)
    state = dnfwi.LineSearch()
    c = (dnfwi.model_mean(vp), dnfwi.model_mean(vs), dnfwi.model_mean(rho))
    g.set_energy_weights(3)
    dh = float(g.p.dh)
    coeff = dnfwi.taper_hor(nx, ny, dh, 3, 10, 1.0)
    n_fwd = [0]

    def misfit(tv, ts, tr, itest):
        g.set_model(tv, ts, tr)
        g.forward()
        n_fwd[0] += 1
        return g.residuals(obs[0], obs[1], None, 1)

    g.set_kernel_sampling(0)
    out = []
    for it in range(1, iters + 1):
        t0 = time.perf_counter()
        g.set_model(vp, vs, rho)
        g.forward()
        l2 = g.residuals(obs[0], obs[1], None, 1)
        g.adjoint()
        g.assemble()
        grads = [g.gradient(3 + k) for k in range(3)]
        grads = [(a * coeff).astype(np.float32)
                 for a in dnfwi.energy_precond(3, dh, g.energy(0), None, *c, *grads)[1:]]
        t1 = time.perf_counter()
        d = lb.direction(it, *c, *grads, vp, vs, rho)
        t2 = time.perf_counter()
        n_fwd[0] = 0
        ls = dnfwi.step_length(par, it, 0.01, 2.0, 6, l2, vp, vs, rho, *d, misfit, state=state)
        t3 = time.perf_counter()
        vp, vs, rho, _ = dnfwi.model_update(par, it, ls.eps_scale, vp, vs, rho, *d)
        t4 = time.perf_counter()
        out.append({"iteration": it, "misfit": l2, "step": ls.eps_scale, "search_ok": bool(ls.step1),
                    "test_forward_runs": n_fwd[0],
                    "gradient_s": t1 - t0, "lbfgs_direction_s": t2 - t1, "step_search_s": t3 - t2,
                    "model_update_s": t4 - t3, "total_s": t4 - t0})
    lb.close()
    return out


def run_gradient(steps, nt=4000, tune_steps=40, sample_every=7, tiles=None, fwi_iters=0):
    """Time of the gradient job's two phases on resident inputs: the forward
    run that stores the snapshots, and the adjoint run whose sweeps
    correlate with them (residuals and assembly timed separately).  Cell
    updates: NX*NY*NT per phase, 2*NX*NY*NT per job.  Algorithmic bytes per
    cell-update: forward 80 + 20/DTINV (5 snapshot fields written every
    DTINV steps), adjoint 80 + 16 + 76/DTINV (the time integral of the
    adjoint velocity, and snapshot + adjoint + model reads and three
    accumulator updates every DTINV steps).  Kernel averages are sampled
    every `sample_every`-th step; an odd interval meets storing and
    non-storing steps alike; they come from one more job after the timed
    ones."""
    import dnhip
    common, bg, true, src, sig, rec = gradient_workload(nt)
    fwd = dnhip.Context(dnhip.Params(mode=0, **common))
    fwd.set_model(*true)
    fwd.set_sources(src, sig)
    fwd.set_receivers(rec)
    fwd.forward()
    obs = (fwd.seismogram(0), fwd.seismogram(1))
    fwd.close()
    g = dnhip.Context(dnhip.Params(mode=1, **common))
    g.set_model(*bg)
    g.set_sources(src, sig)
    g.set_receivers(rec)
    if tiles:
        g.set_tile(*tiles[0], sweep=0)
        g.set_tile(*tiles[1], sweep=1)
    elif tune_steps > 0:
        g.autotune(tune_steps)
    tile = [list(g.tile(0)), list(g.tile(1))]
    t_f, t_r, t_a, t_g, dev_f, dev_a = [], [], [], [], [], []
    for rep in range(steps + 2):
        # job 0: warm-up; jobs 1..steps: timed; last job: kernel averages by
        # HIP events (kept out of the timed jobs, whose time they would add to)
        sampled = rep == steps + 1
        g.set_kernel_sampling(sample_every if sampled else 0)
        t0 = time.perf_counter()
        g.forward()
        t1 = time.perf_counter()
        tf = g.timing()
        l2 = g.residuals(obs[0], obs[1], None, 1)
        t2 = time.perf_counter()
        g.adjoint()
        t3 = time.perf_counter()
        ta = g.timing()
        g.assemble()
        grad = g.gradient(3)
        t4 = time.perf_counter()
        if sampled:
            kf = (tf["v_kernel_ms"], tf["s_kernel_ms"])
            ka = (ta["v_kernel_ms"], ta["s_kernel_ms"])
        elif rep > 0:
            t_f.append(t1 - t0)
            t_r.append(t2 - t1)
            t_a.append(t3 - t2)
            t_g.append(t4 - t3)
            dev_f.append(tf["run_ms"])
            dev_a.append(ta["run_ms"])
    assert np.isfinite(grad).all() and np.isfinite(l2)
    if nt == 4000:      # a short profiling run need not reach the receivers
        assert np.abs(grad).max() > 0 and l2 > 0
    footprint = g.footprint()
    fwi = run_fwi_iterations(g, bg, obs, fwi_iters) if fwi_iters > 0 else None
    g.close()
    cells = common["nx"] * common["ny"]
    cu = cells * nt
    dtinv = common["dtinv"]
    b_fwd, b_adj = 80 + 20 / dtinv, 80 + 16 + 76 / dtinv
    f_s, a_s = float(np.mean(t_f)), float(np.mean(t_a))
    job_s = float(np.mean(np.array(t_f) + np.array(t_r) + np.array(t_a) + np.array(t_g)))
    return {
        "workload": "elastic FWI gradient 2000x600, DH=10 m, FD8, NT=4000, snapshots every 2nd step "
                    "(48 GB), 999 receivers (vx, vy)",
        "jobs": steps, "tile": {"velocity": tile[0], "stress": tile[1]}, "footprint_GB": footprint / 1e9,
        "job_s": job_s, "gcells_per_s": 2 * cu / job_s / 1e9,
        "forward_s": f_s, "forward_gcells_per_s": cu / f_s / 1e9,
        "forward_device_ms": float(np.mean(dev_f)),
        "forward_bytes_per_cell": b_fwd, "forward_GBs": b_fwd * cu / f_s / 1e9,
        "adjoint_s": a_s, "adjoint_gcells_per_s": cu / a_s / 1e9,
        "adjoint_device_ms": float(np.mean(dev_a)),
        "adjoint_bytes_per_cell": b_adj, "adjoint_GBs": b_adj * cu / a_s / 1e9,
        "residuals_s": float(np.mean(t_r)), "assemble_and_copy_s": float(np.mean(t_g)),
        "forward_kernel_ms": list(kf), "adjoint_kernel_ms": list(ka),
        "misfit": l2,
        "fwi_iterations": fwi,
    }


def config1_workload():
    """The reference's CPU-runnable case: elastic P/SV 400 x 200, DH = 10 m,
    FD8 Holberg, DT = 1 ms, NT = 2000, CPML 20 on three sides plus free
    surface, vp linear 1500 -> 4000 m/s with a seeded smooth +-5%
    perturbation, vs = vp/sqrt(3), rho = 310 vp^0.25, one explosive Ricker
    source of 10 Hz at x = 2000 m, z = 20 m, 200 receivers every 20 m at
    z = 20 m."""
    import dnhip
    import synth
    nx, ny, dh, dt, nt, fc = 400, 200, 10.0, 1e-3, 2000, 10.0
    vp, vs, rho = synth.layered_model(nx, ny, dh, seed=1)
    p = dnhip.Params(nx=nx, ny=ny, nt=nt, dh=dh, dt=dt, fdorder=8, maxrelerror=1, fw=20, free_surf=1,
                     L=0, seismo=1, ts=float(np.float32(1.0 / fc)), damping=1500.0, fpml=10.0)
    src = synth.source_table(2000.0, 20.0, dh, fc=fc)
    sig = synth.ricker(nt, dt, fc)[None, :]
    rec = synth.receiver_line(20.0, 20.0, 200, 20.0, dh)
    return p, (vp, vs, rho), src, sig, rec


def run_config1(threads, shots=5):
    """Config 1 on the GPU (shots/s and Gcell-updates/s over `shots` whole
    shots after one warm-up) and on the CPU oracle with 1 and `threads`
    threads (one whole shot each), with the seismograms compared."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import dnhip
    import orc
    p, (vp, vs, rho), src, sig, rec = config1_workload()
    g = dnhip.Context(p)
    g.set_model(vp, vs, rho)
    g.set_sources(src, sig)
    g.set_receivers(rec)
    g.forward()
    t0 = time.perf_counter()
    for _ in range(shots):
        g.forward()
    t_gpu = (time.perf_counter() - t0) / shots
    seis = g.seismogram(0)
    g.close()
    cells = p.nx * p.ny * p.nt
    out = {"workload": "elastic P/SV 400x200, DH=10 m, FD8, DT=1 ms, NT=2000, CPML 20 + free surface, "
                       "Ricker 10 Hz, 200 receivers",
           "gpu_ms_per_shot": t_gpu * 1e3, "gpu_gcells_per_s": cells / t_gpu / 1e9}
    for n in (1, threads):
        orc.set_threads(n)
        o = orc.Oracle(p, vp, vs, rho, prec="f32")
        o.set_sources(src, sig)
        o.set_receivers(rec)
        t0 = time.perf_counter()
        o.forward()
        dt = time.perf_counter() - t0
        out[f"cpu_{n}_threads_s"] = dt
        out[f"cpu_{n}_threads_gcells_per_s"] = cells / dt / 1e9
        if n == 1:
            out["gpu_equals_cpu_oracle"] = bool(np.array_equal(seis, o.seismogram(0)))
        o.close()
    return out


def run_aniso(physics, nt=6000, tune_steps=40, sample_every=20, tiles=None, L=0):
    """One acoustic (physics 2), VTI (3), TTI (4) or SH (5) shot on the
    4000 x 1200 grid of the headline workload (DH=5 m, FD8, DT=0.4 ms, CPML
    20 + free surface / top switch, Ricker 15 Hz, 800 receivers); the
    acoustic and SH shots on the seeded layered vp/vs/rho model, VTI/TTI on
    a seeded layered model with Thomsen epsilon/delta 0.1-0.3 / 0.05-0.15
    (TTI: tilt 30 +- 10 degrees).  Algorithmic bytes per cell-update:
    acoustic velocity 28 B (p, rip, rjp, vx, vy read; vx, vy written),
    pressure 20 B (vx, vy, p, modulus read; p written); SH velocity 20 B
    (sxz, syz, 1/rho, vz read; vz written), stress 28 B (vz, sxz, syz, uip,
    ujp read; sxz, syz written), viscoelastic SH stress (L > 0) 4 (7 + 6 L) B
    (vz, sxz, syz, fipjp, f, dip[L], d[L], r[L], q[L] read; sxz, syz, r[L],
    q[L] written; Qs 20-60 seeded); elastic velocity 36 B, stress VTI 48 B
    (5 wavefields + 4 stiffness planes read, 3 stresses written), TTI 64 B (8
    stiffness planes)."""
    import dnhip
    import synth
    nx, ny, dh, dt, fc = 4000, 1200, 5.0, 4.0e-4, 15.0
    p = dnhip.Params(nx=nx, ny=ny, nt=nt, dh=dh, dt=dt, fdorder=8, maxrelerror=1, fw=20, free_surf=1, seismo=1,
                     ts=float(np.float32(1.0 / fc)), damping=1500.0, fpml=fc, physics=physics, L=L,
                     fl=[3.0, 30.0, 60.0][:L])
    ctx = dnhip.Context(p)
    if physics == 2:
        vp, _, rho = synth.layered_model(nx, ny, dh, 11)
        ctx.set_model(vp, None, rho)
    elif physics == 5:
        _, vs, rho = synth.layered_model(nx, ny, dh, 11)
        taus = None
        if L:
            qs = np.random.default_rng(11).uniform(20.0, 60.0, (nx, ny))
            taus = (2.0 / qs).astype(np.float32)
        ctx.set_model(None, vs, rho, taus)
    else:
        rho, c11, c13, c33, c44, theta = synth.anisotropic_model(nx, ny, dh, 11, tilt=30.0 if physics == 4 else None)
        ctx.set_model_aniso(rho, c11, c13, c33, c44, theta)
    ctx.set_sources(synth.source_table(nx * dh / 2, 2 * dh, dh, fc=fc), synth.ricker(nt, dt, fc)[None, :])
    ctx.set_receivers(synth.receiver_line(2 * dh, 25.0, 800, 2 * dh, dh))
    if tiles:
        ctx.set_tile(*tiles[0], sweep=0)
        ctx.set_tile(*tiles[1], sweep=1)
    elif tune_steps > 0:
        ctx.autotune(tune_steps)
    ctx.forward()
    ctx.set_kernel_sampling(sample_every)
    t0 = time.perf_counter()
    ctx.forward()
    wall = time.perf_counter() - t0
    tm = ctx.timing()
    assert np.isfinite(ctx.seismogram(0)).all()
    cells = nx * ny
    sb = {2: 20, 3: 48, 4: 64, 5: 28 if L == 0 else 4 * (7 + 6 * L)}[physics]
    vb = {2: 28, 5: 20}.get(physics, 36)
    name = {2: "acoustic", 3: "VTI elastic", 4: "TTI elastic", 5: "SH elastic" if L == 0 else f"SH visco L={L}"}[physics]
    out = {"workload": f"{name} 4000x1200, DH=5 m, FD8, NT={nt}",
           "gcells_per_s": cells * nt / wall / 1e9, "ms_per_shot": wall * 1e3,
           "tile": {"velocity": list(ctx.tile(0)), "stress": list(ctx.tile(1))},
           "v_kernel_ms": tm["v_kernel_ms"], "s_kernel_ms": tm["s_kernel_ms"],
           "stress_bytes_per_cell": sb, "vel_bytes_per_cell": vb}
    if tm["s_kernel_ms"] > 0:
        out["s_kernel_GBs"] = sb * cells / (tm["s_kernel_ms"] * 1e-3) / 1e9
        out["v_kernel_GBs"] = vb * cells / (tm["v_kernel_ms"] * 1e-3) / 1e9
        out["s_kernel_frac_of_hbm"] = out["s_kernel_GBs"] / HBM_PEAK_GBS
        out["v_kernel_frac_of_hbm"] = out["v_kernel_GBs"] / HBM_PEAK_GBS
    out["step_frac_of_hbm"] = (sb + vb) * cells * nt / wall / 1e9 / HBM_PEAK_GBS
    ctx.close()
    return out


def run_executable_marmousi(keep_shots=100):
    """The drop-in executable on the shipped Marmousi example's parameter,
    source and receiver files (500 x 174, DH 20 m, DT 2 ms, 6 s = 3000
    steps, FD8, CPML 10 + free surface, the band-limited spike source, 400
    OBC receivers; a synthetic layered model under 400 m of water stands
    where MFILE points, the example's model is not shipped): MODE 0 over
    all its shots, once through the multi-shot engine (the default) and once
    shot by shot (DENISE_HIP_SEQUENTIAL=1).  Wall time of each process
    (start-up, model and SU file I/O included) and the SU files of three
    shots compared byte for byte."""
    import filecmp
    import shutil
    import subprocess
    import tempfile
    import synth
    gold = os.path.join(ROOT, "tests", "golden", "marmousi")
    exe = os.path.join(ROOT, "denise-black-edition_amd", "bin", "denise_hip")
    base = tempfile.mkdtemp(prefix="dnbench_", dir="/tmp")
    out = {"workload": "shipped Marmousi example, MODE 0: 500x174, NT=3000, FD8, QUELLART 6, 400 receivers"}
    try:
        lines = open(os.path.join(gold, "source_OBC_VSP.dat")).read().splitlines()
        shots = lines[1:1 + keep_shots]
        vp, vs, rho = synth.layered_model(500, 174, 20.0, seed=4, sigma=4.0)
        vp[:, :20], vs[:, :20], rho[:, :20] = 1500.0, 0.0, 1000.0
        walls = {}
        for label, env in (("batched", {}), ("sequential", {"DENISE_HIP_SEQUENTIAL": "1"})):
            d = os.path.join(base, label)
            for sub in ("start", "su", "log", "source", "receiver"):
                os.makedirs(os.path.join(d, sub), exist_ok=True)
            shutil.copy(os.path.join(gold, "DENISE_marm_OBC.inp"), d)
            shutil.copy(os.path.join(gold, "receiver_OBC.dat"), os.path.join(d, "receiver"))
            with open(os.path.join(d, "source", "source_OBC_VSP.dat"), "w") as f:
                f.write(f"{len(shots)}\n" + "\n".join(shots) + "\n")
            for ext, a in (("vp", vp), ("vs", vs), ("rho", rho)):
                np.ascontiguousarray(a, np.float32).tofile(os.path.join(d, f"start/marmousi_II_marine.{ext}"))
            t0 = time.perf_counter()
            r = subprocess.run([exe, "DENISE_marm_OBC.inp"], cwd=d, capture_output=True, text=True, timeout=600,
                               env=dict(os.environ, **env))
            walls[label] = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError(f"denise_hip ({label}) failed: {r.stderr[-500:]}")
            loops = [ln for ln in r.stdout.splitlines() if ln.startswith(" time loops:")]
            walls[label + "_loops"] = float(loops[-1].split()[2]) if loops else None
        out["shots"] = len(shots)
        out["batched_s"] = walls["batched"]
        out["sequential_s"] = walls["sequential"]
        out["speedup"] = walls["sequential"] / walls["batched"]
        # the time loops alone (engine phases / dn_run_forward calls), without
        # start-up, host set-up per shot and SU file output
        out["batched_loops_s"] = walls["batched_loops"]
        out["sequential_loops_s"] = walls["sequential_loops"]
        if walls["batched_loops"] and walls["sequential_loops"]:
            out["loops_speedup"] = walls["sequential_loops"] / walls["batched_loops"]
        cells = 500 * 174 * 3000 * len(shots)
        out["batched_gcells_per_s"] = cells / walls["batched"] / 1e9
        out["sequential_gcells_per_s"] = cells / walls["sequential"] / 1e9
        same = True
        for ishot in sorted({1, (len(shots) + 1) // 2, len(shots)}):
            for comp in ("x", "y"):
                name = f"su/DENISE_MARMOUSI_{comp}.su.shot{ishot}"
                same &= filecmp.cmp(os.path.join(base, "batched", name), os.path.join(base, "sequential", name),
                                    shallow=False)
        out["su_files_equal"] = bool(same)
    finally:
        shutil.rmtree(base, ignore_errors=True)
    return out


def host_cpu():
    """The host CPU model and thread counts (lscpu's 'Model name' from
    /proc/cpuinfo)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(L, nsteps, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    p, m, src, sig, rec = workload(L)
    orc.set_threads(threads)
    o = orc.Oracle(p, m["vp"], m["vs"], m["rho"], m["taus"], m["taup"], prec="f32")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    t0 = time.perf_counter()
    o.forward(nsteps)
    dt = time.perf_counter() - t0
    o.close()
    return p.nx * p.ny * nsteps / dt / 1e9, dt


def profiled_traffic(workload, sweep, tile):
    """HBM bytes per launch of a sweep kernel at this tile shape, as measured
    by the FETCH_SIZE / WRITE_SIZE counter runs summarised in profiles/
    (lines "traffic_bytes_per_launch <workload> <sweep> <tile> = <bytes>");
    None when no profile covers it."""
    key = f"traffic_bytes_per_launch {workload} {sweep} {','.join(map(str, tile))} ="
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith(".md"):
            continue
        for line in open(os.path.join(pdir, name)):
            line = line.strip().strip("`")
            if line.startswith(key):
                return float(line[len(key):].strip())
    return None


def calibrate_copy(mib, reps):
    """A plain device copy of known size, for scaling the HBM byte counters
    (FETCH_SIZE / WRITE_SIZE) of a profiler run."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "reps": reps + 1,
                      "ms_per_copy": ms, "copy_GBs": 2 * n * 4 / (ms * 1e-3) / 1e9}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-steps", type=int, default=800)
    ap.add_argument("--cpu1-steps", type=int, default=40, help="steps of the single-thread CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config1", action="store_true",
                    help="skip the 400x200 config (GPU and whole-shot CPU runs on 1 and all threads)")
    ap.add_argument("--no-elastic", action="store_true")
    ap.add_argument("--no-visco", action="store_true", help="profiling: elastic shot only")
    ap.add_argument("--cache-policy", type=str, default="",
                    help="profiling: 'v,s' streaming policy of the two sweeps (0/1), with --tile-v/--tile-s")
    ap.add_argument("--forms-only", action="store_true",
                    help="only the batched and served forms (no single-shot lines, no CPU baseline)")
    ap.add_argument("--batch-shots", type=int, default=64, help="multi-shot workload size (0: skip)")
    ap.add_argument("--serve-saturate-qps", type=float, default=64.0,
                    help="also serve the same requests arriving at this rate (0: skip)")
    ap.add_argument("--graph-steps", type=int, default=32,
                    help="steps per captured hipGraph in the batched and served forms (0: plain launches, "
                         "e.g. under a kernel-trace profiler)")
    ap.add_argument("--serve-requests", type=int, default=120, help="serving workload size (0: skip)")
    ap.add_argument("--serve-qps", type=float, default=8.0)
    ap.add_argument("--grad-jobs", type=int, default=2, help="timed gradient jobs (0: skip)")
    ap.add_argument("--grad-only", action="store_true", help="profiling: only the gradient job")
    ap.add_argument("--executable-only", action="store_true", help="only the executable on the Marmousi example")
    ap.add_argument("--no-aniso", action="store_true", help="skip the acoustic, VTI, TTI and SH shots")
    ap.add_argument("--aniso-only", type=int, default=0, metavar="PHYSICS",
                    help="profiling: only the acoustic (2), VTI (3), TTI (4) or SH (5) shot")
    ap.add_argument("--sh-L", type=int, default=0, help="with --aniso-only 5: the viscoelastic SH shot with L")
    ap.add_argument("--no-executable", action="store_true",
                    help="skip the drop-in executable on the shipped Marmousi example (100 shots, MODE 0)")
    ap.add_argument("--fwi-iters", type=int, default=2,
                    help="FWI iterations (L-BFGS, step search, update) after the gradient jobs (0: skip)")
    ap.add_argument("--nt", type=int, default=6000, help="profiling: shorter shots")
    ap.add_argument("--tune-steps", type=int, default=40)
    ap.add_argument("--tile-v", default="", help="profiling: fixed velocity-sweep shape tbx,by,ty,vw")
    ap.add_argument("--tile-s", default="", help="profiling: fixed stress-sweep shape tbx,by,ty,vw")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB",
                    help="profiling: run only a device copy of this size")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy, 3)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU shot sharding is a "
                                   "design note (one shot per GPU), run with --gpus 1"}))
        sys.exit(2)
    tiles = None
    if args.tile_v and args.tile_s:
        tiles = [tuple(int(v) for v in args.tile_v.split(",")), tuple(int(v) for v in args.tile_s.split(","))]
    L = 2
    policy = tuple(int(v) for v in args.cache_policy.split(",")) if args.cache_policy else None
    if args.forms_only:
        out = {}
        if args.batch_shots > 0:
            out["batched"] = run_batched(args.batch_shots, args.graph_steps)
        if args.serve_requests > 0:
            out["serving"] = run_serving(args.serve_requests, args.serve_qps, args.graph_steps, args.serve_saturate_qps)
        print(json.dumps(out))
        return
    if args.aniso_only:
        print(json.dumps({"aniso": run_aniso(args.aniso_only, nt=args.nt, tune_steps=args.tune_steps, tiles=tiles,
                                             L=args.sh_L if args.aniso_only == 5 else 0)}))
        return
    if args.executable_only:
        print(json.dumps({"executable": run_executable_marmousi()}))
        return
    if args.grad_only:
        print(json.dumps({"gradient": run_gradient(max(1, args.grad_jobs), nt=min(args.nt, 4000),
                                                   tune_steps=args.tune_steps, tiles=tiles,
                                                   fwi_iters=args.fwi_iters)}))
        return
    if args.no_visco:
        elastic = run_gpu(0, args.steps, args.warmup, nt=args.nt, tune_steps=args.tune_steps, tiles=tiles, policy=policy)
        print(json.dumps({"elastic": elastic}))
        return
    res = run_gpu(L, args.steps, args.warmup, nt=args.nt, tune_steps=args.tune_steps, tiles=tiles, policy=policy)
    print(f"# visco: {res}", file=sys.stderr, flush=True)
    elastic = None
    if not args.no_elastic:
        elastic = run_gpu(0, 1, 1, nt=args.nt, tune_steps=args.tune_steps, tiles=tiles)
        print(f"# elastic: {elastic}", file=sys.stderr, flush=True)
    batched = None
    if args.batch_shots > 0:
        batched = run_batched(args.batch_shots, args.graph_steps)
        print(f"# batched: {batched}", file=sys.stderr, flush=True)
    serving = None
    if args.serve_requests > 0:
        serving = run_serving(args.serve_requests, args.serve_qps, args.graph_steps, args.serve_saturate_qps)
        print(f"# serving: {serving}", file=sys.stderr, flush=True)
    gradient = None
    if args.grad_jobs > 0:
        gradient = run_gradient(args.grad_jobs, tune_steps=args.tune_steps, fwi_iters=args.fwi_iters)
        print(f"# gradient: {gradient}", file=sys.stderr, flush=True)
    aniso = None
    if not args.no_aniso:
        aniso = {"acoustic": run_aniso(2, nt=args.nt, tune_steps=args.tune_steps),
                 "vti": run_aniso(3, nt=args.nt, tune_steps=args.tune_steps),
                 "tti": run_aniso(4, nt=args.nt, tune_steps=args.tune_steps),
                 "sh": run_aniso(5, nt=args.nt, tune_steps=args.tune_steps),
                 "sh_visco": run_aniso(5, nt=args.nt, tune_steps=args.tune_steps, L=2),
                 "acoustic_gradient": run_gradient_other(2, tune_steps=args.tune_steps),
                 "sh_gradient": run_gradient_other(5, tune_steps=args.tune_steps)}
        print(f"# aniso: {aniso}", file=sys.stderr, flush=True)
    executable = None
    if not args.no_executable:
        executable = run_executable_marmousi()
        print(f"# executable: {executable}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        # one GPU's share of the host's cores: OMP_NUM_THREADS when it is set
        host = host_cpu()
        threads = int(host["omp_num_threads"] or 0) or min(16, os.cpu_count() or 1)
        gcps, secs = cpu_baseline(L, args.cpu_steps, threads)
        cpu = {"value": gcps, "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
               "sample": f"first {args.cpu_steps} of 6000 time steps of the same 4000x1200 "
                         f"viscoelastic shot, FP32 oracle (OpenMP), {secs:.1f} s wall",
               "host": host}
        if args.cpu1_steps > 0:
            g1, s1 = cpu_baseline(L, args.cpu1_steps, 1)
            cpu["single_thread"] = {"value": g1, "cores": 1,
                                    "sample": f"first {args.cpu1_steps} time steps, {s1:.1f} s wall"}
        if not args.no_config1:
            cpu["config1"] = run_config1(threads)
            print(f"# config1: {cpu['config1']}", file=sys.stderr, flush=True)
    achieved = res["s_kernel_GBs"]
    line = {
        "metric": METRIC,
        "value": res["gcells_per_s"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_shot"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded smooth-random layered model, Ricker source)",
        "config": {"workload": f"single-shot viscoelastic P/SV 4000x1200, DH=5 m, FD8 Holberg, "
                               f"DT=0.4 ms, NT={args.nt}, L=2 (Q=50), CPML 20 + free surface, "
                               "Ricker 15 Hz, 800 receivers",
                   "nx": 4000, "ny": 1200, "nt": args.nt, "tile": res["tile"], "L": 2, "fdorder": 8,
                   "parallelism": "single GPU, whole-grid domain"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None,
                     "traffic": profiled_traffic("visco", "stress", res["tile"]["stress"]),
                     "kernel": "viscoelastic stress + memory-variable update (FD8, L=2, uniform tau)",
                     "bytes_per_cell": res["stress_bytes_per_cell"],
                     "moved_bytes_per_cell": res["stress_moved_bytes_per_cell"],
                     "moved_GBs": res["s_kernel_moved_GBs"],
                     "avg_launch_ms": res["s_kernel_ms"]},
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": serving,
        "gradient": gradient,
        "executable": executable,
        "detail": {"visco": res, "elastic": elastic, "aniso": aniso,
                   "elastic_step_frac_of_hbm": (elastic["step_GBs"] / HBM_PEAK_GBS) if elastic else None,
                   "elastic_kernels_frac_of_hbm": (
                       80 * elastic["cells_per_step"] / ((elastic["v_kernel_ms"] + elastic["s_kernel_ms"]) * 1e-3)
                       / 1e9 / HBM_PEAK_GBS) if elastic and elastic["s_kernel_ms"] else None},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
