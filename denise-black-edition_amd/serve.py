"""Request server for P/SV modelling and FWI-gradient jobs on one GPU.

Continuous batching over the multi-shot engine (dn_engine): requests that
arrive while others run are admitted between engine steps as soon as their
device footprint (wavefields, CPML memory, seismograms and, for gradients,
the stored forward snapshots -- dn_estimate_footprint) fits in the HBM
budget; admitted shots keep their state on the device and advance together
with the running ones.  A gradient job runs its forward phase, has its
residuals formed on the device, then rejoins the engine for its adjoint
phase.  A request whose footprint exceeds the whole budget is rejected.

The host loop only admits, steps and collects; every time step of every
shot is GPU work.  Admitted requests are prepared (device allocation,
model, sources, receivers) by a worker thread while the engine steps the
running shots, so preparation does not stall the GPU; a prepared shot joins
at the next step boundary.  Latency = completion time - arrival time.
"""
from __future__ import annotations

import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

import dnhip


@dataclass
class Request:
    """One job: kind "forward" (seismograms) or "gradient" (vp/vs/rho
    gradient of the L2 misfit against obs)."""
    rid: int
    kind: str
    params: dnhip.Params
    model: tuple                # (vp, vs, rho[, taus, taup]) x-major float32
    srcpos: np.ndarray          # (8, nsrc)
    signals: np.ndarray         # (nsrc, NT)
    recpos: np.ndarray          # (2, ntr)
    obs: tuple = ()             # gradient: (obs_vx, obs_vy, obs_p or None), ntr x NT
    arrival: float = 0.0        # seconds after the start of the run
    # filled in by the server
    result: dict = field(default_factory=dict)
    admitted: float = -1.0
    done: float = -1.0
    error: str = ""


class Server:
    def __init__(self, device: int = 0, budget_bytes: int | None = None, chunk: int = 64,
                 graph_steps: int = 32, reserve_bytes: int = 2 << 30, arena: bool = True):
        self.device = device
        free, total = dnhip.device_memory(device)
        self.budget = budget_bytes if budget_bytes is not None else max(0, free - reserve_bytes)
        self.used = 0
        # one device allocation for every request's state, made once here:
        # contexts are then carved from it without hipMalloc / hipFree
        self.arena = dnhip.Arena(self.budget, device) if arena and self.budget > 0 else None
        self._requeue: list[Request] = []  # admitted, but the arena had no room after all
        self._hold = False                 # admit nothing until a shot releases memory
        self.chunk = chunk
        self.engine = dnhip.Engine(device)
        self.engine.set_graph(graph_steps)
        self.running = {}          # Context -> (Request, phase, bytes)
        self.preparing = 0         # admitted requests not yet attached
        self._todo: queue.Queue = queue.Queue()
        self._ready: queue.Queue = queue.Queue()
        self._worker = threading.Thread(target=self._prepare_loop, daemon=True)
        self._worker.start()

    @staticmethod
    def footprint(req: Request) -> int:
        return dnhip.estimate_footprint(req.params, req.srcpos.shape[1], req.recpos.shape[1])

    def _prepare_loop(self) -> None:
        """Worker thread: build the device state of admitted requests."""
        while True:
            item = self._todo.get()
            if item is None:
                return
            req, nbytes = item
            ctx = None
            try:
                if req.kind == "gradient" and req.params.mode == 0:
                    raise ValueError("gradient requests need params.mode = 1")
                ctx = dnhip.Context(req.params, self.device, arena=self.arena)
                ctx.set_model(*req.model)
                ctx.set_sources(req.srcpos, req.signals)
                ctx.set_receivers(req.recpos)
                self._ready.put((req, nbytes, ctx, None))
            except Exception as ex:  # reported on the request
                if ctx is not None:
                    ctx.close()
                self._ready.put((req, nbytes, None, ex))

    def _admit(self, req: Request, need: int, now: float) -> None:
        self.used += need
        self.preparing += 1
        req.admitted = now
        self._todo.put((req, need))

    def _attach_ready(self, now: float, block: bool) -> None:
        """Attach every prepared shot (waiting for one when block is set)."""
        while self.preparing:
            try:
                req, nbytes, ctx, err = self._ready.get(block=block, timeout=60.0 if block else None)
            except queue.Empty:
                if block:
                    raise RuntimeError("request preparation did not finish within 60 s")
                return
            block = False
            self.preparing -= 1
            if isinstance(err, dnhip.DnError) and err.code == -3 and self.arena is not None:
                # the arena's free blocks were too fragmented for this shot:
                # it waits (first in line) until a running shot releases memory
                self.used -= nbytes
                req.admitted = -1.0
                self._requeue.append(req)
                self._hold = True
                continue
            if err is not None:
                req.error = f"preparation failed: {err}"
                req.done = now
                self.used -= nbytes
                continue
            self.engine.attach(ctx, 0)
            self.running[ctx] = (req, 0, nbytes)

    def _finish(self, ctx, entry, now: float) -> None:
        """Results of a shot that left the engine at time `now` (runs while
        the engine works on the next run of steps)."""
        req, phase, nbytes = entry
        st = self.engine.status(ctx)
        if st != 0:
            req.error = f"shot failed with status {st}"
        elif req.kind == "gradient" and phase == 0:
            ox, oy, op = (list(req.obs) + [None, None, None])[:3]
            req.result["l2"] = ctx.residuals(ox, oy, op, 1)
            self.engine.forget(ctx)
            self.engine.attach(ctx, 1)
            self.running[ctx] = (req, 1, nbytes)
            return
        elif req.kind == "gradient":
            ctx.assemble()
            req.result["gradient"] = [ctx.gradient(w) for w in (3, 4, 5)]
        else:
            comps = {1: (0, 1), 2: (2,), 3: (3, 4), 4: (0, 1, 2, 3, 4)}.get(req.params.seismo, ())
            req.result["seismograms"] = {c: ctx.seismogram(c) for c in comps}
        req.done = now
        self.engine.forget(ctx)
        ctx.close()
        self.used -= nbytes
        self._hold = False

    def serve(self, requests: list[Request], clock=time.perf_counter) -> dict:
        """Run a request stream (arrival times relative to the start) to
        completion; returns latency and throughput statistics."""
        todo = sorted(requests, key=lambda r: r.arrival)
        waiting: list[Request] = []
        t0 = clock()
        k = 0
        t_wait, n_calls, shot_calls = 0.0, 0, 0
        finished = []    # (ctx, entry, time) of shots whose results are not read yet
        while k < len(todo) or waiting or self.running or self.preparing or finished:
            now = clock() - t0
            while k < len(todo) and todo[k].arrival <= now:
                waiting.append(todo[k])
                k += 1
            # admission by footprint, oldest first; a request that can never
            # fit is rejected, one that does not fit now waits
            if self._requeue:
                waiting = self._requeue + waiting
                self._requeue = []
            still = []
            for req in waiting:
                need = self.footprint(req)
                if self._hold:
                    still.append(req)
                elif need > self.budget:
                    req.error = f"footprint {need} B exceeds the budget {self.budget} B"
                    req.done = now
                elif self.used + need <= self.budget:
                    self._admit(req, need, now)
                else:
                    still.append(req)
            waiting = still
            # prepared shots join; with nothing to do, wait for one
            self._attach_ready(clock() - t0, block=not self.running and not finished and self.preparing > 0)
            launched = bool(self.running)
            if launched:
                self.engine.launch(self.chunk)
            # while the GPU runs: results of the shots finished in the last
            # run (a gradient's forward phase turns into its adjoint phase,
            # which joins the next run)
            for ctx, entry, t in finished:
                self._finish(ctx, entry, t)
            finished = []
            if launched:
                ts = clock()
                self.engine.wait()
                t_wait += clock() - ts
                n_calls += 1
                shot_calls += len(self.running)
                now = clock() - t0
                for ctx in [c for c in self.running if self.engine.status(c) != 1]:
                    finished.append((ctx, self.running.pop(ctx), now))
            elif k < len(todo) and not self.preparing:
                time.sleep(max(0.0, min(0.01, todo[k].arrival - (clock() - t0))))
        wall = clock() - t0
        ok = [r for r in requests if not r.error]
        cells = sum(r.params.nx * r.params.ny * r.params.nt * (2 if r.kind == "gradient" else 1) for r in ok)
        lat = np.array([r.done - r.arrival for r in ok]) if ok else np.zeros(1)
        return {
            "requests": len(requests), "completed": len(ok), "rejected": len(requests) - len(ok),
            "wall_s": wall, "shots_per_s": len(ok) / wall if wall > 0 else 0.0,
            "gcells_per_s": cells / wall / 1e9 if wall > 0 else 0.0,
            "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
            "latency_mean_s": float(lat.mean()),
            # host time spent waiting for engine runs (GPU work not overlapped
            # by host work), and the mean number of shots sharing a run
            "engine_wait_s": t_wait, "mean_shots_per_run": shot_calls / n_calls if n_calls else 0.0,
        }

    def close(self):
        self._todo.put(None)
        self._worker.join()
        while not self._ready.empty():
            _, _, ctx, _ = self._ready.get()
            if ctx is not None:
                ctx.close()
        for ctx in list(self.running):
            ctx.close()
        self.running.clear()
        self.engine.close()
        if self.arena is not None:
            self.arena.close()


def solve_alone(req: Request, device: int = 0) -> dict:
    """The same job with the plain one-shot solver (reference for the mix)."""
    ctx = dnhip.Context(req.params, device)
    ctx.set_model(*req.model)
    ctx.set_sources(req.srcpos, req.signals)
    ctx.set_receivers(req.recpos)
    ctx.forward()
    out = {}
    if req.kind == "gradient":
        ox, oy, op = (list(req.obs) + [None, None, None])[:3]
        out["l2"] = ctx.residuals(ox, oy, op, 1)
        ctx.adjoint()
        ctx.assemble()
        out["gradient"] = [ctx.gradient(w) for w in (3, 4, 5)]
    else:
        comps = {1: (0, 1), 2: (2,), 3: (3, 4), 4: (0, 1, 2, 3, 4)}.get(req.params.seismo, ())
        out["seismograms"] = {c: ctx.seismogram(c) for c in comps}
    ctx.close()
    return out
