"""Host logic of the request server (serve.py) on the CPU: admission by
footprint within the budget, rejection of requests that can never fit,
requests that arrive later joining running ones, the two phases of a
gradient job, and a request whose context does not fit the arena after all
(DN_ERR_OOM) waiting for memory instead of failing.  The device library is
replaced by a small simulation with the same interface: an engine that
advances every attached shot by the launched number of steps, and an arena
that counts bytes."""
import threading
import time
import types

import numpy as np
import pytest

import serve


class FakeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def make_fake(budget, oom_first=()):
    """A stand-in for dnhip; oom_first: request ids whose first context
    creation fails with DN_ERR_OOM."""
    m = types.SimpleNamespace()
    m.DnError = FakeError
    lock = threading.Lock()
    m.log = {"live_bytes": 0, "max_live": 0, "attach": [], "oom": set(oom_first), "events": []}

    def footprint(p, nsrc, ntr):
        return p.nx * p.ny * 4 * (20 if p.mode > 0 else 4)

    m.estimate_footprint = footprint
    m.device_memory = lambda device=0: (budget + (2 << 30), budget + (4 << 30))

    class Arena:
        def __init__(self, nbytes, device=0):
            self.total = nbytes
            self.used = 0

        def close(self):
            assert self.used == 0, "contexts still hold arena memory"

    class Context:
        def __init__(self, p, device=0, arena=None):
            self.p, self.arena = p, arena
            self.n = footprint(p, 1, 1)
            with lock:
                if p.rid in m.log["oom"]:
                    m.log["oom"].discard(p.rid)
                    raise FakeError(-3, "no free block")
                if arena is not None:
                    assert arena.used + self.n <= arena.total, "arena over-committed"
                    arena.used += self.n
                m.log["live_bytes"] += self.n
                m.log["max_live"] = max(m.log["max_live"], m.log["live_bytes"])
            self.closed = False

        def set_model(self, *a):
            pass

        def set_sources(self, *a):
            pass

        def set_receivers(self, *a):
            pass

        def residuals(self, ox, oy, op, testshot):
            return float(self.p.rid)

        def assemble(self):
            pass

        def gradient(self, w):
            return np.full(3, self.p.rid * 10 + w, np.float32)

        def seismogram(self, c):
            return np.full(4, self.p.rid * 100 + c, np.float32)

        def close(self):
            if not self.closed:
                with lock:
                    m.log["live_bytes"] -= self.n
                    if self.arena is not None:
                        self.arena.used -= self.n
                self.closed = True

    class Engine:
        def __init__(self, device=0):
            self.jobs = {}      # ctx -> [phase, steps done]
            self.status_ = {}
            self.pending = 0

        def set_graph(self, n):
            pass

        def attach(self, ctx, phase=0):
            assert ctx not in self.jobs
            if phase == 1:
                assert getattr(ctx, "fwd_complete", False), "adjoint phase before a finished forward phase"
            self.jobs[ctx] = [phase, 0]
            self.status_[ctx] = 1
            m.log["attach"].append((ctx.p.rid, phase, len(self.jobs)))

        def launch(self, n):
            self.pending = n

        def wait(self):
            time.sleep(self.pending * 2e-4)   # the run takes time, so later requests arrive meanwhile
            for ctx, st in list(self.jobs.items()):
                st[1] += self.pending
                if st[1] >= ctx.p.nt:
                    del self.jobs[ctx]
                    self.status_[ctx] = 0
                    ctx.fwd_complete = st[0] == 0
            self.pending = 0
            return len(self.jobs)

        def status(self, ctx):
            return self.status_[ctx]

        def forget(self, ctx):
            assert ctx not in self.jobs
            self.status_.pop(ctx, None)

        def close(self):
            pass

    m.Arena, m.Context, m.Engine = Arena, Context, Engine
    return m


class P:
    """Parameters as the server reads them (rid rides along for the fake)."""

    def __init__(self, rid, nx, ny, nt, mode=0, seismo=1):
        self.rid, self.nx, self.ny, self.nt, self.mode, self.seismo = rid, nx, ny, nt, mode, seismo


def req(rid, kind, nx, ny, nt, arrival):
    p = P(rid, nx, ny, nt, mode=1 if kind == "gradient" else 0)
    src = np.zeros((8, 1), np.float32)
    rec = np.zeros((2, 3), np.int32)
    obs = (np.zeros((3, nt), np.float32),) * 2 if kind == "gradient" else ()
    return serve.Request(rid=rid, kind=kind, params=p, model=(), srcpos=src, signals=np.zeros((1, nt)),
                         recpos=rec, obs=obs, arrival=arrival)


@pytest.fixture
def fake(monkeypatch):
    def install(budget, oom_first=()):
        m = make_fake(budget, oom_first)
        monkeypatch.setattr(serve, "dnhip", m)
        return m
    return install


def test_results_routed_and_gradient_phases(fake):
    m = fake(10 << 30)
    reqs = [req(1, "forward", 100, 50, 300, 0.0), req(2, "gradient", 80, 40, 200, 0.0),
            req(3, "forward", 60, 60, 90, 0.02)]
    srv = serve.Server(chunk=32)
    st = srv.serve(reqs)
    srv.close()
    assert st["completed"] == 3 and st["rejected"] == 0
    assert np.all(reqs[0].result["seismograms"][0] == 100) and np.all(reqs[0].result["seismograms"][1] == 101)
    assert reqs[1].result["l2"] == 2.0
    assert [float(g[0]) for g in reqs[1].result["gradient"]] == [23.0, 24.0, 25.0]
    phases = [(rid, ph) for rid, ph, _ in m.log["attach"] if rid == 2]
    assert phases == [(2, 0), (2, 1)]
    # the late request joined while others were attached
    assert any(rid == 3 and n > 1 for rid, _, n in m.log["attach"])
    assert m.log["live_bytes"] == 0


def test_budget_respected_and_oversize_rejected(fake):
    reqs = [req(k, "forward", 100, 100, 100 + 10 * k, 0.0) for k in range(6)]
    need = 100 * 100 * 4 * 4
    m2 = fake(3 * need)   # room for three at a time
    big = req(9, "gradient", 400, 400, 100, 0.0)
    srv = serve.Server(budget_bytes=3 * need, chunk=16)
    st = srv.serve(reqs + [big])
    srv.close()
    assert big.error and "exceeds" in big.error
    assert st["completed"] == 6
    assert m2.log["max_live"] <= 3 * need
    # three ran together at most, and some waited for memory
    assert max(n for _, _, n in m2.log["attach"]) == 3


def test_oom_at_preparation_waits_for_memory(fake):
    m = fake(10 << 30, oom_first={2})
    reqs = [req(1, "forward", 100, 50, 200, 0.0), req(2, "forward", 100, 50, 120, 0.0)]
    srv = serve.Server(chunk=32)
    st = srv.serve(reqs)
    srv.close()
    assert st["completed"] == 2 and not reqs[1].error
    assert np.all(reqs[1].result["seismograms"][0] == 200)
    # the refused request was attached only after the first one finished
    order = [rid for rid, _, _ in m.log["attach"]]
    assert order == [1, 2]
