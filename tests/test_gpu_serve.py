"""Request server on the GPU: continuous batching with admission by HBM
footprint.  Each result is the same alone and in the mix (bit for bit), a
request that arrives while others run joins them, one that does not fit
waits for memory, one larger than the budget is rejected."""
import numpy as np
import pytest

import dnhip
import serve
import synth

pytestmark = pytest.mark.gpu


def _req(rid, kind, nx, ny, nt, arrival, seed, fd=8, L=0):
    dh, dt, fc = 10.0, 1e-3, 12.0
    vp, vs, rho = synth.quick_layered_model(nx, ny, dh, seed=seed)
    p = dnhip.Params(nx=nx, ny=ny, nt=nt, dh=dh, dt=dt, fdorder=fd, fw=10, free_surf=1, seismo=1, L=L,
                     fl=[4.0, 40.0][:L], mode=1 if kind == "gradient" else 0, dtinv=2,
                     ts=float(np.float32(1.0 / fc)))
    model = (vp, vs, rho)
    if L:
        t = np.full((nx, ny), 0.04, np.float32)
        model = (vp, vs, rho, t, t)
    src = synth.source_table(nx * dh / 2, 20.0, dh, fc=fc)
    sig = synth.ricker(nt, dt, fc)[None, :]
    nrec = (nx - 2) * 10 // 20
    rec = synth.receiver_line(20.0, 20.0, nrec, 20.0, dh)
    obs = ()
    if kind == "gradient":
        rng = np.random.default_rng(seed)
        obs = (rng.standard_normal((nrec, nt)).astype(np.float32) * 1e-3,
               rng.standard_normal((nrec, nt)).astype(np.float32) * 1e-3, None)
    return serve.Request(rid=rid, kind=kind, params=p, model=model, srcpos=src, signals=sig, recpos=rec,
                         obs=obs, arrival=arrival)


def _mix():
    return [
        _req(0, "forward", 160, 90, 3000, 0.000, 1),
        _req(1, "gradient", 120, 70, 1500, 0.000, 2),
        _req(2, "forward", 200, 60, 350, 0.005, 3, fd=4),
        _req(3, "forward", 100, 100, 500, 0.010, 4, L=2),
        _req(4, "gradient", 90, 80, 260, 0.015, 5, fd=4),
        _req(5, "forward", 140, 64, 280, 0.020, 6),
    ]


def _check_same(reqs):
    for r in reqs:
        want = serve.solve_alone(r)
        if r.kind == "gradient":
            assert r.result["l2"] == want["l2"], r.rid
            for g, w in zip(r.result["gradient"], want["gradient"]):
                assert np.array_equal(g, w), r.rid
        else:
            for c, w in want["seismograms"].items():
                assert np.array_equal(r.result["seismograms"][c], w), (r.rid, c)


def test_serving_results_equal_alone():
    reqs = _mix()
    srv = serve.Server(chunk=32, graph_steps=8)
    stats = srv.serve(reqs)
    srv.close()
    print(stats, flush=True)
    assert stats["completed"] == len(reqs) and not any(r.error for r in reqs)
    # continuous batching: some request was admitted while another was running
    assert any(a.admitted < b.done and a.admitted > b.admitted for a in reqs for b in reqs if a is not b)
    _check_same(reqs)


def test_admission_waits_for_memory_and_rejects_oversize():
    reqs = _mix()[:4]
    need = [serve.Server.footprint(r) for r in reqs]
    # room for the two largest at once: the others wait for memory
    budget = sorted(need)[-1] + sorted(need)[-2]
    big = _req(9, "gradient", 400, 300, 2000, 0.0, 9)
    srv = serve.Server(budget_bytes=budget, chunk=32)
    stats = srv.serve(reqs + [big])
    srv.close()
    assert big.error and "exceeds" in big.error
    assert stats["completed"] == len(reqs)
    # at no time more than the budget in use: a waiting request is admitted
    # only after an earlier one finished
    order = sorted(reqs, key=lambda r: r.admitted)
    assert order[-1].admitted >= min(r.done for r in order[:-1])
    _check_same(reqs)
