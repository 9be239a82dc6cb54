"""The served workload at the edges of its size range: forward modelling and
FWI gradient requests from 256 x 256 to 4096 x 1024 cells, NT from 1000 to
8000, arriving close together so that they run in one mix.  Each result
equals the same request solved alone (bit for bit), and a gradient job whose
stored forward wavefield cannot fit the device (4096 x 1024, NT = 8000:
5 fields x 4000 snapshots x 4.2 M cells x 4 B = 335 GB) is rejected."""
import numpy as np
import pytest

import dnhip
import serve
import synth

pytestmark = pytest.mark.gpu


def _req(rid, kind, nx, ny, nt, arrival, seed):
    dh, dt, fc = 10.0, 1e-3, 10.0
    vp, vs, rho = synth.quick_layered_model(nx, ny, dh, seed=seed)
    p = dnhip.Params(nx=nx, ny=ny, nt=nt, dh=dh, dt=dt, fdorder=8, fw=20, free_surf=1, seismo=1,
                     mode=1 if kind == "gradient" else 0, dtinv=2, ts=float(np.float32(1.0 / fc)))
    src = synth.source_table(nx * dh / 2, 20.0, dh, fc=fc)
    sig = synth.ricker(nt, dt, fc)[None, :]
    nrec = (nx - 2) * int(dh) // 20
    rec = synth.receiver_line(20.0, 20.0, nrec, 20.0, dh)
    obs = ()
    if kind == "gradient":
        rng = np.random.default_rng(seed)
        obs = (rng.standard_normal((nrec, nt)).astype(np.float32) * 1e-4,
               rng.standard_normal((nrec, nt)).astype(np.float32) * 1e-4, None)
    return serve.Request(rid=rid, kind=kind, params=p, model=(vp, vs, rho), srcpos=src, signals=sig,
                         recpos=rec, obs=obs, arrival=arrival)


def test_config5_size_range_results_equal_alone():
    reqs = [
        _req(0, "forward", 4096, 1024, 8000, 0.00, 11),     # the largest forward request
        _req(1, "forward", 256, 256, 1000, 0.01, 12),       # the smallest
        _req(2, "gradient", 256, 256, 8000, 0.02, 13),      # smallest grid, longest record
        _req(3, "gradient", 2048, 600, 3000, 0.03, 14),     # 37 GB of stored snapshots
        _req(4, "forward", 1024, 512, 5000, 0.04, 15),
        _req(5, "gradient", 4096, 1024, 8000, 0.05, 16),    # 335 GB of snapshots: cannot fit
    ]
    for r in reqs:
        print(f"[config5] request {r.rid}: {r.kind} {r.params.nx}x{r.params.ny} NT={r.params.nt}, "
              f"footprint {serve.Server.footprint(r) / 1e9:.1f} GB", flush=True)
    srv = serve.Server(chunk=128, graph_steps=32)
    stats = srv.serve(reqs)
    srv.close()
    print({k: stats[k] for k in ("completed", "rejected", "shots_per_s", "latency_p50_s", "latency_p99_s")},
          flush=True)
    big = reqs[5]
    assert big.error and "exceeds" in big.error
    done = reqs[:5]
    assert stats["completed"] == len(done) and not any(r.error for r in done)
    # requests ran together: some request was admitted while another was still running
    assert any(a.admitted < b.done and a.admitted > b.admitted for a in done for b in done if a is not b)
    for r in done:
        want = serve.solve_alone(r)
        if r.kind == "gradient":
            assert r.result["l2"] == want["l2"], r.rid
            for g, w in zip(r.result["gradient"], want["gradient"]):
                assert np.array_equal(g, w), r.rid
                assert np.isfinite(g).all() and np.abs(g).max() > 0, r.rid
        else:
            for c, w in want["seismograms"].items():
                assert np.array_equal(r.result["seismograms"][c], w), (r.rid, c)
                assert np.abs(w).max() > 0, (r.rid, c)
        print(f"[config5] request {r.rid} equals its alone run", flush=True)
