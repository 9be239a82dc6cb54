"""The batched workload at full size: 64 independent elastic P/SV shots of
uneven size (NX in [500, 3000], NY in [200, 800], DH = 10 m, FD4 or FD8,
NT in [1500, 5000], seeded smooth-random models, Ricker sources of 5-20 Hz)
advanced together by the multi-shot engine with hipGraph replay.  Every
shot's seismograms equal its single-shot run (dn_run_forward) bit for bit,
and the smallest shot's equal the FP32 oracle over its whole NT."""
import numpy as np
import pytest

import dnhip
import orc
from cases import rel_l2
from workloads import batch_workload

pytestmark = pytest.mark.gpu


def _prepare(shot):
    p, (vp, vs, rho), src, sig, rec = shot
    g = dnhip.Context(p)
    g.set_model(vp, vs, rho)
    g.set_sources(src, sig)
    g.set_receivers(rec)
    return g


def test_config4_engine_equals_single_shots_and_oracle():
    shots = batch_workload(64)
    cells = sum(p.nx * p.ny * p.nt for p, *_ in shots)
    print(f"[config4] 64 shots, {cells / 1e9:.1f} G cell-updates, "
          f"FD orders {sorted(set(p.fdorder for p, *_ in shots))}", flush=True)
    ctxs = [_prepare(s) for s in shots]
    eng = dnhip.Engine()
    eng.set_graph(32)
    for g in ctxs:
        eng.attach(g)
    eng.run(chunk=256)
    print("[config4] engine run done", flush=True)
    for k, g in enumerate(ctxs):
        assert eng.status(g) == 0, k
    # each shot alone on the plain solver, compared and released one at a time
    for k, (shot, g) in enumerate(zip(shots, ctxs)):
        a = _prepare(shot)
        a.forward()
        for c in (0, 1):
            assert np.array_equal(a.seismogram(c), g.seismogram(c)), (k, c)
        a.close()
    print("[config4] 64 shots equal their single-shot runs", flush=True)
    # the smallest shot against the FP32 oracle over its whole NT
    k = min(range(len(shots)), key=lambda i: shots[i][0].nx * shots[i][0].ny * shots[i][0].nt)
    p, (vp, vs, rho), src, sig, rec = shots[k]
    orc.set_threads(16)
    o = orc.Oracle(p, vp, vs, rho, prec="f32")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    for c in (0, 1):
        err = rel_l2(ctxs[k].seismogram(c), o.seismogram(c))
        print(f"[config4] shot {k} ({p.nx}x{p.ny}, NT={p.nt}, FD{p.fdorder}) comp {c} rel-L2 vs FP32 oracle "
              f"{err:.2e}", flush=True)
        assert err <= 1e-5
    o.close()
    for g in ctxs:
        g.close()
    eng.close()
