"""Multi-shot engine on the GPU: shots of different grid size, FD order,
length and physics advanced together give exactly the results of each shot
run alone (dn_run_forward / dn_run_adjoint), with plain launches and with
hipGraph replay, and shots can join while others are running."""
import numpy as np
import pytest

import dnhip
import orc
from cases import make_case, rel_l2

pytestmark = pytest.mark.gpu

FIELDS = [(dnhip.F_VX, "vx"), (dnhip.F_VY, "vy"), (dnhip.F_SXX, "sxx"), (dnhip.F_SYY, "syy"), (dnhip.F_SXY, "sxy")]


def _mix():
    """Five shots: uneven grids, FD4/FD8/FD6, elastic and viscoelastic,
    force and explosive sources, different NT."""
    return [
        make_case(nx=96, ny=64, nt=180, fdorder=8, L=0, seismo=4, seed=1),
        make_case(nx=150, ny=70, nt=240, fdorder=4, L=0, seismo=1, seed=2, stype=3),
        make_case(nx=71, ny=53, nt=130, fdorder=8, L=2, seismo=4, seed=3),
        make_case(nx=128, ny=96, nt=210, fdorder=6, L=0, seismo=2, seed=4, free_surf=0,
                  src_xz=(640.0, 480.0)),
        make_case(nx=200, ny=40, nt=160, fdorder=4, L=1, seismo=1, seed=5, stype=2),
    ]


def _ctx(case, tile=None):
    p, m, src, sig, rec = case
    g = dnhip.Context(p)
    if tile:
        g.set_tile(*tile[0], sweep=0)
        g.set_tile(*tile[1], sweep=1)
    g.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    return g


def _same(a, b, case, tag):
    p = case[0]
    comps = {1: [0, 1], 2: [2], 3: [3, 4], 4: [0, 1, 2, 3, 4]}[p.seismo]
    for c in comps:
        assert np.array_equal(a.seismogram(c), b.seismogram(c)), (tag, "seis", c)
    for fid, name in FIELDS:
        assert np.array_equal(a.field(fid), b.field(fid)), (tag, name)


@pytest.mark.parametrize("graph", [0, 16])
def test_engine_forward_equals_single_shots(graph):
    cases = _mix()
    alone = []
    for case in cases:
        g = _ctx(case)
        g.forward()
        alone.append(g)
    eng = dnhip.Engine()
    eng.set_graph(graph)
    shots = [_ctx(case) for case in cases]
    for g in shots:
        eng.attach(g)
    print(f"[engine graph={graph}] {len(shots)} shots", flush=True)
    eng.run(chunk=50)
    for k, (case, a, b) in enumerate(zip(cases, alone, shots)):
        assert eng.status(b) == 0
        _same(a, b, case, f"shot{k}")
    # and the oracle for one of them
    p, m, src, sig, rec = cases[2]
    o = orc.Oracle(p, m["vp"], m["vs"], m["rho"], m["taus"], m["taup"], prec="f32")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    assert rel_l2(shots[2].seismogram(0), o.seismogram(0)) <= 1e-5


def test_engine_shots_join_while_others_run():
    """Continuous batching: a shot admitted mid-run of others ends with the
    same result as alone, and so do the others."""
    cases = _mix()[:3]
    alone = []
    for case in cases:
        g = _ctx(case)
        g.forward()
        alone.append(g)
    eng = dnhip.Engine()
    eng.set_graph(8)
    a, b, c = (_ctx(case) for case in cases)
    eng.attach(a)
    eng.step(37)
    eng.attach(b)
    eng.step(29)
    eng.attach(c)
    eng.run(chunk=23)
    for k, (case, x, y) in enumerate(zip(cases, alone, (a, b, c))):
        _same(x, y, case, f"join{k}")


def test_engine_launch_wait_attach_in_flight():
    """Asynchronous runs: a shot attached while a run is in flight joins the
    next run; a second launch before the wait is refused; results equal the
    shots alone."""
    cases = _mix()[1:4]
    alone = []
    for case in cases:
        g = _ctx(case)
        g.forward()
        alone.append(g)
    eng = dnhip.Engine()
    eng.set_graph(8)
    a, b, c = (_ctx(case) for case in cases)
    eng.attach(a)
    eng.launch(40)
    with pytest.raises(dnhip.DnError) as ei:
        eng.launch(10)
    assert ei.value.code == -5
    eng.attach(b)              # joins the next run
    assert eng.wait() == 2
    eng.launch(17)
    eng.attach(c)
    eng.wait()
    while eng.step(30) > 0:
        pass
    for k, (case, x, y) in enumerate(zip(cases, alone, (a, b, c))):
        _same(x, y, case, f"async{k}")


def test_engine_gradient_phases():
    """Forward (with snapshots) and adjoint phases through the engine give
    the single-shot gradient of every shot."""
    base = [make_case(nx=72, ny=56, nt=300, fdorder=8, L=0, mode=1, dtinv=2, fc=15.0, nrec=30, seed=7),
            make_case(nx=90, ny=48, nt=260, fdorder=4, L=2, mode=1, dtinv=3, fc=15.0, nrec=30, seed=8)]
    obs = []
    for case in base:
        p, m, src, sig, rec = case
        mt = dict(m)
        mt["vp"] = m["vp"].copy()
        mt["vp"][30:42, 8:16] *= 1.08
        g = dnhip.Context(dnhip.Params(**{**p.__dict__, "mode": 0}))
        g.set_model(mt["vp"], mt["vs"], mt["rho"], mt["taus"], mt["taup"])
        g.set_sources(src, sig)
        g.set_receivers(rec)
        g.forward()
        obs.append((g.seismogram(0), g.seismogram(1)))
    single = []
    for case, (ox, oy) in zip(base, obs):
        g = _ctx(case)
        g.forward()
        g.residuals(ox, oy, None, 1)
        g.adjoint()
        g.assemble()
        single.append([g.gradient(w) for w in range(6)])
    eng = dnhip.Engine()
    shots = [_ctx(case) for case in base]
    for g in shots:
        eng.attach(g, 0)
    eng.run(chunk=64)
    for g, (ox, oy) in zip(shots, obs):
        g.residuals(ox, oy, None, 1)
        eng.attach(g, 1)
    eng.run(chunk=64)
    for k, g in enumerate(shots):
        g.assemble()
        for w in range(6):
            assert np.array_equal(g.gradient(w), single[k][w]), (k, w)


def test_engine_lean_tiles_equal_single_shots():
    """Grids with interior (lean) tiles for the engine's shapes, forward and
    gradient phases: every shot equals its single-shot run bit for bit."""
    cases = [make_case(nx=600, ny=200, nt=160, fdorder=8, L=0, seismo=4, seed=21, src_xz=(3000.0, 900.0)),
             make_case(nx=520, ny=180, nt=140, fdorder=4, L=2, seismo=1, seed=22, src_xz=(2000.0, 700.0))]
    alone = []
    for case in cases:
        g = _ctx(case)
        g.forward()
        alone.append(g)
    eng = dnhip.Engine()
    shots = [_ctx(case) for case in cases]
    for g in shots:
        eng.attach(g)
    eng.run(chunk=40)
    for k, (case, a, b) in enumerate(zip(cases, alone, shots)):
        _same(a, b, case, f"lean{k}")
    gc = make_case(nx=420, ny=150, nt=200, fdorder=8, L=0, mode=1, dtinv=2, fc=15.0, nrec=60, seed=23)
    p, m, src, sig, rec = gc
    mt = dict(m)
    mt["vp"] = m["vp"] * 1.03
    t = dnhip.Context(dnhip.Params(**{**p.__dict__, "mode": 0}))
    t.set_model(mt["vp"], mt["vs"], mt["rho"], mt["taus"], mt["taup"])
    t.set_sources(src, sig)
    t.set_receivers(rec)
    t.forward()
    ox, oy = t.seismogram(0), t.seismogram(1)
    single = _ctx(gc)
    single.forward()
    single.residuals(ox, oy, None, 1)
    single.adjoint()
    single.assemble()
    g = _ctx(gc)
    eng.attach(g, 0)
    eng.run(chunk=64)
    g.residuals(ox, oy, None, 1)
    eng.attach(g, 1)
    eng.run(chunk=64)
    g.assemble()
    for w in range(6):
        assert np.array_equal(g.gradient(w), single.gradient(w)), w


@pytest.mark.parametrize("stress", [(64, 4, 1, 4), (64, 4, 2, 1), (64, 4, 2, 2)])
def test_engine_elastic_stress_shapes(stress):
    """Every stress shape the multi-shot kernels compile for elastic groups
    (the context's shape is kept): shots equal their single-shot runs."""
    cases = [make_case(nx=600, ny=200, nt=150, fdorder=8, L=0, seismo=4, seed=31, src_xz=(3000.0, 900.0)),
             make_case(nx=333, ny=121, nt=170, fdorder=4, L=0, seismo=1, seed=32, stype=2)]
    tile = ((128, 2, 4, 1), stress)
    alone = []
    for case in cases:
        g = _ctx(case, tile)
        g.forward()
        alone.append(g)
    eng = dnhip.Engine()
    shots = [_ctx(case, tile) for case in cases]
    for g in shots:
        eng.attach(g)
        assert g.tile(1) == stress
    print(f"[engine stress {stress}]", flush=True)
    eng.run(chunk=64)
    for k, (case, a, b) in enumerate(zip(cases, alone, shots)):
        _same(a, b, case, f"stress{k}")


def test_engine_rejects_moment_tensor_and_double_attach():
    case = make_case(nx=64, ny=48, nt=50, stype=5)
    g = _ctx(case)
    eng = dnhip.Engine()
    with pytest.raises(dnhip.DnError):
        eng.attach(g)
    h = _ctx(make_case(nx=64, ny=48, nt=50))
    eng.attach(h)
    with pytest.raises(dnhip.DnError) as ei:
        eng.attach(h)
    assert ei.value.code == -5
    eng.run()
    with pytest.raises(dnhip.DnError):
        eng.attach(h, 1)          # no snapshots: params.mode == 0


def test_footprint_estimate_matches_allocation():
    for kw in (dict(L=0, mode=0), dict(L=2, mode=0), dict(L=0, mode=1, dtinv=3), dict(L=1, mode=1, grad_form=2)):
        case = make_case(nx=97, ny=61, nt=120, **kw)
        g = _ctx(case)
        est = dnhip.estimate_footprint(case[0], case[2].shape[1], case[4].shape[1])
        got = g.footprint()
        assert 0 <= got - est <= 64 * 1024, (kw, est, got)     # cell lists are the difference
