"""Contexts in a device-memory arena: the same results as contexts with
their own allocations, memory returned on destruction (free blocks
coalesce back to one), refusal of a context that does not fit (arena left
unchanged), fit prediction, and destruction refused while contexts live."""
import numpy as np
import pytest

import dnhip
from cases import make_case

pytestmark = pytest.mark.gpu


def _ctx(case, arena=None):
    p, m, src, sig, rec = case
    g = dnhip.Context(p, arena=arena)
    g.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    return g


def test_arena_contexts_equal_plain_and_release_memory():
    cases = [make_case(nx=96, ny=64, nt=150, fdorder=8, L=0, seismo=1, seed=1),
             make_case(nx=71, ny=53, nt=120, fdorder=4, L=2, seismo=4, seed=3),
             make_case(nx=64, ny=48, nt=100, fdorder=8, L=0, mode=1, dtinv=2, seed=5)]
    arena = dnhip.Arena(256 << 20)
    free0, largest0, total = arena.info()
    assert free0 == total == largest0
    for case in cases:
        assert arena.fits(case[0], case[2].shape[1], case[4].shape[1])
    inside = [_ctx(c, arena) for c in cases]
    with pytest.raises(dnhip.DnError) as ei:
        arena.close()                       # contexts still hold memory
    assert ei.value.code == -5
    plain = [_ctx(c) for c in cases]
    for a, b, case in zip(inside, plain, cases):
        a.forward()
        b.forward()
        comps = {1: [0, 1], 4: [0, 1, 2, 3, 4]}[case[0].seismo]
        for comp in comps:
            assert np.array_equal(a.seismogram(comp), b.seismogram(comp))
        assert a.footprint() == b.footprint()
    used = total - arena.info()[0]
    assert used >= sum(g.footprint() for g in inside)
    # a context is zeroed when created: reuse of released memory is clean
    inside[0].close()
    again = _ctx(cases[0], arena)
    again.forward()
    assert np.array_equal(again.seismogram(0), plain[0].seismogram(0))
    for g in inside[1:] + [again]:
        g.close()
    free1, largest1, _ = arena.info()
    assert free1 == total and largest1 == total
    arena.close()


def test_arena_refuses_what_does_not_fit():
    case = make_case(nx=200, ny=150, nt=400, fdorder=8, L=0, mode=1, dtinv=1, seed=2)
    need = dnhip.estimate_footprint(case[0], case[2].shape[1], case[4].shape[1])
    arena = dnhip.Arena(need // 2)
    assert not arena.fits(case[0], case[2].shape[1], case[4].shape[1])
    before = arena.info()
    with pytest.raises(dnhip.DnError) as ei:
        dnhip.Context(case[0], arena=arena)
    assert ei.value.code == -3
    assert arena.info() == before
    arena.close()
