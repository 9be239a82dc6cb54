"""dn_calc_residuals_noalloc against dn_calc_residuals through the C ABI:
the same adjoint sources (bit for bit) and misfit on two calls in a row,
the second with a smaller input (fewer receivers, shorter record), so the
shared static scratch is reused with stale contents from the first call;
inputs beyond the static scratch are refused with DN_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest

import dnhip
from cases import make_case

pytestmark = pytest.mark.gpu

DN_NOALLOC_MAX_SAMPLES = 8388608


def _gradient_ctx(nx, ny, nt, nrec, quelltypb, seed):
    p, m, src, sig, rec = make_case(nx=nx, ny=ny, nt=nt, mode=1, quelltypb=quelltypb, seismo=4,
                                    nrec=nrec, seed=seed)
    g = dnhip.Context(p)
    g.set_model(m["vp"], m["vs"], m["rho"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    g.forward()
    rng = np.random.default_rng(seed)
    obs = [(g.seismogram(c) * (1.0 + 0.1 * rng.standard_normal((rec.shape[1], nt)))).astype(np.float32)
           for c in (0, 1, 2)]
    return g, obs


def _adjoint(g):
    return [g.adjoint_sources(c) for c in (0, 1, 2)]


@pytest.mark.parametrize("quelltypb", [1, 4, 7])
def test_noalloc_residuals_equal_the_allocating_entry_point(quelltypb):
    print(f"[noalloc] QUELLTYPB {quelltypb}: two calls, the second smaller", flush=True)
    for nx, ny, nt, nrec, seed in ((120, 80, 400, 50, 3), (90, 60, 250, 20, 4)):
        g, obs = _gradient_ctx(nx, ny, nt, nrec, quelltypb, seed)
        l2_a = g.residuals(*obs, testshot=1)
        adj_a = _adjoint(g)
        l2_b = g.residuals_noalloc(*obs, testshot=1)
        adj_b = _adjoint(g)
        assert l2_b == l2_a and l2_a > 0
        for a, b in zip(adj_a, adj_b):
            assert np.array_equal(a, b)
        assert any(np.abs(a).max() > 0 for a in adj_a)
        # not a test shot: misfit 0, same residuals
        assert g.residuals_noalloc(*obs, testshot=0) == 0.0
        for a, b in zip(adj_a, _adjoint(g)):
            assert np.array_equal(a, b)
        g.close()


def test_noalloc_refuses_inputs_beyond_its_scratch():
    nrec, nt = 70, DN_NOALLOC_MAX_SAMPLES // 70 + 1
    p, m, src, sig, rec = make_case(nx=160, ny=40, nt=nt, mode=0, nrec=nrec)
    g = dnhip.Context(p)
    g.set_model(m["vp"], m["vs"], m["rho"])
    g.set_receivers(rec)
    assert rec.shape[1] * nt > DN_NOALLOC_MAX_SAMPLES
    obs = np.zeros((rec.shape[1], nt), np.float32)
    lib = dnhip.load()
    l2 = C.c_double(-1.0)
    fp = C.POINTER(C.c_float)
    rc = lib.dn_calc_residuals_noalloc(g.h, obs.ctypes.data_as(fp), obs.ctypes.data_as(fp), None, 1, C.byref(l2))
    assert rc == -1 and b"static scratch" in lib.dn_last_error()
    assert l2.value == -1.0
    g.close()
