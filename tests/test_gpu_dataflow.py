"""Dataflow passes (dn_set_dataflow): one launch per forward or adjoint pass
whose (step, sweep, tile) items start when the neighbouring tiles have
finished the preceding sweep.  Every result -- seismograms, final
wavefields, CPML memory, snapshots through the gradient, energy weights --
equals the two-launches-per-step run bit for bit, for every scalar tile
shape (whose dependency radius in tiles differs with the FD order), FD2-12,
with and without the free surface, every source type and SEISMO mode,
DTINV 1-3, GRAD_FORM 1 and 2 and the adjoint-source variants; runs it does
not apply to keep the per-step launches."""
import numpy as np
import pytest

import dnhip
import orc
from cases import make_case, rel_l2

# The dataflow pass did not finish its first case (forward, shape 64,4,1,1,
# FD2, 150 x 90, 220 steps) within the test's 120 s limit; no mistake in
# its code has been shown (agent-scope release/acquire per work item writes
# back and invalidates each XCD's L2, and its spin limit was too long to end
# a stalled pass in time), so the path is not in the library and these tests
# stay skipped.
pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="dataflow pass removed: its first case ran past the 120 s limit; "
                                      "cause not shown")]

SCALAR = [(64, 4, 1, 1), (128, 2, 1, 1), (256, 1, 1, 1), (64, 4, 2, 1), (128, 2, 4, 1), (64, 8, 2, 1)]


def _ctx(p, m, src, sig, rec, tile, flow):
    g = dnhip.Context(p)
    g.set_model(m["vp"], m["vs"], m["rho"], m["taus"], m["taup"])
    g.set_sources(src, sig)
    g.set_receivers(rec)
    g.set_tile(*tile, sweep=-1)
    g.set_dataflow(flow)
    return g


def _forward_pair(case, tile):
    p, m, src, sig, rec = case
    a = _ctx(p, m, src, sig, rec, tile, 0)
    b = _ctx(p, m, src, sig, rec, tile, 1)
    assert a.dataflow() == (0, 0)
    assert b.dataflow() == (1, 1)
    a.forward()
    b.forward()
    return a, b


def _same_forward(a, b):
    for c in range(5):
        x, y = a.seismogram(c), b.seismogram(c)
        assert np.array_equal(x, y), (c, rel_l2(y, x))
    for fid in (dnhip.F_VX, dnhip.F_VY, dnhip.F_SXX, dnhip.F_SYY, dnhip.F_SXY):
        assert np.array_equal(a.field(fid), b.field(fid)), fid
    for k in range(9):
        assert np.array_equal(a.psi(k), b.psi(k)), k


@pytest.mark.parametrize("tile", SCALAR)
@pytest.mark.parametrize("fdorder", [2, 8, 12])
def test_forward_equals_per_step_launches(tile, fdorder):
    print(f"[flow] forward {tile} FD{fdorder}", flush=True)
    case = make_case(nx=150, ny=90, nt=220, fdorder=fdorder, seismo=4)
    a, b = _forward_pair(case, tile)
    _same_forward(a, b)
    assert np.abs(a.seismogram(0)).max() > 0
    a.close()
    b.close()


@pytest.mark.parametrize("free_surf,stype,seismo", [(0, 1, 4), (1, 2, 1), (1, 3, 2), (0, 4, 3), (1, 1, 3)])
def test_forward_boundaries_sources_receivers(free_surf, stype, seismo):
    print(f"[flow] forward free_surf {free_surf} stype {stype} seismo {seismo}", flush=True)
    case = make_case(nx=130, ny=70, nt=240, free_surf=free_surf, stype=stype, seismo=seismo)
    a, b = _forward_pair(case, (64, 4, 2, 1))
    _same_forward(a, b)
    a.close()
    b.close()


def test_forward_equals_oracle():
    case = make_case(nx=110, ny=70, nt=260, seismo=4)
    p, m, src, sig, rec = case
    b = _ctx(p, m, src, sig, rec, (64, 4, 2, 1), 1)
    b.forward()
    o = orc.Oracle(p, m["vp"], m["vs"], m["rho"], prec="f32")
    o.set_sources(src, sig)
    o.set_receivers(rec)
    o.forward()
    for c in range(5):
        assert np.array_equal(b.seismogram(c), o.seismogram(c)), c
    for fid in (dnhip.F_VX, dnhip.F_VY, dnhip.F_SXX, dnhip.F_SYY, dnhip.F_SXY):
        assert np.array_equal(b.field(fid), o.field(fid)), fid
    b.close()


def _gradient_run(case, tile, flow, eprecond=0):
    p, m, src, sig, rec = case
    g = _ctx(p, m, src, sig, rec, tile, flow)
    if eprecond:
        g.set_energy_weights(eprecond)
    g.forward()
    obs = [0.9 * g.seismogram(c) for c in range(3)]
    l2 = g.residuals(obs[0], obs[1], obs[2], 1)
    g.adjoint()
    g.assemble()
    out = {"l2": l2, "grad": [g.gradient(k) for k in range(6)]}
    if eprecond:
        out["ew"] = [g.energy(0), g.energy(1)]
    for fid in (dnhip.F_VX, dnhip.F_VY, dnhip.F_SXX, dnhip.F_SYY, dnhip.F_SXY):
        out[f"f{fid}"] = g.field(fid)
    g.close()
    return out


@pytest.mark.parametrize("tile", SCALAR)
def test_gradient_equals_per_step_launches(tile):
    """Forward with the snapshot store, residuals, adjoint with the
    correlation in both sweeps, assembly: the six gradient planes."""
    print(f"[flow] gradient {tile}", flush=True)
    case = make_case(nx=120, ny=70, nt=240, mode=1, seismo=1, dtinv=2)
    a = _gradient_run(case, tile, 0)
    b = _gradient_run(case, tile, 1)
    assert a["l2"] == b["l2"]
    for k in range(6):
        assert np.array_equal(a["grad"][k], b["grad"][k]), k
        assert np.abs(a["grad"][k]).max() > 0, k
    for key in a:
        if key.startswith("f"):
            assert np.array_equal(a[key], b[key]), key


@pytest.mark.parametrize("dtinv,grad_form,quelltypb,eprecond",
                         [(1, 1, 1, 0), (3, 2, 1, 3), (2, 1, 4, 1), (2, 2, 7, 3), (1, 1, 2, 1)])
def test_gradient_variants(dtinv, grad_form, quelltypb, eprecond):
    print(f"[flow] gradient DTINV {dtinv} GRAD_FORM {grad_form} QUELLTYPB {quelltypb} EPRECOND {eprecond}",
          flush=True)
    case = make_case(nx=110, ny=60, nt=200, mode=1, seismo=4 if quelltypb >= 4 else 1, dtinv=dtinv,
                     grad_form=grad_form, quelltypb=quelltypb, fdorder=6)
    a = _gradient_run(case, (64, 8, 2, 1), 0, eprecond)
    b = _gradient_run(case, (64, 8, 2, 1), 1, eprecond)
    assert a["l2"] == b["l2"]
    for k in range(6):
        assert np.array_equal(a["grad"][k], b["grad"][k]), k
    if eprecond:
        for w in range(2):
            assert np.array_equal(a["ew"][w], b["ew"][w]), w
        assert np.abs(b["ew"][0]).max() > 0


def test_full_size_gradient_slice():
    """The gradient job's 2000 x 600 grid, 120 steps: equal to the per-step
    launches."""
    case = make_case(nx=2000, ny=600, nt=120, mode=1, seismo=1, dtinv=2, nrec=400)
    a = _gradient_run(case, (64, 4, 2, 1), 0)
    b = _gradient_run(case, (64, 4, 2, 1), 1)
    for k in range(6):
        assert np.array_equal(a["grad"][k], b["grad"][k]), k


def test_runs_it_does_not_apply_to_keep_per_step_launches():
    p, m, src, sig, rec = make_case(nx=100, ny=60, nt=120)
    g = _ctx(p, m, src, sig, rec, (64, 4, 1, 4), 1)
    assert g.dataflow() == (1, 0)  # vector shape
    g.set_tile(64, 4, 2, 1, sweep=-1)
    assert g.dataflow() == (1, 1)
    g.set_tile(128, 2, 1, 1, sweep=0)
    assert g.dataflow() == (1, 0)  # the sweeps' shapes differ
    g.set_tile(64, 4, 2, 1, sweep=-1)
    g.set_kernel_sampling(5)
    assert g.dataflow() == (1, 0)  # per-launch timing needs the per-step launches
    g.set_kernel_sampling(0)
    g.close()
    pv, mv, srcv, sigv, recv = make_case(nx=100, ny=60, nt=120, L=2)
    v = _ctx(pv, mv, srcv, sigv, recv, (64, 4, 2, 1), 1)
    assert v.dataflow() == (1, 0)  # viscoelastic
    v.close()
    p5, m5, src5, sig5, rec5 = make_case(nx=100, ny=60, nt=120, stype=5)
    g5 = _ctx(p5, m5, src5, sig5, rec5, (64, 4, 2, 1), 1)
    assert g5.dataflow() == (1, 0)  # moment-tensor source
    g5.forward()
    g0 = _ctx(p5, m5, src5, sig5, rec5, (64, 4, 2, 1), 0)
    g0.forward()
    _same_forward(g0, g5)
    g0.close()
    g5.close()
